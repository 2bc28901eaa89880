"""AO integral-direct CCSD on the GPU (method "direct": mpqc_ccsd_set_direct_ao,
U = (pq|rs) [Ca tau Ca^T] each iteration, ccsd.h:573-603) against the direct oracle
(oracle/ccsd_direct.py) and the reference's own direct run (h2o-eom-ccsd-direct-631g.out:
its iteration table and its EOM singlets).

Bar: residuals 1e-13 relative to their largest element in FP64 (2e-6 in FP32);
energies on the golden table as tests/test_golden.py; singlets 2e-9 Eh.
"""
import numpy as np
import pytest

from oracle import ccsd as O
from oracle import ccsd_direct as DR
from test_golden import FIX, _inputs, check_trace

pytestmark = pytest.mark.gpu


def _direct_blocks(inp):
    return {k: a for k, a in inp.blocks.items() if k != "Gabcd"}


def _coefs(scf, inp):
    return scf.C[:, inp.n_frozen:scf.ndocc], scf.C[:, scf.ndocc:]


@pytest.mark.parametrize("prec,layout,tol", [("f64", "dense", 1e-13), ("f64", "ta", 1e-13), ("f32", "dense", 2e-6)])
def test_direct_residual_matches_oracle(prec, layout, tol):
    from mpqc_amd.ccsd import CCSD
    scf, inp = _inputs(FIX["cases"][1])          # H2O cc-pVDZ, all electrons: nbf 24, o 5, v 19
    Ci, Ca = _coefs(scf, inp)
    rng = np.random.default_rng(3)
    t1 = rng.normal(0, 0.02, (inp.v, inp.o))
    t2 = rng.normal(0, 0.02, (inp.v, inp.v, inp.o, inp.o))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    ints = O.Integrals(**{**_direct_blocks(inp), "Gabcd": None}, Ci=Ci, Ca=Ca)
    tau = O.make_tau(t1, t2)
    r1r, r2r = DR.direct_residual(scf.eri)(t1, t2, tau, ints)
    print(f"{prec} {layout} nbf={scf.eri.shape[0]} o={inp.o} v={inp.v}", flush=True)
    with CCSD(inp.o, inp.v, precision=prec, occ_block_size=2, unocc_block_size=5) as c:
        c.set_integrals(_direct_blocks(inp), layout=layout)
        c.set_direct_ao(scf.eri, Ci, Ca)
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        r1, r2 = c.compute_residuals(t1, t2)
    e1 = np.abs(r1 - r1r).max() / np.abs(r1r).max()
    e2 = np.abs(r2 - r2r).max() / np.abs(r2r).max()
    print(f"  r1 {e1:.2e}  r2 {e2:.2e}")
    assert e1 < tol and e2 < tol


def test_direct_trace_and_eom_match_reference_direct_run():
    from mpqc_amd.ccsd import CCSD
    case = FIX["cases"][0]
    eom = FIX["eom_cases"][0]
    assert "direct" in case["source"] and eom["source"] == case["source"]
    scf, inp = _inputs(case)
    Ci, Ca = _coefs(scf, inp)
    tp = case["precision"] * case["precision_scale_for_current_rule"]
    with CCSD(inp.o, inp.v, precision="f64", occ_block_size=4, unocc_block_size=4, max_iter=100) as c:
        c.set_integrals(_direct_blocks(inp), layout="ta")
        c.set_direct_ao(scf.eri, Ci, Ca)
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        E = c.compute_ccsd(target_precision=tp)
        rows = c.trace()
        check_trace(rows, case, 2e-12)
        assert len(rows) == len(case["trace"])
        assert abs(E - case["e_ccsd"]) < 2e-12
        # EOM on the direct residual's Jacobian (the reference's EOM of the same run)
        c.compute_ccsd(target_precision=1e-12)
        w, it = c.eom_ee(len(eom["singlets"]), target_precision=1e-10, max_iter=200)
    print("GPU direct", E, "singlets", w, "iterations", it)
    assert np.abs(w - np.array(eom["singlets"])).max() < 2e-9


def test_direct_rejects_misuse():
    from mpqc_amd import _lib as L
    from mpqc_amd.ccsd import CCSD
    scf, inp = _inputs(FIX["cases"][0])
    Ci, Ca = _coefs(scf, inp)
    with CCSD(inp.o, inp.v) as c:
        with pytest.raises(L.CCSDError):
            c.set_option("method", 2)                      # no AO integrals yet
        with pytest.raises(ValueError):
            c.set_direct_ao(scf.eri, Ca, Ci)               # coefficient shapes swapped
    with CCSD(inp.o, inp.v, precision="f16x3") as c:
        c.set_integrals(_direct_blocks(inp))
        c.set_direct_ao(scf.eri, Ci, Ca)
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        with pytest.raises(L.CCSDError):
            c.compute_ccsd()                               # split forms: standard and df only
    with CCSD(inp.o, inp.v) as c:
        c.set_integrals(_direct_blocks(inp))
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        with pytest.raises(L.CCSDError):
            c.compute_ccsd()                               # standard method without Gabcd


def _direct_df_case():
    from mpqc_amd.chem.basis import BasisSet
    from mpqc_amd.chem.molecule import Molecule
    from test_oracle_df import _df_inputs
    dcase = FIX["df_cases"][0]
    d, inp = _df_inputs(dcase)
    eri = BasisSet(dcase["basis"], Molecule.from_xyz_text(FIX["geometry_angstrom"])).eri()  # exact (mu nu|la si)
    s = d.scf
    nfz = inp.base.n_frozen
    return inp, eri, s.C[:, nfz:s.ndocc].copy(), s.C[:, s.ndocc:].copy()


def _dfints(inp, Ci, Ca):
    from oracle import ccsd_df as D
    return D.DFIntegrals(**inp.df_blocks(), Ci=Ci, Ca=Ca)


@pytest.mark.parametrize("layout", ["dense", "ta"])
def test_direct_df_residual_matches_oracle(layout):
    """method direct_df (ccsd.h:73; the df residual with U, :436/:457) against the oracle."""
    from mpqc_amd.ccsd import CCSD
    inp, eri, Ci, Ca = _direct_df_case()
    rng = np.random.default_rng(9)
    t1 = rng.normal(0, 0.02, (inp.v, inp.o))
    t2 = rng.normal(0, 0.02, (inp.v, inp.v, inp.o, inp.o))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    tau = O.make_tau(t1, t2)
    r1r, r2r = DR.direct_df_residual(eri)(t1, t2, tau, _dfints(inp, Ci, Ca))
    print(f"direct_df {layout} o={inp.o} v={inp.v} naux={inp.naux}", flush=True)
    with CCSD(inp.o, inp.v, precision="f64", occ_block_size=2, unocc_block_size=3) as c:
        c.set_df_basis(inp.naux, aux_block_size=7 if layout == "ta" else 0)
        c.set_integrals(inp.df_blocks(), layout=layout)
        c.set_direct_ao(eri, Ci, Ca)
        c.set_option("method", 3)
        c.set_orbital_energies(inp.base.eps_occ, inp.base.eps_vir)
        r1, r2 = c.compute_residuals(t1, t2)
    e1 = np.abs(r1 - r1r).max() / np.abs(r1r).max()
    e2 = np.abs(r2 - r2r).max() / np.abs(r2r).max()
    print(f"  r1 {e1:.2e}  r2 {e2:.2e}")
    assert e1 < 1e-13 and e2 < 1e-13


def test_direct_df_solve_and_eom_match_oracle():
    """Converged direct_df energy against the oracle's solve, and its EOM singlets
    (the Jacobian of the direct_df residual) against the oracle's central differences."""
    from oracle import eom as EO
    from mpqc_amd.ccsd import CCSD
    inp, eri, Ci, Ca = _direct_df_case()
    res = DR.direct_df_residual(eri)
    E_ref, t1r, t2r, _ = O.compute_ccsd(_dfints(inp, Ci, Ca), inp.base.eps_occ, inp.base.eps_vir,
                                        target_precision=1e-12, max_iter=100, residual_fn=res)
    with CCSD(inp.o, inp.v, precision="f64", max_iter=100) as c:
        c.set_df_basis(inp.naux)
        c.set_integrals(inp.df_blocks())
        c.set_direct_ao(eri, Ci, Ca)
        c.set_option("method", 3)
        c.set_orbital_energies(inp.base.eps_occ, inp.base.eps_vir)
        E = c.compute_ccsd(target_precision=1e-12)
        t1, t2 = c.get_amplitudes()
        w, it = c.eom_ee(3, target_precision=1e-10, max_iter=200)
    ints = _dfints(inp, Ci, Ca)
    w_ref = EO.eom_ee_singlets(lambda a, b: res(a, b, O.make_tau(a, b), ints), t1, t2, 3)
    print(f"direct_df E {E:.14f} oracle {E_ref:.14f}; singlets {w} oracle {w_ref} ({it} iterations)")
    assert abs(E - E_ref) < 1e-11
    assert np.abs(w - w_ref).max() < 1e-8
