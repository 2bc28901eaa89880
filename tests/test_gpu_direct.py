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
