"""EOM-EE-CCSD singlets on the GPU (mpqc_ccsd_eom_ee, EOM_CCSD eom_ccsd_impl.h:236-478)
against the reference's converged excitation energies (h2o-eom-ccsd-direct-631g.out,
h2o-eom-ccsd-631g-pvdz.out) and against the oracle (oracle/eom.py: eigenvalues of the
central-difference Jacobian on the symmetric subspace) for a random Hamiltonian.

Bar: 2e-9 Eh against the reference's 12 printed digits (its CCSD stopped at its own
convergence); 1e-8 against the oracle's central differences (truncation ~h^2 = 1e-10
times the residual's third derivative).
"""
import numpy as np
import pytest

from test_golden import FIX, _inputs

pytestmark = pytest.mark.gpu

TOL_REF = 2e-9


def _converged(c, blocks, eo, ev, prec=1e-12):
    from mpqc_amd.ccsd import Integrals
    c.set_integrals(Integrals(**blocks))
    c.set_orbital_energies(eo, ev)
    return c.compute_ccsd(target_precision=prec)


@pytest.mark.parametrize("max_vector,cis_guess", [(8, 0), (3, 0), (8, 1)])
def test_eom_singlets_standard_water(max_vector, cis_guess):
    from mpqc_amd.ccsd import CCSD
    case = FIX["eom_cases"][0]
    gcase = FIX["cases"][0]
    assert case["basis"] == gcase["basis"] and case["frozen_core"] == gcase["frozen_core"]
    _, inp = _inputs(gcase)
    n = len(case["singlets"])
    print(f"water {case['basis']} o={inp.o} v={inp.v} roots={n} max_vector={max_vector} cis_guess={cis_guess}",
          flush=True)
    with CCSD(inp.o, inp.v, precision="f64", occ_block_size=2, unocc_block_size=3, max_iter=100) as c:
        _converged(c, inp.blocks, inp.eps_occ, inp.eps_vir)
        c.set_option("eom_max_vector", max_vector)
        c.set_option("eom_cis_guess", cis_guess)
        w, it = c.eom_ee(n, target_precision=1e-10, max_iter=200)
        nprod, _ = c.eom_stats()
    print("GPU", w, "iterations", it, "products", nprod)
    print("reference", case["singlets"])
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - np.array(case["singlets"])).max() < TOL_REF


def test_eom_singlets_df_water_and_context_restored():
    from mpqc_amd.ccsd import CCSD
    from test_oracle_df import _df_inputs
    case = FIX["eom_cases"][1]
    dcase = FIX["df_cases"][0]
    assert case["basis"] == dcase["basis"] and case["df_basis"] == dcase["df_basis"]
    _, inp = _df_inputs(dcase)
    n = len(case["singlets"])
    print(f"DF water o={inp.o} v={inp.v} naux={inp.naux} roots={n}", flush=True)
    with CCSD(inp.o, inp.v, precision="f64", occ_block_size=2, unocc_block_size=2, max_iter=100) as c:
        c.set_df_basis(inp.naux, aux_block_size=4)
        c.set_integrals(inp.df_blocks(), layout="ta")     # no Giabc / Gabcd: built from the factors
        c.set_orbital_energies(inp.base.eps_occ, inp.base.eps_vir)
        E = c.compute_ccsd(target_precision=1e-12)
        t1, t2 = c.get_amplitudes()
        w, it = c.eom_ee(n, target_precision=1e-10, max_iter=200)
        # the context is back on the df method: the residual at the converged amplitudes
        r1, r2 = c.compute_residuals(t1, t2)
    print("GPU", w, "iterations", it)
    print("reference", case["singlets"])
    assert np.abs(w - np.array(case["singlets"])).max() < TOL_REF
    assert abs(E - dcase["e_ccsd"]) < 2e-12
    assert np.abs(r1).max() < 1e-9 and np.abs(r2).max() < 1e-9


@pytest.mark.parametrize("cis_guess", [0, 1])
def test_eom_matches_oracle_random_hamiltonian(cis_guess):
    from oracle import ccsd as O
    from oracle import eom as EO
    from mpqc_amd.ccsd import CCSD
    from spinorb_ccsd import mpqc_blocks, random_hamiltonian
    no, nv = 3, 6
    f, g = random_hamiltonian(no, nv, seed=11, eri_scale=0.03)
    b = mpqc_blocks(f, g, no)
    eo, ev = np.diag(f)[:no].copy(), np.diag(f)[no:].copy()
    with CCSD(no, nv, precision="f64", max_iter=200) as c:
        _converged(c, b, eo, ev, prec=1e-14)
        t1, t2 = c.get_amplitudes()
        c.set_option("eom_cis_guess", cis_guess)
        w, it = c.eom_ee(3, target_precision=1e-10, max_iter=200)
    ref = EO.eom_ee_singlets(EO.standard_residual(O.Integrals(**b)), t1, t2, 3)
    print("GPU", w, "oracle", ref, "iterations", it)
    assert np.abs(w - ref).max() < 1e-8


@pytest.mark.parametrize("prec", ["f16x3", "bf16x3"])
def test_eom_singlets_split_form_context(prec):
    """A split-form context converges its amplitudes on the 16-bit plane
    contractions; the Jacobian products of EOM are reverse sweeps of the FP64
    tape at those amplitudes.  The singlets agree with the reference's within
    the amplitudes' own difference from FP64 (target precision 1e-12)."""
    from mpqc_amd.ccsd import CCSD
    case = FIX["eom_cases"][0]
    _, inp = _inputs(FIX["cases"][0])
    n = len(case["singlets"])
    with CCSD(inp.o, inp.v, precision=prec, max_iter=100) as c:
        _converged(c, inp.blocks, inp.eps_occ, inp.eps_vir)
        w, it = c.eom_ee(n, target_precision=1e-10, max_iter=200)
    d = np.abs(w - np.array(case["singlets"])).max()
    print(f"{prec} context: singlets {w}, max |dw| vs reference {d:.2e}")
    assert d < 1e-7


def test_eom_rejects_misuse():
    from mpqc_amd import _lib as L
    from mpqc_amd.ccsd import CCSD
    _, inp = _inputs(FIX["cases"][0])
    with CCSD(inp.o, inp.v, precision="f64") as c:
        _converged(c, inp.blocks, inp.eps_occ, inp.eps_vir, prec=1e-8)
        with pytest.raises(L.CCSDError):
            c.eom_ee(inp.o * inp.v + 1)                     # more roots than singles
        with pytest.raises(L.CCSDError):
            c.eom_ee(2, max_iter=1)                         # not converged in one iteration
        with pytest.raises(L.CCSDError):
            c.set_option("eom_max_vector", 1)
