"""The pair-symmetric particle-particle ladder and hole-hole G tau term on the GPU
(option pp_symmetric, the default) against the full contractions (pp_symmetric 0) and
the oracle, in the FP64 and split forms, at amplitudes without the (ab)(ij) symmetry
(ccsd_r1_r2.h:275, ccsd_intermediates.h:182-193).

Bars: FP64 1e-12 relative between the two forms and against the oracle (round-off of
sums of different order); split forms 1e-6 (f16x3, measured 1.1e-7) and 5e-5 (bf16x3,
measured 4.9e-6) relative against the oracle, about ten times the errors the split
residuals show at this size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case():
    from oracle import ccsd as O
    from spinorb_ccsd import mpqc_blocks, random_hamiltonian
    no, nv = 5, 29
    f, g = random_hamiltonian(no, nv, seed=31, eri_scale=0.03)
    b = mpqc_blocks(f, g, no)
    rng = np.random.default_rng(8)
    t1 = rng.normal(0, 0.01, (nv, no))
    t2 = rng.normal(0, 0.01, (nv, nv, no, no))
    ints = O.Integrals(**b)
    tau = O.make_tau(t1, t2)
    return no, nv, f, b, t1, t2, O.cs_ccsd_r1(t1, t2, tau, ints), O.cs_ccsd_r2(t1, t2, tau, ints)


def _residuals(prec, sym, no, nv, f, b, t1, t2):
    from mpqc_amd.ccsd import CCSD, Integrals
    with CCSD(no, nv, precision=prec) as c:
        c.set_option("pp_symmetric", sym)
        c.set_integrals(Integrals(**b))
        c.set_orbital_energies(np.diag(f)[:no], np.diag(f)[no:])
        return c.compute_residuals(t1, t2)


@pytest.mark.parametrize("prec,bar", [("f64", 1e-12), ("f16x3", 1e-6), ("bf16x3", 5e-5)])
def test_pair_symmetric_ladder_residual(prec, bar):
    no, nv, f, b, t1, t2, r1r, r2r = _case()
    print(f"pp_symmetric residual {prec} o={no} v={nv}", flush=True)
    res = {s: _residuals(prec, s, no, nv, f, b, t1, t2) for s in (0, 1)}
    for s, (r1, r2) in res.items():
        e1 = np.abs(r1 - r1r).max() / np.abs(r1r).max()
        e2 = np.abs(r2 - r2r).max() / np.abs(r2r).max()
        print(f"  pp_symmetric={s}: r1 {e1:.2e} r2 {e2:.2e}")
        assert e1 < bar and e2 < bar
    d2 = np.abs(res[1][1] - res[0][1]).max() / np.abs(r2r).max()
    print(f"  symmetric vs full r2 {d2:.2e}")
    assert d2 < bar
    if prec == "f64":
        assert d2 < 1e-12


def test_pair_symmetric_ladder_solve_energy():
    from mpqc_amd.ccsd import CCSD, Integrals
    from spinorb_ccsd import mpqc_blocks, random_hamiltonian
    no, nv = 4, 20
    f, g = random_hamiltonian(no, nv, seed=5, eri_scale=0.03)
    b = mpqc_blocks(f, g, no)
    E = {}
    for s in (0, 1):
        with CCSD(no, nv, precision="f64", max_iter=100) as c:
            c.set_option("pp_symmetric", s)
            c.set_integrals(Integrals(**b))
            c.set_orbital_energies(np.diag(f)[:no], np.diag(f)[no:])
            E[s] = c.compute_ccsd(target_precision=1e-11)
    print("E", E)
    assert abs(E[1] - E[0]) < 1e-11


@pytest.mark.parametrize("block", ["Gabcd", "Gijab", "Giabc"])
@pytest.mark.parametrize("prec,bar", [("f64", 1e-12), ("f16x3", 1e-6)])
def test_asymmetric_blocks_take_the_full_contraction(prec, bar, block):
    # Gabcd (pp ladder) or Gijab (hole-hole intermediate) without the pair symmetry, or
    # Giabc without Giabc(i,c,a,b) = Giabc(i,b,a,c) (Rd's T1 dressing term taken from the j
    # term's Giabc t1), after a symmetry-breaking perturbation: the context detects it and
    # runs the reference's contraction, so the residual still matches the oracle
    from oracle import ccsd as O
    no, nv, f, b, t1, t2, _, _ = _case()
    rng = np.random.default_rng(9)
    b = dict(b)
    b[block] = b[block] + 1e-3 * rng.normal(size=b[block].shape)
    ints = O.Integrals(**b)
    tau = O.make_tau(t1, t2)
    r1r, r2r = O.cs_ccsd_r1(t1, t2, tau, ints), O.cs_ccsd_r2(t1, t2, tau, ints)
    r1, r2 = _residuals(prec, 1, no, nv, f, b, t1, t2)
    e1 = np.abs(r1 - r1r).max() / np.abs(r1r).max()
    e2 = np.abs(r2 - r2r).max() / np.abs(r2r).max()
    print(f"asymmetric {block} {prec}: r1 {e1:.2e} r2 {e2:.2e}")
    assert e1 < bar and e2 < bar
