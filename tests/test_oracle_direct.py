"""AO integral-direct CCSD (method "direct", oracle/ccsd_direct.py) against the
reference's own direct run (h2o-eom-ccsd-direct-631g.out iteration table) and
the closed form that links it to the standard residual."""
import numpy as np
import pytest

from oracle import ccsd as O
from oracle import ccsd_direct as DR
from test_golden import FIX, _inputs, check_trace


def _direct_ints(scf, inp):
    b = dict(inp.blocks)
    b["Gabcd"] = None          # the direct method never forms (ab|cd) (ccsd.h:291-299)
    nfz = inp.n_frozen
    return O.Integrals(**b, Ci=scf.C[:, nfz:scf.ndocc].copy(), Ca=scf.C[:, scf.ndocc:].copy())


def test_direct_trace_matches_reference_direct_run():
    case = FIX["cases"][0]
    assert "direct" in case["source"]
    scf, inp = _inputs(case)
    ints = _direct_ints(scf, inp)
    tp = case["precision"] * case["precision_scale_for_current_rule"]
    E, t1, t2, rows = DR.compute_ccsd_direct(ints, scf.eri, inp.eps_occ, inp.eps_vir, target_precision=tp)
    print(f"direct oracle {E:.15f} reference {case['e_ccsd']:.15f}")
    check_trace(rows, case, 2e-12)
    assert len(rows) == len(case["trace"])
    assert abs(E - case["e_ccsd"]) < 2e-12


def _random_amplitudes(o, v, seed):
    rng = np.random.default_rng(seed)
    t1 = rng.normal(0, 0.02, (v, o))
    t2 = rng.normal(0, 0.02, (v, v, o, o))
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


@pytest.mark.parametrize("idx", [0, 1])
def test_direct_residual_equals_standard_at_consistent_integrals(idx):
    """U transformed by (Ca, Ca) is tau Gabcd and by (Ci, Ca) the Giabc tau
    contraction when every block is a transform of the same AO integrals, so the
    two residuals agree at any amplitudes; a transposed AO index is caught."""
    scf, inp = _inputs(FIX["cases"][idx])
    t1, t2 = _random_amplitudes(inp.o, inp.v, seed=idx)
    tau = O.make_tau(t1, t2)
    std = O.Integrals(**inp.blocks)
    r1s = O.cs_ccsd_r1(t1, t2, tau, std)
    r2s = O.cs_ccsd_r2(t1, t2, tau, std)
    dint = _direct_ints(scf, inp)
    r1d, r2d = DR.direct_residual(scf.eri)(t1, t2, tau, dint)
    e1 = np.abs(r1d - r1s).max() / np.abs(r1s).max()
    e2 = np.abs(r2d - r2s).max() / np.abs(r2s).max()
    print(f"direct vs standard: r1 {e1:.2e} r2 {e2:.2e}")
    assert e1 < 1e-12 and e2 < 1e-12
    # planted mistake: (pr|qs) in place of (pq|rs)
    bad = DR.direct_residual(scf.eri.transpose(0, 2, 1, 3))(t1, t2, tau, _direct_ints(scf, inp))
    assert np.abs(bad[1] - r2s).max() / np.abs(r2s).max() > 1e-4


def _exact_factors(scf, inp):
    """Three-index factors X(K,p,q) that reproduce the active-MO integrals exactly:
    the eigen-factorisation of the (pq|rs) supermatrix."""
    from oracle import ccsd_df as D
    nfz = inp.n_frozen
    C = scf.C[:, nfz:]
    mo = np.einsum("mnls,mp,nq,lr,st->pqrt", scf.eri, C, C, C, C, optimize=True)
    n = C.shape[1]
    w, U = np.linalg.eigh(mo.reshape(n * n, n * n))
    keep = w > 1e-12 * w.max()
    X = (U[:, keep] * np.sqrt(w[keep])).T.reshape(-1, n, n)
    o = inp.o
    Xai, Xij, Xab = X[:, o:, :o].copy(), X[:, :o, :o].copy(), X[:, o:, o:].copy()
    g = D.four_index_from_factors(Xai, Xij, Xab)
    return Xai, Xij, Xab, g


def test_direct_df_equals_standard_with_exact_factors():
    """direct_df (the df residual with U) at factors that reproduce the integrals
    exactly is the standard residual; a transposed AO index in U is caught."""
    from oracle import ccsd_df as D
    scf, inp = _inputs(FIX["cases"][0])
    Xai, Xij, Xab, g = _exact_factors(scf, inp)
    for k in ("Gijab", "Gijkl", "Giajb", "Gijka"):
        assert np.abs(g[k] - inp.blocks[k]).max() < 1e-12
    t1, t2 = _random_amplitudes(inp.o, inp.v, seed=5)
    tau = O.make_tau(t1, t2)
    std = O.Integrals(**inp.blocks)
    r1s = O.cs_ccsd_r1(t1, t2, tau, std)
    r2s = O.cs_ccsd_r2(t1, t2, tau, std, fov_sign=-1.0)   # the df residual's g_ac convention
    b = inp.blocks
    nfz = inp.n_frozen

    def dints():
        return D.DFIntegrals(Fia=b["Fia"], Fij=b["Fij"], Fab=b["Fab"], Gijab=g["Gijab"], Gijkl=g["Gijkl"],
                             Giajb=g["Giajb"], Gijka=g["Gijka"], Xai=Xai, Xij=Xij, Xab=Xab,
                             Ci=scf.C[:, nfz:scf.ndocc].copy(), Ca=scf.C[:, scf.ndocc:].copy())
    r1d, r2d = DR.direct_df_residual(scf.eri)(t1, t2, tau, dints())
    e1 = np.abs(r1d - r1s).max() / np.abs(r1s).max()
    e2 = np.abs(r2d - r2s).max() / np.abs(r2s).max()
    print(f"direct_df vs standard: r1 {e1:.2e} r2 {e2:.2e}")
    assert e1 < 1e-11 and e2 < 1e-11
    bad = DR.direct_df_residual(scf.eri.transpose(0, 2, 1, 3))(t1, t2, tau, dints())
    assert np.abs(bad[1] - r2s).max() / np.abs(r2s).max() > 1e-4
