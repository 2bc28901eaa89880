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
