"""EOM-EE-CCSD singlets (oracle/eom.py): eigenvalues of the CCSD Jacobian on the
(ab)(ij)-symmetric subspace against the reference's converged excitation
energies (h2o-eom-ccsd-direct-631g.out, h2o-eom-ccsd-631g-pvdz.out)."""
import numpy as np
import pytest

from oracle import ccsd as O
from oracle import ccsd_df as D
from oracle import eom as E
from test_golden import FIX, _inputs

TOL = 2e-9  # the reference's CCSD stopped at its convergence precision (1e-9 rule)


def test_symmetric_basis_is_orthonormal_and_complete():
    v, o = 3, 2
    n_sym, embed, project = E.symmetric_basis_apply(v, o)
    B = np.stack([embed(np.eye(n_sym)[k]) for k in range(n_sym)], axis=1)
    assert np.abs(B.T @ B - np.eye(n_sym)).max() < 1e-14
    rng = np.random.default_rng(0)
    t2 = rng.normal(size=(v, v, o, o))
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    x = np.concatenate([rng.normal(size=v * o), t2.ravel()])
    assert np.abs(embed(project(x)) - x).max() < 1e-13


def test_eom_singlets_standard_water():
    case = FIX["eom_cases"][0]
    gcase = FIX["cases"][0]
    _, inp = _inputs(gcase)
    ints = O.Integrals(**inp.blocks)
    _, t1, t2, _ = O.compute_ccsd(ints, inp.eps_occ, inp.eps_vir, target_precision=1e-12, max_iter=100)
    w = E.eom_ee_singlets(E.standard_residual(ints), t1, t2, len(case["singlets"]))
    print("oracle", w, "reference", case["singlets"])
    assert np.abs(w - np.array(case["singlets"])).max() < TOL


def test_eom_singlets_df_water():
    from test_oracle_df import _df_inputs
    case = FIX["eom_cases"][1]
    d, inp = _df_inputs(FIX["df_cases"][0])
    ints = D.DFIntegrals(**inp.df_blocks())
    _, t1, t2, _ = D.compute_ccsd_df(ints, inp.base.eps_occ, inp.base.eps_vir, target_precision=1e-12,
                                     max_iter=100)

    def res(a, b):
        tau = O.make_tau(a, b)
        r1 = D.cs_ccsd_r1_df(a, b, tau, ints)
        return r1, D.cs_ccsd_r2_df(a, b, tau, ints)
    w = E.eom_ee_singlets(res, t1, t2, len(case["singlets"]))
    print("oracle", w, "reference", case["singlets"])
    assert np.abs(w - np.array(case["singlets"])).max() < TOL
