"""(T) oracle (oracle/ccsd_t.py: ccsd_t.h:1127-1167, 2274-2335) pinned by the
independent spin-orbital (T) of the textbook equations and by the reference's
golden (T) energy of h2o-ccsd_t-631g-pvdz.out."""
import numpy as np
import pytest

from oracle import ccsd_t as T
from spinorb_ccsd import mpqc_blocks, random_hamiltonian, spin_to_spatial, spinorb_ccsd, spinorb_triples
from test_golden import FIX

GOLDEN_T = FIX["triples_cases"][0]["e_triples"]


@pytest.mark.parametrize("no,nv,seed,offdiag", [(3, 5, 3, 0.0), (2, 6, 7, 0.02), (4, 4, 9, 0.02)])
def test_closed_shell_triples_equal_spin_orbital_triples(no, nv, seed, offdiag):
    f, g = random_hamiltonian(no, nv, seed=seed, eri_scale=0.05, fock_offdiag=offdiag)
    _, t1s, t2s, _ = spinorb_ccsd(f, g, no)
    t1, t2 = spin_to_spatial(t1s, t2s, no, nv)
    b = mpqc_blocks(f, g, no)
    et = T.ccsd_t_straight(t1, t2, b["Gijab"], b["Giabc"], b["Gijka"], np.diag(f)[:no], np.diag(f)[no:])
    ref = spinorb_triples(f, g, no, t1s, t2s)
    assert abs(et - ref) < 1e-14 * max(1.0, abs(ref)) and abs(ref) > 1e-6
    # planted mistake: the hole term with the wrong sign is caught
    wrong = T.ccsd_t_straight(t1, t2, b["Gijab"], b["Giabc"], -b["Gijka"], np.diag(f)[:no], np.diag(f)[no:])
    assert abs(wrong - ref) > 1e-3 * abs(ref)


def test_golden_triples_of_the_df_water():
    from oracle import ccsd_df as D
    from test_oracle_df import _df_inputs
    case = FIX["df_cases"][0]
    d, inp = _df_inputs(case)
    ints = D.DFIntegrals(**inp.df_blocks())
    E, t1, t2, _ = D.compute_ccsd_df(ints, inp.base.eps_occ, inp.base.eps_vir,
                                     target_precision=case["precision"])
    b = inp.base.blocks  # the [df] four-index blocks incl. <ia|bc>
    et = T.ccsd_t_straight(t1, t2, b["Gijab"], b["Giabc"], b["Gijka"], inp.base.eps_occ, inp.base.eps_vir)
    print(f"(T) {et:.15e} reference {GOLDEN_T:.15e}")
    assert abs(et - GOLDEN_T) < 2e-12
