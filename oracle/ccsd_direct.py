"""CPU oracle for MPQC's AO integral-direct CCSD (method "direct") (numpy).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline.  The product path (``mpqc_amd``) never
imports it.

The direct method (src/mpqc/chemistry/qc/lcao/cc/ccsd.h:73, set-up :246-247 and
:313-322) keeps the MO blocks with at most three virtual indices and the AO
integrals (mu nu|ka la); each iteration (:423-426) forms

  U(p,r,i,j) = [(t2(a,b,i,j) Ca(q,a)) Ca(s,b) + tc(i,q) tc(j,s)] (p q|r s),
  tc(i,q) = Ca(q,c) t1(c,i)                                (compute_u2_u11, :573-603)

and hands it to the residual, where it replaces the g_iabc_bar tau term of r1
(ccsd_r1_r2.h:116-117) and the particle-particle ladder with its Giabc terms in
r2 (:282-286); see ``oracle.ccsd.cs_ccsd_r1 / cs_ccsd_r2`` with ``u``.

``direct_df_residual`` is method "direct_df" (ccsd.h:73; :423-459): U as above,
then the standard r1 with U on the density-fitted four-index blocks (:436) and
the df r2 with U (:457).

Pinning (tests/test_oracle_direct.py): the iteration table of
h2o-eom-ccsd-direct-631g.out, which the reference ran with method "direct"; the
closed form that U transformed by (Ca, Ca) is the pp-ladder tau Gabcd and by
(Ci, Ca) the Giabc tau contraction when the MO blocks come from the same AO
integrals (so the direct and standard residuals agree), with a planted mistake;
direct_df equals the standard residual when the three-index factors reproduce
the MO integrals exactly (an eigen-factorisation of the (pq|rs) supermatrix).
"""
from __future__ import annotations

import numpy as np

from . import ccsd as O
from . import ccsd_df as D

__all__ = ["compute_u2_u11", "direct_residual", "compute_ccsd_direct", "direct_df_residual"]


def _e(subscripts, *ops):
    return np.einsum(subscripts, *ops, optimize=True)


def compute_u2_u11(t2, t1, ao_eri, Ca):
    """ccsd.h:581-596.  ao_eri[p,q,r,s] = (pq|rs) (chemist notation)."""
    tc = _e("qc,ci->iq", Ca, t1)
    tau_ao = _e("abij,qa,sb->qsij", t2, Ca, Ca) + _e("iq,js->qsij", tc, tc)
    return _e("qsij,pqrs->prij", tau_ao, ao_eri)


def direct_residual(ao_eri):
    """``residual_fn`` of oracle.ccsd.compute_ccsd for the direct method
    (ccsd.h:423-459): U first, then r1 and r2 with U."""
    def res(t1, t2, tau, ints):
        u = compute_u2_u11(t2, t1, ao_eri, ints.Ca)
        r1 = O.cs_ccsd_r1(t1, t2, tau, ints, u=u)
        return r1, O.cs_ccsd_r2(t1, t2, tau, ints, u=u)
    return res


def compute_ccsd_direct(ints, ao_eri, eps_occ, eps_vir, **kw):
    """ccsd.h:268-477 with method = "direct"."""
    return O.compute_ccsd(ints, eps_occ, eps_vir, residual_fn=direct_residual(ao_eri), **kw)


def direct_df_residual(ao_eri):
    """``residual_fn`` for method "direct_df" (ccsd.h:423-459); ``ints`` is a
    DFIntegrals with Ci and Ca."""
    def res(t1, t2, tau, ints):
        u = compute_u2_u11(t2, t1, ao_eri, ints.Ca)
        std = O.Integrals(Fia=ints.Fia, Fij=ints.Fij, Fab=ints.Fab, Gijab=ints.Gijab, Gijkl=ints.Gijkl,
                          Giajb=ints.Giajb, Giabc=None, Gijka=ints.Gijka, Gabcd=None, Ci=ints.Ci, Ca=ints.Ca)
        r1 = O.cs_ccsd_r1(t1, t2, tau, std, u=u)            # :436
        ints.FIJ, ints.FAB = std.FIJ, std.FAB
        return r1, D.cs_ccsd_r2_df(t1, t2, tau, ints, u=u)  # :457
    return res
