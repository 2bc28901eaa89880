"""CPU oracle for MPQC's AO -> MO integral transform (numpy restatement).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline.  The product path (``mpqc_amd``) never
imports it.

Restates LCAOFactory::compute4 (src/mpqc/chemistry/qc/lcao/factory/
lcao_factory.h:378-497):

* ``ccsd_blocks_ao``  the branch without density fitting (:409-493): the first
  bra index transformed first, "i, rho, mu, nu" = C("sigma, i") AO("sigma, rho,
  mu, nu") (:439-441), then the second bra index, then the ket indices; a
  physical-notation request <pq|rs> reads the chemist-notation (pr|qs)
  (:455-461).  The Fock blocks are <p|F|q> = C_p^T F C_q.
* ``ccsd_blocks_df``  the density-fitted branch (:386-406):
  result("i,k,j,l") = left("q,i,j") * right("q,k,l") for physical notation, with
  left = right = (K|G|p q)[inv_sqr] = sum_L L(L,K) (L|p q) (ccsd.h:481-493).

Pinning (tests/test_oracle_ao2mo.py): the Hartree-Fock energy rebuilt from the
MO blocks equals the SCF energy; the reference's golden CCSD energies
(standard and DF) follow from these blocks.
"""
from __future__ import annotations

import numpy as np

__all__ = ["ccsd_blocks_ao", "ccsd_blocks_df", "hf_energy_from_blocks"]


def _e(subscripts, *ops):
    return np.einsum(subscripts, *ops, optimize=True)


def _fock_blocks(F, Ci, Ca):
    return dict(Fia=Ci.T @ F @ Ca, Fij=Ci.T @ F @ Ci, Fab=Ca.T @ F @ Ca)


def ccsd_blocks_ao(eri, F, Ci, Ca):
    """Nine cc::Integrals blocks (ccsd.h:281-305) from AO (mu nu|la si)."""
    def phys(C1, C2, C3, C4):
        # <pq|rs> = (pr|qs): transform (mu nu|la si) with (C1, C3, C2, C4), index by index
        x = _e("sp,sabc->pabc", C1, eri)      # first bra index  (:439-441)
        x = _e("ar,pabc->prbc", C3, x)        # second index of (pr|qs) (:457)
        x = _e("bq,prbc->prqc", C2, x)        # first ket index  (:467)
        x = _e("cs,prqc->prqs", C4, x)        # second ket index (:472)
        return np.ascontiguousarray(x.transpose(0, 2, 1, 3))
    out = _fock_blocks(F, Ci, Ca)
    out.update(Gijab=phys(Ci, Ci, Ca, Ca), Gijkl=phys(Ci, Ci, Ci, Ci), Giajb=phys(Ci, Ca, Ci, Ca),
               Giabc=phys(Ci, Ca, Ca, Ca), Gijka=phys(Ci, Ci, Ci, Ca), Gabcd=phys(Ca, Ca, Ca, Ca))
    return out


def ccsd_blocks_df(B, L, F, Ci, Ca, reduced_abcd=True):
    """The df method's integral set from three-centre (K|mu nu) and a metric
    factor L (L L^T = V^-1): X = L^T B in the MO blocks, and the four-index
    blocks as products of two factors over K."""
    X = B if L is None else _e("LK,Lmn->Kmn", L, B)

    def three(C1, C2):
        return _e("Kmn,mp,nq->Kpq", X, C1, C2)
    Xai, Xij, Xab, Xia = three(Ca, Ci), three(Ci, Ci), three(Ca, Ca), three(Ci, Ca)

    def phys(left, right):
        # result("i,k,j,l") = left("q,i,j") * right("q,k,l")   (:399)
        return np.ascontiguousarray(_e("qij,qkl->ikjl", left, right))
    out = _fock_blocks(F, Ci, Ca)
    out.update(Xai=Xai, Xij=Xij, Xab=Xab, Gijab=phys(Xia, Xia), Gijkl=phys(Xij, Xij),
               Giajb=phys(Xij, Xab), Gijka=phys(Xij, Xia))
    if not reduced_abcd:
        out.update(Giabc=phys(Xia, Xab), Gabcd=phys(Xab, Xab))
    return out


def hf_energy_from_blocks(b, e_nuc, frozen_core_energy=0.0):
    """Closed-shell HF energy over the orbitals of the Fij/Gijkl blocks:
    E = sum_i 2 F_ii - sum_ij (2 <ij|ij> - <ij|ji>) + E_nuc (all occupied active)."""
    G = b["Gijkl"]
    o = G.shape[0]
    ii = np.arange(o)
    J = G[ii[:, None], ii[None, :], ii[:, None], ii[None, :]]
    K = G[ii[:, None], ii[None, :], ii[None, :], ii[:, None]]
    return 2.0 * float(np.trace(b["Fij"])) - float(np.sum(2.0 * J - K)) + e_nuc + frozen_core_energy
