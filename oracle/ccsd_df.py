"""CPU oracle for MPQC's density-fitted closed-shell CCSD residual (numpy).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline.  The product path (``mpqc_amd``) never
imports it.

Restates, term by term and in the reference's evaluation order
(paths relative to the MPQC source tree):

* ``DFIntegrals``          src/mpqc/chemistry/qc/lcao/cc/ccsd_r1_r2.h:24-56 (DF members
                           Xai, Xij, Xab; filled at ccsd.h:281-305, X = (K|G|p q)[inv_sqr],
                           ccsd.h:481-493)
* ``cs_ccsd_r1_df``        ccsd_r1_r2.h:136-184
* ``cs_ccsd_r2_df``        ccsd_r1_r2.h:309-520: with ``u`` unset the Gabcd branch
                           (:418-430) when Gabcd is given, otherwise the lazy
                           density-fitted ladder of reduced_abcd_memory (:497-508);
                           with ``u`` (method direct_df) the U terms (:509-514)
* ``df_abcd``              direct_task_integrals.h:255-300 with integral_builder.h:228-325
                           (Physical notation: g(a,b,c,d) = sum_K bra(K,a,c) ket(K,b,d))

Pinning (tests/test_oracle_df.py, tests/test_golden.py): the DF residuals
equal the standard residuals (oracle/ccsd.py, itself pinned by the golden
traces and spin-orbital CCSD) evaluated with the four-index integrals the same
three-index factors define, on random symmetric factors; the DF solve
reproduces the reference's h2o-ccsd-631g-pvdz trace and energy.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .ccsd import F_OO, W_oooo, compute_ccsd

__all__ = ["DFIntegrals", "cs_ccsd_r1_df", "cs_ccsd_r2_df", "df_abcd", "compute_ccsd_df",
           "four_index_from_factors"]


def _e(subscripts: str, *operands):
    return np.einsum(subscripts, *operands, optimize=True)


@dataclasses.dataclass
class DFIntegrals:
    """``cc::Integrals`` as the DF method fills it (ccsd.h:281-305): with
    reduced_abcd_memory (the default) Giabc and Gabcd stay empty."""

    Fia: np.ndarray
    Fij: np.ndarray
    Fab: np.ndarray
    Gijab: np.ndarray
    Gijkl: np.ndarray
    Giajb: np.ndarray
    Gijka: np.ndarray
    Xai: np.ndarray  # (K, a, i)
    Xij: np.ndarray  # (K, i, j)
    Xab: np.ndarray  # (K, a, b)
    Giabc: Optional[np.ndarray] = None
    Gabcd: Optional[np.ndarray] = None
    FIJ: Optional[np.ndarray] = None
    FAB: Optional[np.ndarray] = None
    # active occupied / virtual MO coefficients of method direct_df (ccsd.h:313-322)
    Ci: Optional[np.ndarray] = None
    Ca: Optional[np.ndarray] = None

    @property
    def o(self) -> int:
        return self.Fij.shape[0]

    @property
    def v(self) -> int:
        return self.Fab.shape[0]


def df_abcd(bra, ket):
    """Physical-notation DF four-index tensor of two three-index factors:
    g(a,b,c,d) = sum_K bra(K,a,c) ket(K,b,d)."""
    return _e("Kac,Kbd->abcd", bra, ket)


def cs_ccsd_r1_df(t1, t2, tau, ints: DFIntegrals):
    """ccsd_r1_r2.h:136-184.  Returns r1(a,i); writes ints.FAB and ints.FIJ."""
    Xai, Xij, Xab = ints.Xai, ints.Xij, ints.Xab
    # :147-148
    X_ai_tau = 2.0 * _e("Xbj,abij->Xai", Xai, tau) - _e("Xbj,abji->Xai", Xai, tau)
    # :152
    r1 = ints.Fia.T - 2.0 * _e("kc,ci,ak->ai", ints.Fia, t1, t1)
    # :155-157
    h_ac = ints.Fab - _e("Kal,Kcl->ac", X_ai_tau, Xai)
    ints.FAB = h_ac
    r1 = r1 + h_ac @ t1
    # :161-163
    h_ki = ints.Fij + _e("Kci,Kck->ki", X_ai_tau, Xai)
    ints.FIJ = h_ki
    r1 = r1 - t1 @ h_ki
    # :167-171
    h_kc = ints.Fia + _e("klcd,dl->kc", 2.0 * ints.Gijab - ints.Gijab.transpose(0, 1, 3, 2), t1)
    r1 = r1 + _e("kc,caki->ai", h_kc, 2.0 * t2 - t2.transpose(0, 1, 3, 2)) \
        + _e("kc,ci,ak->ai", h_kc, t1, t1)
    # :174-175
    r1 = r1 + _e("kica,ck->ai", 2.0 * ints.Gijab, t1) - _e("kaic,ck->ai", ints.Giajb, t1)
    # :177
    r1 = r1 + _e("Kci,Kac->ai", X_ai_tau, Xab)
    # :179
    r1 = r1 - _e("Kal,Kli->ai", X_ai_tau, Xij)
    return r1


def cs_ccsd_r2_df(t1, t2, tau, ints: DFIntegrals, u=None):
    """ccsd_r1_r2.h:309-520.  Requires ints.FIJ / ints.FAB from the r1 of the
    method.  ``u`` (U(p,r,i,j), AO indices p, r) replaces the ladder block."""
    G = ints
    Xai, Xij, Xab = G.Xai, G.Xij, G.Xab
    # :317-318
    X_ab_t1 = _e("Kab,bi->Kai", Xab, t1)
    # :321-323
    r2 = _e("Kbj,Kai->abij", X_ab_t1, Xai - _e("Kki,ak->Kai", Xij, t1)) \
        - _e("jika,bk->abij", G.Gijka, t1) - _e("ikac,cj,bk->abij", G.Gijab, t1, t1)
    # :329-338
    g_ijka_bar = 2.0 * G.Gijka - G.Gijka.transpose(1, 0, 2, 3)
    g_ki = F_OO(G.FIJ, G.Fia, g_ijka_bar, t1)
    g_ac = (G.FAB - _e("kc,ak->ac", G.Fia, t1)
            + 2.0 * _e("K,Kac->ac", _e("Kdk,dk->K", Xai, t1), Xab)
            - _e("Kak,Kck->ac", X_ab_t1, Xai))
    # :340
    r2 = r2 + _e("ac,cbij->abij", g_ac, t2) - _e("ki,abkj->abij", g_ki, t2)
    # :349
    T = 0.5 * t2 + _e("di,bl->dbil", t1, t1)
    # :351-356
    j_akic = (G.Gijab.transpose(2, 1, 0, 3)
              - _e("lkic,al->akic", G.Gijka, t1)
              + _e("Kai,Kck->akic", X_ab_t1, Xai)
              - _e("xdl,dail,xck->akic", Xai, T, Xai)
              + 0.5 * _e("klcd,adil->akic", 2.0 * G.Gijab - G.Gijab.transpose(0, 1, 3, 2), t2))
    # :358-361
    k_kaic = (G.Giajb
              - _e("klic,al->kaic", G.Gijka, t1)
              + _e("Kdk,di,Kac->kaic", Xai, t1, Xab)
              - _e("kldc,dail->kaic", G.Gijab, T))
    # :364-365
    r2 = r2 + 0.5 * _e("akic,cbkj->abij", 2.0 * j_akic - k_kaic.transpose(1, 0, 2, 3),
                       2.0 * t2 - t2.transpose(1, 0, 2, 3))
    # :367-369
    tmp = _e("kaic,cbjk->abij", k_kaic, t2)
    r2 = r2 - (0.5 * tmp + tmp.transpose(1, 0, 2, 3))
    # :373-374
    r2 = r2 + r2.transpose(1, 0, 3, 2)
    r2 = r2 + G.Gijab.transpose(2, 3, 0, 1)
    # :378-381
    a_klij = W_oooo(G.Gijkl, G.Gijka, G.Gijab, tau, t1)
    r2 = r2 + _e("klij,abkl->abij", a_klij, tau)
    # particle-particle ladder
    if u is not None:
        # :509-514 (as ccsd_r1_r2.h:282-286)
        b_abij = (_e("prij,rb,pa->abij", u, G.Ca, G.Ca)
                  - _e("rk,bk,prij,pa->abij", G.Ci, t1, u, G.Ca)
                  - _e("prij,pk,rb,ak->abij", u, G.Ci, G.Ca, t1))
    elif G.Gabcd is not None:
        # :418-430
        b_abij = _e("cdij,abcd->abij", tau, G.Gabcd)
        tmp = _e("kacd,cdij->kaij", G.Giabc, tau)
        b_abij = b_abij - _e("kaji,bk->abij", tmp, t1) - _e("kbij,ak->abij", tmp, t1)
    else:
        # :497-508
        Y = 0.5 * Xab - _e("Kbi,ai->Kab", Xai, t1)
        b_abij = _e("cdij,abcd->abij", tau, df_abcd(Xab, Y))
        b_abij = b_abij + b_abij.transpose(1, 0, 3, 2)
    return r2 + b_abij


def compute_ccsd_df(ints: DFIntegrals, eps_occ, eps_vir, **kw):
    """ccsd.h:268-477 with method = "df" (residuals at :435 and :454)."""
    def residual_fn(t1, t2, tau, ints_):
        r1 = cs_ccsd_r1_df(t1, t2, tau, ints_)
        return r1, cs_ccsd_r2_df(t1, t2, tau, ints_)
    return compute_ccsd(ints, eps_occ, eps_vir, residual_fn=residual_fn, **kw)


def four_index_from_factors(Xai, Xij, Xab):
    """The four-index blocks (physicist notation) that symmetric three-index
    factors define through <pq|rs> = (pr|qs) = sum_K X(K,p,r) X(K,q,s).
    Xia(K,i,a) = Xai(K,a,i)."""
    Xia = Xai.transpose(0, 2, 1)
    return dict(
        Gijab=_e("Kia,Kjb->ijab", Xia, Xia),
        Gijkl=_e("Kik,Kjl->ijkl", Xij, Xij),
        Giajb=_e("Kij,Kab->iajb", Xij, Xab),
        Giabc=_e("Kib,Kac->iabc", Xia, Xab),
        Gijka=_e("Kik,Kja->ijka", Xij, Xia),
        Gabcd=_e("Kac,Kbd->abcd", Xab, Xab))
