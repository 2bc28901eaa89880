"""CPU oracle for MPQC's perturbative triples correction (T) (numpy restatement).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline.  The product path (``mpqc_amd``) never
imports it.

``ccsd_t_straight`` restates CCSD_T::compute_ccsd_t_straight
(src/mpqc/chemistry/qc/lcao/cc/ccsd_t.h:1127-1167) with its integral layouts
(get_aijk / get_abci / get_abij, :2210-2244) and CCSD_T_Reduce (:2274-2335):

  t3(a,b,c,i,j,k) = g_dabi(d,a,b,i) t2(d,c,j,k) - g_cjkl(c,j,k,l) t2(a,b,i,l)
  t3 <- the sum of its six simultaneous permutations of (a,i), (b,j), (c,k)
  v3(a,b,c,i,j,k) = g_abij(a,b,i,j) t1(c,k);  v3 <- v3 + v3(b,c,a,j,k,i) + v3(a,c,b,i,k,j)
  E(T) = 1/3 sum (t3 + v3)(4 t3 + t3(abckij) + t3(abcjki)
                            - 2 (t3(abckji) + t3(abcikj) + t3(abcjik))) / D
  D = e_i + e_j + e_k - e_a - e_b - e_c

Pinning (tests/test_oracle_ccsd_t.py): the reference's golden (T) energy of
h2o-ccsd_t-631g-pvdz.out (DF integrals, frozen core) to 1e-12 Eh, and the
spin-orbital (T) of the textbook equations (tests/spinorb_ccsd.py, independent
code) at converged spin-orbital CCSD amplitudes of random Hamiltonians.
"""
from __future__ import annotations

import numpy as np

__all__ = ["ccsd_t_straight", "ccsd_t_flops"]


def _e(subscripts, *ops):
    return np.einsum(subscripts, *ops, optimize=True)


def ccsd_t_straight(t1, t2, Gijab, Giabc, Gijka, eps_occ, eps_vir):
    """E(T) of ccsd_t.h:1127-1167 from converged t1(a,i), t2(a,b,i,j) and the
    physical-notation blocks <ij|ab>, <ia|bc>, <ij|ka>."""
    g_cjkl = _e("ijka->aijk", Gijka)   # get_aijk: result("a,i,j,k") = <i j|G|k a>   (:2210-2220)
    g_dabi = _e("iabc->abci", Giabc)   # get_abci: result("a,b,c,i") = <i a|G|b c>   (:2224-2234)
    g_abij = _e("ijab->abij", Gijab)   # get_abij                                     (:2238-2244)
    # :1141-1142
    t3 = _e("dabi,dcjk->abcijk", g_dabi, t2) - _e("cjkl,abil->abcijk", g_cjkl, t2)
    # :1143-1145
    t3 = (t3 + _e("acbikj->abcijk", t3) + _e("cabkij->abcijk", t3) + _e("cbakji->abcijk", t3)
          + _e("bcajki->abcijk", t3) + _e("bacjik->abcijk", t3))
    # :1148-1150
    v3 = _e("abij,ck->abcijk", g_abij, t1)
    v3 = v3 + _e("bcajki->abcijk", v3) + _e("acbikj->abcijk", v3)
    # :1160-1165 with CCSD_T_Reduce (:2314-2333): divide each element by D
    x = (4.0 * t3 + _e("abckij->abcijk", t3) + _e("abcjki->abcijk", t3)
         - 2.0 * (_e("abckji->abcijk", t3) + _e("abcikj->abcijk", t3) + _e("abcjik->abcijk", t3)))
    eo, ev = np.asarray(eps_occ), np.asarray(eps_vir)
    D = (eo[None, None, None, :, None, None] + eo[None, None, None, None, :, None]
         + eo[None, None, None, None, None, :] - ev[:, None, None, None, None, None]
         - ev[None, :, None, None, None, None] - ev[None, None, :, None, None, None])
    return float(np.sum((t3 + v3) * x / D)) / 3.0


def ccsd_t_flops(o: int, v: int) -> float:
    """Contraction FLOPs of the t3 build (the two terms of ccsd_t.h:1141-1142)."""
    return 2.0 * o**3 * v**4 + 2.0 * o**4 * v**3
