"""CPU oracle for MPQC's closed-shell CCSD iteration (numpy restatement).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline.  The product path (``mpqc_amd``) never
imports it.

Every function restates the reference algorithm term by term, with the same
index orders, coefficients and evaluation order, and cites the reference line
it follows (paths relative to the MPQC source tree):

* ``Integrals``            src/mpqc/chemistry/qc/lcao/cc/ccsd_r1_r2.h:24-56
* ``F_oo/F_vv/F_ov``       src/mpqc/chemistry/qc/lcao/cc/ccsd_intermediates.h:57-95
* ``F_OO/F_VV``            ccsd_intermediates.h:105-112, 141-148
* ``W_oooo``               ccsd_intermediates.h:182-193
* ``cs_ccsd_r1``           ccsd_r1_r2.h:70-124   (``u`` given: the direct method's terms)
* ``cs_ccsd_r2``           ccsd_r1_r2.h:198-293  (``u`` given: the direct method's terms)
* ``jacobi_update_*``      src/mpqc/chemistry/qc/lcao/cc/solvers.h:165-258
* ``error``                src/mpqc/chemistry/qc/cc/solvers.h:60-63
* ``is_converged``         src/mpqc/chemistry/qc/cc/solvers.h:141-152
* ``energy``               src/mpqc/chemistry/qc/lcao/cc/ccsd.h:376-380
* ``compute_ccsd``         src/mpqc/chemistry/qc/lcao/cc/ccsd.h:268-477
* ``DIIS``                 see ``oracle/diis.py`` (TiledArray's ``TA::DIIS``)

Tensor conventions (all float64, dense, row-major): ``t1[a,i]``,
``t2[a,b,i,j]``, ``Fia[i,a]``, ``Fij[i,j]``, ``Fab[a,b]``,
``Gijab[i,j,a,b] = <ij|ab>``, ``Gijkl``, ``Giajb[i,a,j,b]``,
``Giabc[i,a,b,c]``, ``Gijka[i,j,k,a]``, ``Gabcd[a,b,c,d]`` (physicist
notation, ``<pq|rs> = (pr|qs)``).

Pinning (what checks this module, all under ``tests/``): golden final CCSD
energies and per-iteration traces from the reference's validation outputs;
spin-orbital CCSD (textbook equations, independent code) on random Hamiltonians;
two-electron CCSD == FCI; brute-force per-term einsum on tiny shapes;
closed forms (t=0 residuals, MP2 first step).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional

import numpy as np

from .diis import DIIS


def _e(subscripts: str, *operands):
    return np.einsum(subscripts, *operands, optimize=True)


@dataclasses.dataclass
class Integrals:
    """Mirror of ``cc::Integrals`` (ccsd_r1_r2.h:24-56), standard method."""

    Fia: np.ndarray
    Fij: np.ndarray
    Fab: np.ndarray
    Gijab: np.ndarray
    Gijkl: np.ndarray
    Giajb: np.ndarray
    Giabc: Optional[np.ndarray]
    Gijka: np.ndarray
    Gabcd: Optional[np.ndarray]
    # effective one-body Hamiltonian, written by cs_ccsd_r1 (ccsd_r1_r2.h:90,96)
    FIJ: Optional[np.ndarray] = None
    FAB: Optional[np.ndarray] = None
    # active occupied / virtual MO coefficients C[mu, p] of the direct method (ccsd.h:313-322)
    Ci: Optional[np.ndarray] = None
    Ca: Optional[np.ndarray] = None

    @property
    def o(self) -> int:
        return self.Fij.shape[0]

    @property
    def v(self) -> int:
        return self.Fab.shape[0]


# --------------------------------------------------------------------------
# one-particle intermediates (ccsd_intermediates.h)
# --------------------------------------------------------------------------
def F_oo(f_oo, g_ijab_bar, tau):
    """ccsd_intermediates.h:57-63: F_oo(k,i) = f(k,i) + gbar(k,l,a,b) tau(a,b,i,l)."""
    return f_oo + _e("klab,abil->ki", g_ijab_bar, tau)


def F_vv(f_vv, g_ijab_bar, tau):
    """ccsd_intermediates.h:71-77: F_vv(a,c) = f(a,c) - gbar(k,l,c,d) tau(a,d,k,l)."""
    return f_vv - _e("klcd,adkl->ac", g_ijab_bar, tau)


def F_ov(f_ov, g_ijab_bar, t1):
    """ccsd_intermediates.h:89-95: F_ov(k,c) = f(k,c) + gbar(k,l,c,d) t1(d,l)."""
    return f_ov + _e("klcd,dl->kc", g_ijab_bar, t1)


def F_OO(F_oo_, f_ov, g_ijka_bar, t1):
    """ccsd_intermediates.h:105-112 (4-argument overload)."""
    return F_oo_ + f_ov @ t1 + _e("klic,cl->ki", g_ijka_bar, t1)


def F_VV(F_vv_, f_ov, g_iabc_bar, t1, fov_sign=1.0):
    """ccsd_intermediates.h:141-148 (4-argument overload).

    F_VV(a,c) = F_vv(a,c) + f_ov(k,c) t1(a,k) + gbar(k,a,d,c) t1(d,k).
    ``fov_sign=+1`` follows the reference as written.  Exact CCSD (equality
    with spin-orbital CCSD when f_ov != 0, i.e. a non-Brillouin reference)
    needs ``fov_sign=-1``; for Hartree-Fock references f_ov = 0 and the two
    agree.
    """
    return F_vv_ + fov_sign * (t1 @ f_ov) + _e("kadc,dk->ac", g_iabc_bar, t1)


def W_oooo(g_ijkl, g_ijka, g_ijab, tau, t1):
    """ccsd_intermediates.h:182-193."""
    return (g_ijkl + _e("klic,cj->klij", g_ijka, t1)
            + _e("lkjc,ci->klij", g_ijka, t1)
            + _e("klcd,cdij->klij", g_ijab, tau))


# --------------------------------------------------------------------------
# residuals
# --------------------------------------------------------------------------
def make_tau(t1, t2):
    """ccsd.h:342,389: tau(a,b,i,j) = t2(a,b,i,j) + t1(a,i) t1(b,j)."""
    return t2 + _e("ai,bj->abij", t1, t1)


def cs_ccsd_r1(t1, t2, tau, ints: Integrals, u=None):
    """ccsd_r1_r2.h:70-124.  Returns r1(a,i); writes ints.FAB and ints.FIJ.
    ``u`` (U(p,r,i,j) of the direct method, AO indices p, r) replaces the
    g_iabc_bar tau term (:110-118)."""
    Fia = ints.Fia
    # :83-84
    r1 = Fia.T - 2.0 * _e("kc,ci,ak->ai", Fia, t1, t1)
    # :87
    g_ijab_bar = 2.0 * ints.Gijab - ints.Gijab.transpose(1, 0, 2, 3)
    # :89-91
    h_ac = F_vv(ints.Fab, g_ijab_bar, tau)
    ints.FAB = h_ac
    r1 = r1 + h_ac @ t1
    # :95-97
    h_ki = F_oo(ints.Fij, g_ijab_bar, tau)
    ints.FIJ = h_ki
    r1 = r1 - t1 @ h_ki
    # :101-104
    h_kc = F_ov(Fia, g_ijab_bar, t1)
    r1 = r1 + (2.0 * _e("kc,caki->ai", h_kc, t2) - _e("kc,caik->ai", h_kc, t2)
               + _e("kc,ci,ak->ai", h_kc, t1, t1))
    # :107-108
    r1 = r1 + (2.0 * _e("kica,ck->ai", ints.Gijab, t1) - _e("kaic,ck->ai", ints.Giajb, t1))
    if u is None:
        # :110-114
        g_iabc_bar = 2.0 * ints.Giabc - ints.Giabc.transpose(0, 1, 3, 2)
        r1 = r1 + _e("kacd,cdki->ai", g_iabc_bar, tau)
    else:
        # :116-117
        r1 = r1 + _e("prki,pk,ra->ai", 2.0 * u - u.transpose(0, 1, 3, 2), ints.Ci, ints.Ca)
    # :120-121
    g_lkic_bar = 2.0 * ints.Gijka - ints.Gijka.transpose(1, 0, 2, 3)  # (l,k,i,c)
    r1 = r1 - _e("lkic,cakl->ai", g_lkic_bar, tau)
    return r1


def cs_ccsd_r2(t1, t2, tau, ints: Integrals, fov_sign=1.0, u=None):
    """ccsd_r1_r2.h:198-293.  Requires ints.FIJ / ints.FAB from cs_ccsd_r1.
    ``fov_sign``: see ``F_VV``.  ``u`` (direct method) replaces the
    particle-particle ladder and its Giabc terms (:282-286)."""
    G = ints
    # "permutation part" :206-212
    r2 = _e("icab,cj->abij", G.Giabc, t1) - _e("kbic,ak,cj->abij", G.Giajb, t1, t1)
    r2 = r2 - (_e("jika,bk->abij", G.Gijka, t1) + _e("ikac,cj,bk->abij", G.Gijab, t1, t1))
    # g intermediates :218-226
    g_ijka_bar = 2.0 * G.Gijka - G.Gijka.transpose(1, 0, 2, 3)
    g_ki = F_OO(G.FIJ, G.Fia, g_ijka_bar, t1)
    g_iabc_bar = 2.0 * G.Giabc - G.Giabc.transpose(0, 1, 3, 2)
    g_ac = F_VV(G.FAB, G.Fia, g_iabc_bar, t1, fov_sign)
    r2 = r2 + _e("ac,cbij->abij", g_ac, t2) - _e("ki,abkj->abij", g_ki, t2)
    # j and k intermediates :233-249
    T = 0.5 * t2 + _e("di,bl->dbil", t1, t1)
    g_klcd_bar = 2.0 * G.Gijab - G.Gijab.transpose(0, 1, 3, 2)
    j_akic = (G.Gijab.transpose(2, 1, 0, 3)            # Gijab(i,k,a,c) -> (a,k,i,c)
              - _e("lkic,al->akic", G.Gijka, t1)
              + _e("kacd,di->akic", G.Giabc, t1)
              - _e("klcd,dail->akic", G.Gijab, T)
              + 0.5 * _e("klcd,adil->akic", g_klcd_bar, t2))
    k_kaic = (G.Giajb
              - _e("klic,al->kaic", G.Gijka, t1)
              + _e("kadc,di->kaic", G.Giabc, t1)
              - _e("kldc,dail->kaic", G.Gijab, T))
    # :251-252
    x_akic = 2.0 * j_akic - k_kaic.transpose(1, 0, 2, 3)
    y_cbkj = 2.0 * t2 - t2.transpose(1, 0, 2, 3)
    r2 = r2 + 0.5 * _e("akic,cbkj->abij", x_akic, y_cbkj)
    # :254-256
    tmp = _e("kaic,cbjk->abij", k_kaic, t2)
    r2 = r2 - (0.5 * tmp + tmp.transpose(1, 0, 2, 3))
    # permutation :260
    r2 = r2 + r2.transpose(1, 0, 3, 2)
    # :262
    r2 = r2 + G.Gijab.transpose(2, 3, 0, 1)
    # hole-hole ladder :266-269
    a_klij = W_oooo(G.Gijkl, G.Gijka, G.Gijab, tau, t1)
    r2 = r2 + _e("klij,abkl->abij", a_klij, tau)
    if u is None:
        # particle-particle ladder :275-280
        b_abij = _e("cdij,abcd->abij", tau, G.Gabcd)
        tmp = _e("kacd,cdij->kaij", G.Giabc, tau)
        b_abij = b_abij - _e("kaji,bk->abij", tmp, t1)
        b_abij = b_abij - _e("kbij,ak->abij", tmp, t1)
    else:
        # :282-286: (U Ca(r,b) - Ci(r,k) t1(b,k) U) Ca(p,a) - U Ci(p,k) Ca(r,b) t1(a,k)
        b_abij = (_e("prij,rb,pa->abij", u, G.Ca, G.Ca)
                  - _e("rk,bk,prij,pa->abij", G.Ci, t1, u, G.Ca)
                  - _e("prij,pk,rb,ak->abij", u, G.Ci, G.Ca, t1))
    r2 = r2 + b_abij
    return r2


def pp_ladder(tau, Gabcd):
    """The particle-particle ladder alone (ccsd_r1_r2.h:275):
    b(a,b,i,j) = tau(c,d,i,j) * Gabcd(a,b,c,d)."""
    o = tau.shape[2]
    v = tau.shape[0]
    return (Gabcd.reshape(v * v, v * v) @ tau.reshape(v * v, o * o)).reshape(v, v, o, o)


# --------------------------------------------------------------------------
# solver pieces
# --------------------------------------------------------------------------
def jacobi_update_t1(r1, eps_occ, eps_vir):
    """lcao/cc/solvers.h:223-258: delta_t1(a,i) = r1(a,i) / (e_i - e_a)."""
    return r1 / (eps_occ[None, :] - eps_vir[:, None])


def jacobi_update_t2(r2, eps_occ, eps_vir):
    """lcao/cc/solvers.h:165-210: delta_t2(a,b,i,j) = r2 / (e_i + e_j - e_a - e_b)."""
    d = (eps_occ[None, None, :, None] + eps_occ[None, None, None, :]
         - eps_vir[:, None, None, None] - eps_vir[None, :, None, None])
    return r2 / d


def error(r1, r2):
    """qc/cc/solvers.h:60-63: sqrt(|r1|^2 + |r2|^2) / (size(r1) + size(r2))."""
    n1 = np.linalg.norm(r1.ravel())
    n2 = np.linalg.norm(r2.ravel())
    return math.sqrt(n1 * n1 + n2 * n2) / (r1.size + r2.size)


def is_converged(target_precision, err, dE):
    """qc/cc/solvers.h:141-152 (precision_margin_energy = 0.1)."""
    return dE <= target_precision * 0.1 and err <= target_precision


def energy(Fia, Gijab, t1, tau, r1, r2):
    """ccsd.h:376-380:
    E = 2 <f_ai + r1, t1> + <g_abij + r2, 2 tau(a,b,i,j) - tau(b,a,i,j)>."""
    f_ai = Fia.T
    g_abij = Gijab.transpose(2, 3, 0, 1)
    return (2.0 * float(np.vdot(f_ai + r1, t1))
            + float(np.vdot(g_abij + r2, 2.0 * tau - tau.transpose(1, 0, 2, 3))))


class JacobiDIISSolver:
    """lcao/cc/solvers.h:1128-1169 on top of qc/cc/solvers.h:76-180.

    KeyVal defaults (qc/cc/solvers.h:93-95): diis_start=1, n_diis=8,
    diis_damp=0, diis_ngroup=1, diis_group_nstart=1.
    """

    def __init__(self, eps_occ, eps_vir, diis_start=1, n_diis=8, diis_damp=0.0,
                 diis_ngroup=1, diis_group_nstart=1, **diis_kw):
        self.eps_occ = np.asarray(eps_occ, dtype=np.float64)
        self.eps_vir = np.asarray(eps_vir, dtype=np.float64)
        self.diis = DIIS(diis_start, n_diis, diis_damp, diis_ngroup, diis_group_nstart,
                         **diis_kw)

    error = staticmethod(error)
    is_converged = staticmethod(is_converged)

    def update_only(self, t1, t2, r1, r2):
        """lcao/cc/solvers.h:1153-1158 (truncate() is a no-op for dense data)."""
        t1 = t1 + jacobi_update_t1(r1, self.eps_occ, self.eps_vir)
        t2 = t2 + jacobi_update_t2(r2, self.eps_occ, self.eps_vir)
        return t1, t2

    def update(self, t1, t2, r1, r2):
        """qc/cc/solvers.h:112-121: Jacobi update, then DIIS extrapolation with
        the pre-update residuals as error vectors."""
        t1, t2 = self.update_only(t1, t2, r1, r2)
        x = [t1, t2]
        self.diis.extrapolate(x, [r1.copy(), r2.copy()])
        return x[0], x[1]


@dataclasses.dataclass
class IterRow:
    iter: int
    dE: float
    error: float
    energy: float


def compute_ccsd(ints: Integrals, eps_occ, eps_vir, max_iter=30,
                 target_precision=1e-9, solver: Optional[JacobiDIISSolver] = None,
                 callback: Optional[Callable] = None, residual_fn=None,
                 fov_sign=1.0):
    """ccsd.h:268-477 with method = "standard".

    Returns (E_corr, t1, t2, rows) where ``rows`` holds the printed iteration
    table (ccsd.h:28-36): (iter, dE, error, energy).  ``callback`` (if given) is
    called as callback(iter, t1, t2, r1, r2, E) after each energy evaluation.
    ``residual_fn(t1, t2, tau, ints) -> (r1, r2)`` overrides the residual.
    """
    if solver is None:
        solver = JacobiDIISSolver(eps_occ, eps_vir)
    if residual_fn is None:
        def residual_fn(t1_, t2_, tau_, ints_):
            r1_ = cs_ccsd_r1(t1_, t2_, tau_, ints_)
            return r1_, cs_ccsd_r2(t1_, t2_, tau_, ints_, fov_sign)
    o, v = ints.o, ints.v
    f_ai = ints.Fia.T.copy()                       # :331-332
    g_abij = ints.Gijab.transpose(2, 3, 0, 1).copy()  # :336-337
    t1 = np.zeros((v, o))                          # :334-335
    t2 = np.zeros((v, v, o, o))                    # :338-339
    tau = make_tau(t1, t2)                         # :341-342
    E1 = 0.0
    it = 0
    r1 = f_ai.copy()                               # :351
    r2 = g_abij.copy()                             # :352
    rows: List[IterRow] = []
    while it < max_iter:                           # :367
        err = solver.error(r1, r2)                 # :373
        E0 = E1                                    # :376
        E1 = (2.0 * float(np.vdot(f_ai + r1, t1))
              + float(np.vdot(g_abij + r2, 2.0 * tau - tau.transpose(1, 0, 2, 3))))
        dE = abs(E0 - E1)                          # :380
        if callback is not None:
            callback(it, t1, t2, r1, r2, E1)
        if it == 0 or not solver.is_converged(target_precision, err, dE):  # :382
            t1, t2 = solver.update(t1, t2, r1, r2)  # :386
            tau = make_tau(t1, t2)                  # :389
            if it > 0:
                rows.append(IterRow(it, dE, err, E1))  # :399-406
        else:
            rows.append(IterRow(it, dE, err, E1))   # :410-415
            break
        r1, r2 = residual_fn(t1, t2, tau, ints)     # :433, :452
        it += 1                                      # :467
    return E1, t1, t2, rows


def ccsd_flops(o: int, v: int) -> float:
    """Algorithmic FLOPs of one residual evaluation of the reference algorithm
    (2 * prod(extents) per binary contraction, counted from ccsd_r1_r2.h:70-293)."""
    return float(2 * o**2 * v**4 + 12 * o**3 * v**3 + 4 * o**4 * v**2 + 14 * o**2 * v**3
                 + 18 * o**3 * v**2 + 4 * o**4 * v + 2 * o * v**3 + 6 * o**2 * v**2
                 + 2 * o**3 * v + 4 * o * v**2 + 8 * o**2 * v)
