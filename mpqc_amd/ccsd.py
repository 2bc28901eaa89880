"""Host-side mirror of MPQC's CCSD solver interface on top of the C-ABI.

Names, keywords and argument meanings follow the reference
(src/mpqc/chemistry/qc/lcao/cc/ccsd.h:65-115 keywords, ccsd.h:268-477 loop,
ccsd_r1_r2.h:24-56 ``cc::Integrals``, qc/cc/solvers.h:93-95 DIIS keywords,
lcao/wfn/lcao_wfn.h:53-72 block sizes, lcao/expression/trange1_engine.cpp:10-20
tiling).  All arithmetic runs in ``libmpqc_ccsd.so`` on the GPU.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L

PRECISIONS = {"f64": L.PREC_F64, "f32": L.PREC_F32, "bf16x3": L.PREC_BF16X3,
              "f16x3": L.PREC_F16X3, "bf16x6": L.PREC_BF16X6}


def compute_trange1(n: int, block_size: int) -> List[int]:
    """Hashmarks of a TiledRange1 (trange1_engine.cpp:10-20): uniform blocks of
    ``block_size`` with a short last tile."""
    if block_size <= 0:
        raise ValueError("block_size must be positive")
    marks = [0]
    i = block_size
    while i < n:
        marks.append(i)
        i += block_size
    marks.append(n)
    return marks


def to_ta_tiles(x: np.ndarray, hashmarks: Sequence[Sequence[int]]) -> np.ndarray:
    """Dense row-major tensor -> concatenation of row-major tiles in row-major
    tile-ordinal order (the TA::DistArray tile layout)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = []
    grids = [list(zip(hm[:-1], hm[1:])) for hm in hashmarks]
    import itertools
    for tiles in itertools.product(*grids):
        sl = tuple(slice(lo, hi) for lo, hi in tiles)
        out.append(x[sl].ravel())
    return np.concatenate(out) if out else np.zeros(0)


def from_ta_tiles(buf: np.ndarray, shape, hashmarks: Sequence[Sequence[int]]) -> np.ndarray:
    import itertools
    x = np.empty(shape)
    grids = [list(zip(hm[:-1], hm[1:])) for hm in hashmarks]
    off = 0
    for tiles in itertools.product(*grids):
        sl = tuple(slice(lo, hi) for lo, hi in tiles)
        n = int(np.prod([hi - lo for lo, hi in tiles]))
        x[sl] = buf[off:off + n].reshape([hi - lo for lo, hi in tiles])
        off += n
    return x


@dataclasses.dataclass
class Integrals:
    """``cc::Integrals<Array>`` (ccsd_r1_r2.h:24-56) as dense host arrays in
    the reference's index orders."""
    Fia: np.ndarray
    Fij: np.ndarray
    Fab: np.ndarray
    Gijab: np.ndarray
    Gijkl: np.ndarray
    Giajb: np.ndarray
    Giabc: np.ndarray
    Gijka: np.ndarray
    Gabcd: np.ndarray

    @property
    def o(self):
        return self.Fij.shape[0]

    @property
    def v(self):
        return self.Fab.shape[0]

    def items(self):
        return [(k, getattr(self, k)) for k in L.INTEGRAL_IDS]


@dataclasses.dataclass
class TraceRow:
    iter: int
    dE: float
    error: float
    energy: float
    seconds: float


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class CCSD:
    """Closed-shell CCSD on one GPU with MPQC's solver semantics.

    Keywords mirror the reference: ``max_iter`` (30), ``occ_block_size`` /
    ``unocc_block_size`` (24, lcao_wfn.h:68-70; tile layout of host buffers),
    DIIS ``diis_start``=1, ``n_diis``=8, ``diis_damp``=0, ``diis_ngroup``=1,
    ``diis_group_nstart``=1 (qc/cc/solvers.h:93-95), plus ``precision`` for the
    contractions ("f64", "f32", "bf16x3").
    """

    def __init__(self, o: int, v: int, precision: str = "f64", max_iter: int = 30,
                 occ_block_size: int = 24, unocc_block_size: int = 24, diis_start: int = 1,
                 n_diis: int = 8, diis_damp: float = 0.0, diis_ngroup: int = 1,
                 diis_group_nstart: int = 1, fov_sign: float = 1.0,
                 occ_hashmarks: Optional[Sequence[int]] = None,
                 vir_hashmarks: Optional[Sequence[int]] = None):
        self.lib = L.load()
        self.o, self.v = int(o), int(v)
        self.max_iter = int(max_iter)
        # a TiledRange1 given as hashmarks (e.g. uneven tiles) replaces the block size
        self.occ_hm = (list(map(int, occ_hashmarks)) if occ_hashmarks is not None
                       else compute_trange1(self.o, occ_block_size))
        self.vir_hm = (list(map(int, vir_hashmarks)) if vir_hashmarks is not None
                       else compute_trange1(self.v, unocc_block_size))
        occ = (ctypes.c_int64 * len(self.occ_hm))(*self.occ_hm)
        vir = (ctypes.c_int64 * len(self.vir_hm))(*self.vir_hm)
        ctx = ctypes.c_void_p()
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        st = self.lib.mpqc_ccsd_ctx_create(self.o, self.v, occ, len(self.occ_hm) - 1, vir,
                                           len(self.vir_hm) - 1, PRECISIONS[precision],
                                           ctypes.byref(ctx))
        L.check(st, None)
        self.ctx = ctx
        self._check(self.lib.mpqc_ccsd_solver_init(self.ctx, diis_start, n_diis, diis_damp,
                                                   diis_ngroup, diis_group_nstart))
        if fov_sign != 1.0:
            self.set_option("fov_sign", fov_sign)

    # -- lifecycle
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mpqc_ccsd_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, status):
        L.check(status, self.ctx)

    # -- configuration
    def set_option(self, key: str, value: float):
        self._check(self.lib.mpqc_ccsd_set_option(self.ctx, key.encode(), float(value)))

    def _modes(self, name):
        O, V = self.occ_hm, self.vir_hm
        K = getattr(self, "aux_hm", None)
        return {"Fia": [O, V], "Fij": [O, O], "Fab": [V, V], "Gijab": [O, O, V, V],
                "Gijkl": [O, O, O, O], "Giajb": [O, V, O, V], "Giabc": [O, V, V, V],
                "Gijka": [O, O, O, V], "Gabcd": [V, V, V, V], "t1": [V, O],
                "t2": [V, V, O, O], "Xai": [K, V, O], "Xij": [K, O, O], "Xab": [K, V, V]}[name]

    def set_df_basis(self, naux: int, aux_block_size: int = 0):
        """Density fitting with ``naux`` fitting functions: selects the "df"
        residual (ccsd.h:93-99).  ``aux_block_size`` tiles the fitting index
        of TA-layout buffers (0: one tile)."""
        self.naux = int(naux)
        self.aux_hm = compute_trange1(self.naux, aux_block_size) if aux_block_size else [0, self.naux]
        hm = (ctypes.c_int64 * len(self.aux_hm))(*self.aux_hm)
        self._check(self.lib.mpqc_ccsd_set_df_basis(self.ctx, self.naux, hm, len(self.aux_hm) - 1))

    def set_direct_ao(self, ao_eri, c_occ, c_vir):
        """AO integral-direct method (ccsd.h:73): AO integrals (mu nu|la si) and the
        active occupied / virtual MO coefficients C[mu, p]; selects the "direct"
        residual (U = (pq|rs) Ca tau Ca^T each iteration, ccsd.h:573-603)."""
        eri = np.ascontiguousarray(ao_eri, dtype=np.float64)
        ci = np.ascontiguousarray(c_occ, dtype=np.float64)
        ca = np.ascontiguousarray(c_vir, dtype=np.float64)
        nbf = eri.shape[0]
        if eri.shape != (nbf,) * 4 or ci.shape != (nbf, self.o) or ca.shape != (nbf, self.v):
            raise ValueError("set_direct_ao: shapes must be (nbf,)*4, (nbf, o), (nbf, v)")
        vp = ctypes.c_void_p
        self._check(self.lib.mpqc_ccsd_set_direct_ao(self.ctx, nbf, eri.ctypes.data_as(vp), ci.ctypes.data_as(vp),
                                                     ca.ctypes.data_as(vp), L.MEM_HOST))

    def set_integral(self, name: str, arr: np.ndarray, layout: str = "dense"):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if layout == "ta":
            a = to_ta_tiles(a, self._modes(name))
        ids = {**L.INTEGRAL_IDS, **L.DF_INTEGRAL_IDS}
        self._check(self.lib.mpqc_ccsd_set_integral(
            self.ctx, ids[name], a.ctypes.data_as(ctypes.c_void_p),
            L.LAYOUT_TA_TILES if layout == "ta" else L.LAYOUT_DENSE, L.MEM_HOST))

    def set_integral_device(self, name: str, dev_ptr: int, layout: str = "dense"):
        """A block already in device memory at ``dev_ptr`` (dense, or TA tiles of
        this context's TiledRange1s)."""
        self._check(self.lib.mpqc_ccsd_set_integral(self.ctx, L.INTEGRAL_IDS[name],
                                                    ctypes.c_void_p(dev_ptr),
                                                    L.LAYOUT_TA_TILES if layout == "ta" else L.LAYOUT_DENSE,
                                                    L.MEM_DEVICE))

    def set_integrals(self, ints, layout: str = "dense"):
        """All blocks present in ``ints`` (an Integrals, or a dict of the
        cc::Integrals member names, e.g. the df set without Giabc / Gabcd)."""
        get = ints.get if isinstance(ints, dict) else (lambda k: getattr(ints, k, None))
        for name in list(L.INTEGRAL_IDS) + list(L.DF_INTEGRAL_IDS):
            a = get(name)
            if a is not None:
                self.set_integral(name, a, layout)

    def triples(self):
        """(T) correction from the context's amplitudes (CCSD_T::compute_ccsd_t,
        ccsd_t.h:144-176).  Returns (E(T), device seconds)."""
        e = ctypes.c_double()
        sec = ctypes.c_double()
        self._check(self.lib.mpqc_ccsd_triples(self.ctx, ctypes.byref(e), ctypes.byref(sec)))
        return e.value, sec.value

    def eom_ee(self, n_roots: int, target_precision: float = 1e-8, max_iter: int = 100,
               verbose: bool = False):
        """EOM-EE-CCSD singlet excitation energies from the context's amplitudes
        (EOM_CCSD::evaluate, eom_ccsd_impl.h:395-478, with the Davidson solver of
        :236-390; target_precision is the reference's convergence of the roots).
        Returns (energies ascending, iterations)."""
        e = np.empty(int(n_roots))
        it = ctypes.c_int()
        self._check(self.lib.mpqc_ccsd_eom_ee(self.ctx, int(n_roots), int(max_iter), float(target_precision),
                                              int(bool(verbose)), e.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.byref(it)))
        return e, it.value

    def eom_stats(self):
        """(J^T products, their device seconds) of the last eom_ee."""
        n = ctypes.c_int()
        sec = ctypes.c_double()
        self._check(self.lib.mpqc_ccsd_eom_stats(self.ctx, ctypes.byref(n), ctypes.byref(sec)))
        return n.value, sec.value

    def get_integral(self, name: str, layout: str = "dense") -> np.ndarray:
        """An integral block held by the context (as set or as transformed)."""
        o, v, k = self.o, self.v, getattr(self, "naux", 0)
        shapes = {"Fia": (o, v), "Fij": (o, o), "Fab": (v, v), "Gijab": (o, o, v, v),
                  "Gijkl": (o, o, o, o), "Giajb": (o, v, o, v), "Giabc": (o, v, v, v),
                  "Gijka": (o, o, o, v), "Gabcd": (v, v, v, v), "Xai": (k, v, o), "Xij": (k, o, o),
                  "Xab": (k, v, v)}
        ids = {**L.INTEGRAL_IDS, **L.DF_INTEGRAL_IDS}
        buf = np.empty(int(np.prod(shapes[name])))
        self._check(self.lib.mpqc_ccsd_get_integral(self.ctx, ids[name], buf.ctypes.data_as(ctypes.c_void_p),
                                                    L.LAYOUT_TA_TILES if layout == "ta" else L.LAYOUT_DENSE,
                                                    L.MEM_HOST))
        if layout == "ta":
            return from_ta_tiles(buf, shapes[name], self._modes(name))
        return buf.reshape(shapes[name])

    def transform_ao(self, eri, fock_ao, c_occ, c_vir):
        """Every integral block from AO (mu nu|la si), the AO Fock matrix and the
        active occupied / virtual MO coefficients, transformed on the device
        (LCAOFactory::compute4, lcao_factory.h:378-497)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (eri, fock_ao, c_occ, c_vir)]
        nbf = arrs[1].shape[0]
        if arrs[0].shape != (nbf,) * 4 or arrs[2].shape != (nbf, self.o) or arrs[3].shape != (nbf, self.v):
            raise ValueError("AO integral / coefficient shapes do not match nbf, o, v")
        self._check(self.lib.mpqc_ccsd_transform_ao(self.ctx, nbf, *[a.ctypes.data_as(ctypes.c_void_p)
                                                                      for a in arrs], L.MEM_HOST))

    def transform_df(self, three_center, metric, fock_ao, c_occ, c_vir):
        """The df integral set from (K|mu nu) and a metric factor L (L L^T = V^-1,
        or None when ``three_center`` is already fitted), on the device."""
        B = np.ascontiguousarray(three_center, dtype=np.float64)
        F = np.ascontiguousarray(fock_ao, dtype=np.float64)
        Co = np.ascontiguousarray(c_occ, dtype=np.float64)
        Cv = np.ascontiguousarray(c_vir, dtype=np.float64)
        nbf = F.shape[0]
        if B.shape != (self.naux, nbf, nbf) or Co.shape != (nbf, self.o) or Cv.shape != (nbf, self.v):
            raise ValueError("three-centre integral / coefficient shapes do not match naux, nbf, o, v")
        Lm = None if metric is None else np.ascontiguousarray(metric, dtype=np.float64)
        if Lm is not None and Lm.shape != (self.naux, self.naux):
            raise ValueError("metric factor must be naux x naux")
        vp = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        self._check(self.lib.mpqc_ccsd_transform_df(self.ctx, nbf, vp(B), vp(Lm), vp(F), vp(Co), vp(Cv),
                                                    L.MEM_HOST))

    def set_orbital_energies(self, eps_occ_act, eps_vir):
        eo = np.ascontiguousarray(eps_occ_act, dtype=np.float64)
        ev = np.ascontiguousarray(eps_vir, dtype=np.float64)
        if eo.shape != (self.o,) or ev.shape != (self.v,):
            raise ValueError("orbital energy shapes do not match o, v")
        self._check(self.lib.mpqc_ccsd_set_orbital_energies(self.ctx, _dptr(eo), _dptr(ev)))

    # -- amplitudes
    def set_amplitudes(self, t1, t2, layout: str = "dense"):
        a = np.ascontiguousarray(t1, dtype=np.float64)
        b = np.ascontiguousarray(t2, dtype=np.float64)
        if layout == "ta":
            a = to_ta_tiles(a, self._modes("t1"))
            b = to_ta_tiles(b, self._modes("t2"))
        self._check(self.lib.mpqc_ccsd_set_amplitudes(
            self.ctx, a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
            L.LAYOUT_TA_TILES if layout == "ta" else L.LAYOUT_DENSE, L.MEM_HOST))

    def _get2(self, fn, layout):
        o, v = self.o, self.v
        a = np.empty(v * o)
        b = np.empty(v * v * o * o)
        self._check(fn(self.ctx, a.ctypes.data_as(ctypes.c_void_p),
                       b.ctypes.data_as(ctypes.c_void_p),
                       L.LAYOUT_TA_TILES if layout == "ta" else L.LAYOUT_DENSE, L.MEM_HOST))
        if layout == "ta":
            return (from_ta_tiles(a, (v, o), self._modes("t1")),
                    from_ta_tiles(b, (v, v, o, o), self._modes("t2")))
        return a.reshape(v, o), b.reshape(v, v, o, o)

    def get_amplitudes(self, layout: str = "dense"):
        return self._get2(self.lib.mpqc_ccsd_get_amplitudes, layout)

    def get_residuals(self, layout: str = "dense"):
        return self._get2(self.lib.mpqc_ccsd_get_residuals, layout)

    def get_dressed_fock(self):
        FIJ = np.empty((self.o, self.o))
        FAB = np.empty((self.v, self.v))
        self._check(self.lib.mpqc_ccsd_get_dressed_fock(self.ctx, FIJ.ctypes.data_as(ctypes.c_void_p),
                                                        FAB.ctypes.data_as(ctypes.c_void_p)))
        return FIJ, FAB

    # -- the operator API of the reference
    def compute_residuals(self, t1, t2):
        """cc::compute_cs_ccsd_r1 + compute_cs_ccsd_r2 at (t1, t2); returns
        (r1, r2)."""
        self.set_amplitudes(t1, t2)
        self._check(self.lib.mpqc_ccsd_residual(self.ctx))
        return self.get_residuals()

    def init_guess(self):
        self._check(self.lib.mpqc_ccsd_init_guess(self.ctx))

    def residual(self):
        self._check(self.lib.mpqc_ccsd_residual(self.ctx))

    def energy_error(self):
        e = ctypes.c_double()
        err = ctypes.c_double()
        self._check(self.lib.mpqc_ccsd_energy_error(self.ctx, ctypes.byref(e), ctypes.byref(err)))
        return e.value, err.value

    def update(self):
        self._check(self.lib.mpqc_ccsd_update(self.ctx))

    def compute_ccsd(self, target_precision: float = 1e-9, verbose: bool = False):
        """CCSD::compute_ccsd (ccsd.h:268-477).  Returns E_corr."""
        e = ctypes.c_double()
        n = ctypes.c_int()
        self._check(self.lib.mpqc_ccsd_solve(self.ctx, self.max_iter, float(target_precision),
                                             int(verbose), ctypes.byref(e), ctypes.byref(n)))
        self.n_iter = n.value
        return e.value

    # -- adjoint (no counterpart in the reference)
    def jacobian_transpose(self, y1, y2):
        """(dR/dt)^T y at the current amplitudes; returns (x1, x2)."""
        a = np.ascontiguousarray(y1, dtype=np.float64)
        b = np.ascontiguousarray(y2, dtype=np.float64)
        o, v = self.o, self.v
        x1 = np.empty((v, o))
        x2 = np.empty((v, v, o, o))
        self._check(self.lib.mpqc_ccsd_jacobian_transpose(
            self.ctx, _dptr(a), _dptr(b), _dptr(x1), _dptr(x2), L.LAYOUT_DENSE, L.MEM_HOST))
        return x1, x2

    def lambda_solve(self, target_precision: float = 1e-10, max_iter: int = 100,
                     verbose: bool = False) -> int:
        """Solves (dR/dt)^T lambda = -dE/dt at the converged amplitudes."""
        n = ctypes.c_int()
        self._check(self.lib.mpqc_ccsd_lambda_solve(self.ctx, int(max_iter), float(target_precision),
                                                    int(verbose), ctypes.byref(n)))
        return n.value

    def get_lambda(self):
        return self._get2(self.lib.mpqc_ccsd_get_lambda, "dense")

    def energy_gradient(self):
        """dE/d(integral block) for all nine blocks; returns {name: array}."""
        self._check(self.lib.mpqc_ccsd_energy_gradient(self.ctx))
        out = {}
        o, v = self.o, self.v
        shapes = {"Fia": (o, v), "Fij": (o, o), "Fab": (v, v), "Gijab": (o, o, v, v),
                  "Gijkl": (o, o, o, o), "Giajb": (o, v, o, v), "Giabc": (o, v, v, v),
                  "Gijka": (o, o, o, v), "Gabcd": (v, v, v, v)}
        for name, idx in L.INTEGRAL_IDS.items():
            g = np.empty(shapes[name])
            self._check(self.lib.mpqc_ccsd_get_integral_gradient(self.ctx, idx, _dptr(g),
                                                                 L.LAYOUT_DENSE, L.MEM_HOST))
            out[name] = g
        return out

    def trace(self) -> List[TraceRow]:
        n = ctypes.c_int()
        self._check(self.lib.mpqc_ccsd_get_trace(self.ctx, 0, ctypes.byref(n), None, None, None,
                                                 None, None))
        m = n.value
        it = (ctypes.c_int * max(m, 1))()
        arrs = [np.zeros(max(m, 1)) for _ in range(4)]
        self._check(self.lib.mpqc_ccsd_get_trace(self.ctx, m, ctypes.byref(n), it,
                                                 *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
        return [TraceRow(it[k], arrs[0][k], arrs[1][k], arrs[2][k], arrs[3][k]) for k in range(m)]


def _kernel_times(self):
    n = ctypes.c_int()
    self._check(self.lib.mpqc_ccsd_get_kernel_times(self.ctx, 0, ctypes.byref(n), None, None,
                                                    None, None))
    m = max(n.value, 1)
    names = ctypes.create_string_buffer(64 * m)
    calls = np.zeros(m, dtype=np.int64)
    ms = np.zeros(m)
    fl = np.zeros(m)
    self._check(self.lib.mpqc_ccsd_get_kernel_times(
        self.ctx, m, ctypes.byref(n), names, calls.ctypes.data_as(ctypes.c_void_p),
        ms.ctypes.data_as(ctypes.c_void_p), fl.ctypes.data_as(ctypes.c_void_p)))
    out = {}
    for k in range(n.value):
        tag = names.raw[64 * k:64 * (k + 1)].split(b"\0", 1)[0].decode()
        out[tag] = {"calls": int(calls[k]), "ms": float(ms[k]), "flops": float(fl[k])}
    return out


CCSD.kernel_times = _kernel_times


def pp_ladder(gabcd_ptr: int, tau_ptr: int, r2_ptr: int, o: int, v: int, alpha=1.0, beta=0.0,
              precision: str = "f64", stream: Optional[int] = None):
    """b("a,b,i,j") = alpha * tau("c,d,i,j") * Gabcd("a,b,c,d") + beta * b, on
    device pointers (ccsd_r1_r2.h:275)."""
    lib = L.load()
    L.check(lib.mpqc_ccsd_pp_ladder(o, v, ctypes.c_void_p(gabcd_ptr), ctypes.c_void_p(tau_ptr),
                                    ctypes.c_void_p(r2_ptr), alpha, beta, PRECISIONS[precision],
                                    ctypes.c_void_p(stream or 0)))


def gemm(transa, transb, m, n, k, alpha, a_ptr, lda, b_ptr, ldb, beta, c_ptr, ldc,
         stride_a=0, stride_b=0, stride_c=0, batch=1, precision="f64", stream=None):
    lib = L.load()
    L.check(lib.mpqc_ccsd_gemm(int(transa), int(transb), m, n, k, alpha, ctypes.c_void_p(a_ptr),
                               lda, stride_a, ctypes.c_void_p(b_ptr), ldb, stride_b, beta,
                               ctypes.c_void_p(c_ptr), ldc, stride_c, batch, PRECISIONS[precision],
                               ctypes.c_void_p(stream or 0)))
