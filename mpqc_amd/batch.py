"""Batched CCSD residuals of many independent molecules (``mpqc_ccsd_batch_*``).

Each molecule is evaluated with the reference's operator API
(``cc::compute_cs_ccsd_r1`` + ``compute_cs_ccsd_r2``, ccsd_r1_r2.h:70-293);
the batch runs all molecules' residuals as grouped launches captured in one
hipGraph.  Tile edges and split-K targets of the grouped contractions come from
the library's autotuner (``tune``).
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Sequence

import numpy as np

from . import _lib as L
from .ccsd import PRECISIONS


def _vp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class CCSDBatch:
    def __init__(self, o: Sequence[int], v: Sequence[int], precision: str = "f64"):
        if len(o) != len(v) or len(o) == 0:
            raise ValueError("o and v must be non-empty and of equal length")
        self.lib = L.load()
        self.o = [int(x) for x in o]
        self.v = [int(x) for x in v]
        oa = (ctypes.c_int * len(o))(*self.o)
        va = (ctypes.c_int * len(v))(*self.v)
        h = ctypes.c_void_p()
        st = self.lib.mpqc_ccsd_batch_create(len(o), oa, va, PRECISIONS[precision], ctypes.byref(h))
        if st != L.OK:
            raise L.CCSDError(st, self.lib.mpqc_ccsd_batch_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpqc_ccsd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != L.OK:
            raise L.CCSDError(st, self.lib.mpqc_ccsd_batch_last_error(self.h).decode(errors="replace"))

    def __len__(self):
        return len(self.o)

    def set_integrals(self, mol: int, blocks: Dict[str, np.ndarray]):
        for name, idx in L.INTEGRAL_IDS.items():
            a = np.ascontiguousarray(blocks[name], dtype=np.float64)
            self._check(self.lib.mpqc_ccsd_batch_set_integral(self.h, mol, idx, _vp(a), L.LAYOUT_DENSE,
                                                              L.MEM_HOST))

    def set_amplitudes(self, mol: int, t1, t2):
        a = np.ascontiguousarray(t1, dtype=np.float64)
        b = np.ascontiguousarray(t2, dtype=np.float64)
        o, v = self.o[mol], self.v[mol]
        if a.shape != (v, o) or b.shape != (v, v, o, o):
            raise ValueError("amplitude shapes do not match the molecule")
        self._check(self.lib.mpqc_ccsd_batch_set_amplitudes(self.h, mol, _vp(a), _vp(b), L.LAYOUT_DENSE,
                                                            L.MEM_HOST))

    def residual(self):
        """One residual per molecule (grouped launches, graph replay)."""
        self._check(self.lib.mpqc_ccsd_batch_residual(self.h))

    def get_residuals(self, mol: int):
        o, v = self.o[mol], self.v[mol]
        r1 = np.empty((v, o))
        r2 = np.empty((v, v, o, o))
        self._check(self.lib.mpqc_ccsd_batch_get_residuals(self.h, mol, _vp(r1), _vp(r2), L.LAYOUT_DENSE,
                                                           L.MEM_HOST))
        return r1, r2

    def tune(self, reps: int = 3, verbose: bool = False) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.mpqc_ccsd_batch_tune(self.h, int(reps), int(verbose), ctypes.byref(ms)))
        return ms.value

    def time(self, reps: int = 5) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.mpqc_ccsd_batch_time(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    def time_stages(self, reps: int = 3):
        self._check(self.lib.mpqc_ccsd_batch_time_stages(self.h, int(reps)))

    def stages(self) -> List[dict]:
        n = ctypes.c_int()
        self._check(self.lib.mpqc_ccsd_batch_get_stages(self.h, 0, ctypes.byref(n), None, None, None, None,
                                                        None, None))
        m = n.value
        tags = ctypes.create_string_buffer(64 * max(m, 1))
        kinds = np.zeros(max(m, 1), dtype=np.int32)
        bm = np.zeros(max(m, 1), dtype=np.int32)
        target = np.zeros(max(m, 1), dtype=np.int64)
        ms = np.zeros(max(m, 1))
        flops = np.zeros(max(m, 1))
        self._check(self.lib.mpqc_ccsd_batch_get_stages(self.h, m, ctypes.byref(n), tags, _vp(kinds),
                                                        _vp(bm), _vp(target), _vp(ms), _vp(flops)))
        kind_names = {0: "copy", 1: "permute", 2: "contraction", 3: "tau", 4: "TP"}
        out = []
        for k in range(m):
            tag = tags.raw[64 * k:64 * (k + 1)].split(b"\0", 1)[0].decode()
            out.append({"tag": tag, "kind": kind_names.get(int(kinds[k]), str(kinds[k])),
                        "bm": int(bm[k]), "target": int(target[k]), "ms": float(ms[k]),
                        "flops": float(flops[k])})
        return out
