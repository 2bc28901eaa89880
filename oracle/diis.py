"""CPU oracle: Pulay DIIS with the parameter semantics of TiledArray's ``TA::DIIS``.

TEST INFRASTRUCTURE ONLY (see ``oracle/ccsd.py``).

``TA::DIIS`` lives in TiledArray (third-party, not vendored in the MPQC tree;
MPQC pins TiledArray master >= 6d8d538, cmake/modules/ValidateTiledArray.cmake:30).
MPQC uses it from src/mpqc/chemistry/qc/cc/solvers.h:93-95,112-121,178 with
``TPack`` (src/mpqc/chemistry/qc/cc/tpack.h:20-83) as the vector type.  The
algorithm restated here is the published Pulay DIIS (P. Pulay, Chem. Phys.
Lett. 73, 393 (1980); J. Comput. Chem. 3, 556 (1982)) in TiledArray's form:

* each ``extrapolate(x, e)`` call increments ``iter``; the error ``e`` is
  appended to a history of at most ``ndiis`` vectors (oldest dropped) and the
  overlap matrix ``B(i,j) = <e_i|e_j>`` is updated;
* ``x`` is appended to a parallel history of at most ``ndiis`` vectors;
* extrapolation happens when ``iter > start`` and
  ``(iter - start) % ngroup < ngroupdiis``: solve the bordered system
  ``[[0, -1^T], [-1, B * s]] c = [-1, 0...]`` with ``s = 1/|B(nskip,nskip)|``
  and the diagonal scaled by ``1 + damping``, by column-pivoted Householder QR;
  while ``|det| < zero_determinant`` drop the oldest vector (``nskip += 1``);
  then ``x = sum_k c_k x_k``.

Parity of the per-iteration trajectory with the reference is pinned by the
reference's recorded iteration tables (tests/golden), which fix which
``zero_determinant`` threshold reproduces them.
"""
from __future__ import annotations

from collections import deque

import numpy as np
import scipy.linalg


def _dot(a, b):
    """TPack dot product (tpack.h:52-62): sum over members."""
    return float(sum(np.vdot(x, y) for x, y in zip(a, b)))


def colpiv_qr_solve(A, rhs):
    """Column-pivoted Householder QR solve (the factorisation TA::DIIS uses,
    Eigen::ColPivHouseholderQR): at step k pick the remaining column of
    largest norm, reflect rows k.. with v = x - alpha e_k,
    alpha = -sign(x_k) |x|; rank = #{k : |R_kk| > eps * n * max|R_ii|};
    components beyond the rank are zero.  Returns (x, |det A|)."""
    A = np.array(A, dtype=np.float64)
    b = np.array(rhs, dtype=np.float64)
    n = A.shape[0]
    perm = np.arange(n)
    diag = np.zeros(n)
    for k in range(n):
        norms = np.sum(A[k:, k:] ** 2, axis=0)
        p = k + int(np.argmax(norms))
        if p != k:
            A[:, [k, p]] = A[:, [p, k]]
            perm[[k, p]] = perm[[p, k]]
        norm = np.sqrt(np.sum(A[k:, k] ** 2))
        if norm == 0.0:
            continue
        alpha = -norm if A[k, k] > 0 else norm
        vh = A[k:, k].copy()
        vh[0] -= alpha
        vn = float(vh @ vh)
        if vn > 0.0:
            A[k:, k:] -= np.outer(vh, (2.0 / vn) * (vh @ A[k:, k:]))
            b[k:] -= (2.0 / vn) * (vh @ b[k:]) * vh
        diag[k] = A[k, k]
    absdet = float(np.prod(np.abs(diag)))
    maxpiv = float(np.max(np.abs(diag))) if n else 0.0
    rank = int(np.sum(np.abs(diag) > maxpiv * np.finfo(np.float64).eps * n))
    z = np.zeros(n)
    if rank:
        z[:rank] = scipy.linalg.solve_triangular(np.triu(A[:rank, :rank]), b[:rank])
    x = np.zeros(n)
    x[perm] = z
    return x, absdet


class DIIS:
    def __init__(self, start=1, ndiis=8, damping=0.0, ngroup=1, ngroupdiis=1,
                 zero_determinant=1.0e-15, zero_norm=1.0e-10):
        self.start = int(start)
        self.ndiis = int(ndiis)
        self.damping = float(damping)
        self.ngroup = int(ngroup)
        self.ngroupdiis = int(ngroupdiis)
        self.zero_determinant = float(zero_determinant)
        self.zero_norm = float(zero_norm)
        self.iter = 0
        self.errors = deque()
        self.xs = deque()
        self.B = np.zeros((0, 0))
        self.C = None
        self.nskip = 0
        self.last_coefficients = None

    def _do_extrapolate(self):
        return self.iter > self.start and ((self.iter - self.start) % self.ngroup) < self.ngroupdiis

    def compute_extrapolation_parameters(self, error):
        scale = 1.0 + self.damping
        if len(self.errors) == self.ndiis:
            self.errors.popleft()
            self.B = self.B[1:, 1:]
        self.errors.append([np.array(e, copy=True) for e in error])
        nvec = len(self.errors)
        B = np.zeros((nvec, nvec))
        B[: nvec - 1, : nvec - 1] = self.B
        for i in range(nvec - 1):
            B[i, nvec - 1] = B[nvec - 1, i] = _dot(self.errors[i], self.errors[nvec - 1])
        B[nvec - 1, nvec - 1] = _dot(self.errors[nvec - 1], self.errors[nvec - 1])
        self.B = B
        if not self._do_extrapolate():
            return
        nskip = 0
        while True:
            rank = nvec - nskip + 1
            A = np.zeros((rank, rank))
            A[0, :] = -1.0
            A[:, 0] = -1.0
            A[0, 0] = 0.0
            rhs = np.zeros(rank)
            rhs[0] = -1.0
            norm = 1.0
            if abs(B[nskip, nskip]) > self.zero_norm:
                norm = 1.0 / abs(B[nskip, nskip])
            A[1:, 1:] = B[nskip:, nskip:] * norm
            A[np.diag_indices(rank)] *= scale
            C, absdet = colpiv_qr_solve(A, rhs)
            nskip += 1
            if not (absdet < self.zero_determinant and nskip < nvec):
                break
        if absdet < self.zero_determinant:
            raise FloatingPointError(
                "DIIS::extrapolate: poorly-conditioned system, |A| = %g" % absdet)
        self.nskip = nskip - 1
        self.C = C

    def extrapolate_x(self, x):
        if len(self.xs) == self.ndiis:
            self.xs.popleft()
        self.xs.append([np.array(m, copy=True) for m in x])
        nvec = len(self.xs)
        if self._do_extrapolate():
            coeffs = self.C
            self.last_coefficients = (self.nskip, np.array(coeffs[1:]))
            for m in range(len(x)):
                acc = np.zeros_like(x[m])
                for k, kk in zip(range(self.nskip, nvec), range(1, nvec - self.nskip + 1)):
                    acc += coeffs[kk] * self.xs[k][m]
                x[m] = acc
        else:
            self.last_coefficients = None

    def extrapolate(self, x, error):
        """In place on the list ``x`` (a TPack)."""
        self.iter += 1
        self.compute_extrapolation_parameters(error)
        self.extrapolate_x(x)
