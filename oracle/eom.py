"""CPU oracle for MPQC's EOM-EE-CCSD singlet excitation energies (numpy).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline.  The product path (``mpqc_amd``) never
imports it.

The reference's EOM-CCSD (src/mpqc/chemistry/qc/lcao/cc/eom/eom_ccsd.h:17 with
eom_ccsd_impl.h, H-bar intermediates in ccsd_hbar.h:422-500) finds the lowest
eigenvalues of the similarity-transformed Hamiltonian in the space of singles
and doubles.  At converged amplitudes that matrix, minus E_CC on its diagonal,
is the Jacobian of the CCSD residual, dR_mu/dt_nu = <mu|[H-bar, tau_nu]|0>:
the excitation energies are its eigenvalues.  For closed shells the residual
is equivariant under the simultaneous exchange (a,i) <-> (b,j) of the doubles,
so the Jacobian at symmetric amplitudes commutes with that exchange; the
singlet roots are the eigenvalues of its block on (ab)(ij)-symmetric doubles.

``eom_ee_singlets`` builds that block by central differences of the oracle's
residual (ccsd_r1_r2.h:70-293, or the df residual) along an orthonormal basis of
the symmetric subspace and returns its lowest eigenvalues.

Pinning (tests/test_oracle_eom.py): the reference's converged singlet
excitation energies of h2o-eom-ccsd-direct-631g.out and h2o-eom-ccsd-631g-pvdz.out
(12 printed digits).
"""
from __future__ import annotations

import numpy as np

from . import ccsd as O

__all__ = ["symmetric_basis_apply", "eom_ee_singlets"]


def _sym_pairs(v, o):
    """Index pairs (p, q) of t2 elements (a,b,i,j) and (b,a,j,i), p <= q."""
    idx = np.arange(v * v * o * o).reshape(v, v, o, o)
    sw = idx.transpose(1, 0, 3, 2)
    flat, flat_sw = idx.ravel(), sw.ravel()
    keep = flat <= flat_sw
    return flat[keep], flat_sw[keep]


def symmetric_basis_apply(v, o):
    """(n_sym, embed, project): embed maps coefficients in the orthonormal
    basis of the symmetric subspace to full (t1, t2) vectors; project is its
    transpose."""
    n1 = v * o
    p, q = _sym_pairs(v, o)
    diag = p == q
    w = np.where(diag, 1.0, 1.0 / np.sqrt(2.0))
    n_sym = n1 + p.size

    def embed(c):
        x = np.zeros(n1 + v * v * o * o)
        x[:n1] = c[:n1]
        c2 = c[n1:] * w
        x[n1 + p] += c2
        x[n1 + q[~diag]] += c2[~diag]
        return x

    def project(x):
        x2 = x[n1:]
        c2 = (x2[p] + np.where(diag, 0.0, x2[q])) * w
        return np.concatenate([x[:n1], c2])
    return n_sym, embed, project


def eom_ee_singlets(residual, t1, t2, nroots, h=1e-5):
    """Lowest ``nroots`` singlet excitation energies at converged (t1, t2);
    ``residual(t1, t2) -> (r1, r2)`` is the CCSD residual."""
    v, o = t1.shape
    n_sym, embed, project = symmetric_basis_apply(v, o)
    x0 = np.concatenate([t1.ravel(), t2.ravel()])
    n1 = v * o

    def R(x):
        r1, r2 = residual(x[:n1].reshape(v, o), x[n1:].reshape(v, v, o, o))
        return np.concatenate([r1.ravel(), r2.ravel()])
    J = np.empty((n_sym, n_sym))
    for k in range(n_sym):
        e = np.zeros(n_sym)
        e[k] = h
        d = embed(e)
        J[:, k] = project((R(x0 + d) - R(x0 - d)) / (2.0 * h))
    w = np.linalg.eigvals(J)
    w = np.sort(w.real[np.abs(w.imag) < 1e-9 * np.maximum(1.0, np.abs(w.real))])
    return w[:nroots]


def standard_residual(ints):
    """The standard residual of the oracle (ccsd_r1_r2.h:70-293) as a closure."""
    def res(t1, t2):
        tau = O.make_tau(t1, t2)
        r1 = O.cs_ccsd_r1(t1, t2, tau, ints)
        return r1, O.cs_ccsd_r2(t1, t2, tau, ints)
    return res
