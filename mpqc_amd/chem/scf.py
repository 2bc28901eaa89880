"""Closed-shell Hartree-Fock (host, numpy).

Symmetric orthogonalisation (as the reference RHF prints: "Using Symmetric
Orthogonalization"), core-Hamiltonian guess, Pulay DIIS on the commutator
FDS - SDF, converged until the energy change and the commutator norm are tight
enough for CCSD energies at the 1e-11 level.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .basis import BasisSet
from .molecule import Molecule


@dataclasses.dataclass
class RHFResult:
    energy: float
    e_nuc: float
    C: np.ndarray          # MO coefficients (nbf x nmo)
    eps: np.ndarray        # orbital energies
    F: np.ndarray          # AO Fock matrix
    H: np.ndarray          # core Hamiltonian
    S: np.ndarray
    eri: np.ndarray        # AO (ij|kl)
    ndocc: int
    iterations: int
    basis: BasisSet


def _jk(eri, D):
    from .basis import jk
    return jk(eri, D)


def rhf(mol: Molecule, basis_name: str, conv_e=1e-13, conv_g=1e-11, max_iter=100,
        eri=None, extra_h=None, verbose=False) -> RHFResult:
    """Closed-shell RHF.  ``extra_h`` (AO matrix) is added to the core
    Hamiltonian (e.g. a uniform electric field)."""
    bs = BasisSet(basis_name, mol)
    S, T, V = bs.one_electron()
    H = T + V
    if extra_h is not None:
        H = H + extra_h
    if eri is None:
        eri = bs.eri()
    nel = mol.n_electrons
    if nel % 2:
        raise ValueError("closed-shell RHF needs an even electron count")
    nocc = nel // 2
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    e_nuc = mol.nuclear_repulsion()

    def diag(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return e, X @ Cp

    eps, C = diag(H)
    D = C[:, :nocc] @ C[:, :nocc].T
    E_old = 0.0
    fs, es = [], []
    for it in range(1, max_iter + 1):
        J, K = _jk(eri, D)
        F = H + 2.0 * J - K
        E = float(np.sum(D * (H + F))) + e_nuc
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        gnorm = float(np.abs(err).max())
        if verbose:
            print(f"RHF iter {it}: E={E:.14f} dE={E - E_old:.2e} |[F,D]|={gnorm:.2e}")
        if abs(E - E_old) < conv_e and gnorm < conv_g:
            break
        E_old = E
        fs.append(F)
        es.append(err)
        if len(fs) > 8:
            fs.pop(0)
            es.pop(0)
        if len(fs) >= 2:
            m = len(fs)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    B[a, b] = float(np.sum(es[a] * es[b]))
            # scale the error overlaps to O(1) so the bordered system stays
            # well conditioned as the errors shrink
            B[:m, :m] /= np.max(np.diag(B[:m, :m]))
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            c = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
            F = sum(ci * fi for ci, fi in zip(c, fs))
        eps, C = diag(F)
        D = C[:, :nocc] @ C[:, :nocc].T
    else:
        raise RuntimeError("RHF did not converge")
    # final canonical orbitals from the converged density's Fock matrix
    J, K = _jk(eri, D)
    F = H + 2.0 * J - K
    eps, C = diag(F)
    D = C[:, :nocc] @ C[:, :nocc].T
    E = float(np.sum(D * (H + F))) + e_nuc
    return RHFResult(E, e_nuc, C, eps, F, H, S, eri, nocc, it, bs)
