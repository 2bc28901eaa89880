"""Density fitting for the host front-end: RI-RHF and the DF-CCSD integral set.

The reference builds three-index integrals ``(K|G|p q)[inv_sqr]``
(src/mpqc/chemistry/qc/lcao/cc/ccsd.h:481-493): Coulomb three-centre
integrals over the fitting basis contracted with the inverse square root of the
Coulomb metric, X(K,p,q) = sum_L [V^-1/2]_KL (L|pq), so that the four-index
integrals are (pq|rs)[df] = sum_K X(K,p,q) X(K,r,s).  Every DF quantity the
CCSD residual uses pairs two X factors over K, so the result does not depend on
which factor L (L L^T = V^-1) is chosen; the symmetric one is used here.

RI-RHF (the reference's "RI-RHF" wavefunction) is RHF with the DF four-index
integrals; its Fock matrix and MO integrals are the "[df]" blocks the CCSD
receives (ccsd.h:281-305).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .basis import BasisSet
from .mo import CCSDInputs, ccsd_integrals
from .molecule import Molecule
from .scf import RHFResult, rhf


def inv_sqrt(V: np.ndarray) -> np.ndarray:
    """Symmetric V^-1/2 of the (positive definite) Coulomb metric."""
    w, U = np.linalg.eigh(V)
    if w.min() <= 0:
        raise ValueError("Coulomb metric is not positive definite")
    return (U * w ** -0.5) @ U.T


def df_factors(obs: BasisSet, dfbs: BasisSet) -> np.ndarray:
    """X(K, mu, nu) = sum_L [V^-1/2]_KL (L|mu nu) in the AO basis: (naux, nbf, nbf)."""
    B = obs.eri3(dfbs)
    Vih = inv_sqrt(dfbs.eri2())
    naux, n, _ = B.shape
    return (Vih @ B.reshape(naux, n * n)).reshape(naux, n, n)


@dataclasses.dataclass
class DFResult:
    scf: RHFResult
    X_ao: np.ndarray  # (naux, nbf, nbf)
    dfbs: BasisSet


def ri_rhf(mol: Molecule, basis_name: str, df_basis_name: str, **kw) -> DFResult:
    obs = BasisSet(basis_name, mol)
    dfbs = BasisSet(df_basis_name, mol)
    X = df_factors(obs, dfbs)
    naux, n, _ = X.shape
    Xm = X.reshape(naux, n * n)
    eri = (Xm.T @ Xm).reshape(n, n, n, n)
    return DFResult(rhf(mol, basis_name, eri=eri, **kw), X, dfbs)


@dataclasses.dataclass
class DFCCSDInputs:
    """Inputs of the DF residual (ccsd_r1_r2.h:136-184, 309-520): the
    four-index [df] blocks in ``blocks`` (all nine, Giabc and Gabcd included
    for the standard path) and the three-index factors in MO blocks."""
    base: CCSDInputs
    Xai: np.ndarray  # (naux, v, o)
    Xij: np.ndarray  # (naux, o, o)
    Xab: np.ndarray  # (naux, v, v)

    @property
    def o(self):
        return self.base.o

    @property
    def v(self):
        return self.base.v

    @property
    def naux(self):
        return self.Xai.shape[0]

    def df_blocks(self) -> dict:
        """The integral set the reference's DF method receives with
        reduced_abcd_memory (ccsd.h:292-305): no Giabc, no Gabcd."""
        b = {k: self.base.blocks[k] for k in ("Fia", "Fij", "Fab", "Gijab", "Gijkl", "Giajb",
                                              "Gijka")}
        b.update(Xai=self.Xai, Xij=self.Xij, Xab=self.Xab)
        return b


def df_ccsd_integrals(df: DFResult, frozen_core: bool = False) -> DFCCSDInputs:
    base = ccsd_integrals(df.scf, frozen_core=frozen_core)
    C = df.scf.C
    nocc = df.scf.ndocc
    Ci = C[:, base.n_frozen:nocc]
    Ca = C[:, nocc:]
    naux, n, _ = df.X_ao.shape

    def mo3(C1, C2):
        x = df.X_ao.reshape(naux * n, n) @ C2             # (K mu, q)
        x = x.reshape(naux, n, -1).transpose(0, 2, 1)       # (K, q, mu)
        x = x.reshape(-1, n) @ C1                           # (K q, p)
        return np.ascontiguousarray(x.reshape(naux, C2.shape[1], C1.shape[1]).transpose(0, 2, 1))

    return DFCCSDInputs(base, mo3(Ca, Ci), mo3(Ci, Ci), mo3(Ca, Ca))
