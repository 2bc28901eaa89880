"""AO -> MO transform into the CCSD integral blocks.

Produces the members of ``cc::Integrals`` (reference:
src/mpqc/chemistry/qc/lcao/cc/ccsd_r1_r2.h:24-56, filled at ccsd.h:281-305)
in physicist notation <pq|rs> = (pr|qs), the active occupied / virtual
orbital energies that JacobiDIISSolver receives (ccsd.h:223-229; diagonal of
the MO Fock matrix, mo_build.h:20-26) and the frozen-core count
(core_electrons/2 when frozen_core, lcao_wfn.h:101-107).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .scf import RHFResult


@dataclasses.dataclass
class CCSDInputs:
    blocks: dict            # Fia, Fij, Fab, Gijab, Gijkl, Giajb, Giabc, Gijka, Gabcd
    eps_occ: np.ndarray     # active occupied
    eps_vir: np.ndarray
    n_frozen: int
    e_ref: float

    @property
    def o(self):
        return len(self.eps_occ)

    @property
    def v(self):
        return len(self.eps_vir)


def _transform(eri, C1, C2, C3, C4):
    """(pq|rs) = sum C1[mu,p] C2[nu,q] C3[la,r] C4[si,s] (mu nu|la si).

    Four quarter transforms, each one GEMM that contracts the leading index
    and appends the new one: (mu nu la si) -> (nu la si p) -> ... -> (p q r s).
    """
    x = eri
    for Cq in (C1, C2, C3, C4):
        n0 = x.shape[0]
        rest = x.shape[1:]
        x = (x.reshape(n0, -1).T @ Cq).reshape(*rest, Cq.shape[1])
    return x


def ccsd_integrals(scf: RHFResult, frozen_core: bool = False, fock_ao=None) -> CCSDInputs:
    """MO integrals for CCSD.  ``fock_ao`` (default: the SCF Fock matrix)
    allows a perturbed one-body operator in the fixed MO basis."""
    C = scf.C
    nocc = scf.ndocc
    nfz = scf.basis.mol.core_electrons // 2 if frozen_core else 0
    Ci = C[:, nfz:nocc]
    Ca = C[:, nocc:]
    F = scf.F if fock_ao is None else fock_ao
    Fmo_ii = Ci.T @ F @ Ci
    Fmo_ia = Ci.T @ F @ Ca
    Fmo_aa = Ca.T @ F @ Ca
    # full MO-basis (pq|rs) once, then the blocks
    nmo = C.shape[1]
    mo = _transform(scf.eri, C, C, C, C)
    occ = slice(nfz, nocc)
    vir = slice(nocc, nmo)

    def phys(A, B, Cc, D):
        # <pq|rs> = (pr|qs)
        return np.ascontiguousarray(mo[A, Cc, B, D].transpose(0, 2, 1, 3))

    Ci, Ca = occ, vir

    blocks = dict(
        Fia=np.ascontiguousarray(Fmo_ia), Fij=np.ascontiguousarray(Fmo_ii),
        Fab=np.ascontiguousarray(Fmo_aa),
        Gijab=phys(Ci, Ci, Ca, Ca), Gijkl=phys(Ci, Ci, Ci, Ci), Giajb=phys(Ci, Ca, Ci, Ca),
        Giabc=phys(Ci, Ca, Ca, Ca), Gijka=phys(Ci, Ci, Ci, Ca), Gabcd=phys(Ca, Ca, Ca, Ca))
    return CCSDInputs(blocks, np.diag(Fmo_ii).copy(), np.diag(Fmo_aa).copy(), nfz, scf.energy)
