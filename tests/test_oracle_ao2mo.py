"""AO -> MO transform oracle (oracle/ao2mo.py: lcao_factory.h:378-497) pinned by
the Hartree-Fock energy rebuilt from its MO blocks and by the reference's
golden CCSD energies computed from those blocks (standard and density-fitted)."""
import numpy as np
import pytest

from oracle import ao2mo as A
from oracle import ccsd as O
from oracle import ccsd_df as D
from test_golden import FIX


def _mol():
    from mpqc_amd.chem.molecule import Molecule
    return Molecule.from_xyz_text(FIX["geometry_angstrom"])


@pytest.mark.parametrize("idx", [0, 2])
def test_ao_blocks_reproduce_hf_and_golden_ccsd(idx):
    from mpqc_amd.chem.scf import rhf
    case = FIX["cases"][idx]
    mol = _mol()
    scf = rhf(mol, case["basis"])
    nocc = scf.ndocc
    b = A.ccsd_blocks_ao(scf.eri, scf.F, scf.C[:, :nocc], scf.C[:, nocc:])
    assert abs(A.hf_energy_from_blocks(b, scf.e_nuc) - scf.energy) < 1e-10
    nfz = mol.core_electrons // 2 if case["frozen_core"] else 0
    b = A.ccsd_blocks_ao(scf.eri, scf.F, scf.C[:, nfz:nocc], scf.C[:, nocc:])
    tp = case["precision"] * case["precision_scale_for_current_rule"]
    E, *_ = O.compute_ccsd(O.Integrals(**b), np.diag(b["Fij"]), np.diag(b["Fab"]), target_precision=tp)
    assert abs(E - case["e_ccsd"]) < (2e-12 if case["basis"] == "6-31G" else 2e-11)


def test_df_blocks_reproduce_hf_and_golden_df_ccsd():
    from mpqc_amd.chem.basis import BasisSet
    from mpqc_amd.chem.df import inv_sqrt, ri_rhf
    case = FIX["df_cases"][0]
    mol = _mol()
    d = ri_rhf(mol, case["basis"], case["df_basis"])
    obs, dfbs = BasisSet(case["basis"], mol), BasisSet(case["df_basis"], mol)
    B, L = obs.eri3(dfbs), inv_sqrt(dfbs.eri2())
    s = d.scf
    full = A.ccsd_blocks_df(B, L, s.F, s.C[:, :s.ndocc], s.C[:, s.ndocc:])
    assert abs(A.hf_energy_from_blocks(full, s.e_nuc) - s.energy) < 1e-10
    # a non-symmetric metric factor (L Q with Q orthogonal) gives the same blocks
    Qm, _ = np.linalg.qr(np.random.default_rng(0).normal(size=L.shape))
    rot = A.ccsd_blocks_df(B, L @ Qm, s.F, s.C[:, :s.ndocc], s.C[:, s.ndocc:])
    assert np.abs(rot["Gijab"] - full["Gijab"]).max() < 1e-12
    b = A.ccsd_blocks_df(B, L, s.F, s.C[:, 1:s.ndocc], s.C[:, s.ndocc:])
    E, *_ = D.compute_ccsd_df(D.DFIntegrals(**b), np.diag(b["Fij"]), np.diag(b["Fab"]),
                              target_precision=case["precision"])
    assert abs(E - case["e_ccsd"]) < 2e-12
