"""AO -> MO transform on the GPU (mpqc_ccsd_transform_ao / _df, LCAOFactory::compute4,
lcao_factory.h:378-497) against the oracle (oracle/ao2mo.py), and the reference's
golden CCSD energies from integrals that never leave the device.

Bar: every block within 1e-12 of the oracle relative to its largest element;
golden energies as in tests/test_golden.py.
"""
import numpy as np
import pytest

from oracle import ao2mo as A
from test_golden import FIX, check_trace

pytestmark = pytest.mark.gpu


def _mol():
    from mpqc_amd.chem.molecule import Molecule
    return Molecule.from_xyz_text(FIX["geometry_angstrom"])


def _rel(a, b, name=""):
    # Fia of a converged RHF is round-off (~1e-12): Fock blocks are compared on the
    # scale of the Fock matrix (O(1) Eh)
    scale = max(np.abs(b).max(), 1.0 if name.startswith("F") else 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("idx,layout", [(0, "ta"), (1, "dense")])
def test_transform_ao_matches_oracle_and_golden(idx, layout):
    from mpqc_amd.ccsd import CCSD
    from mpqc_amd.chem.scf import rhf
    case = FIX["cases"][idx]
    mol = _mol()
    scf = rhf(mol, case["basis"])
    nfz = mol.core_electrons // 2 if case["frozen_core"] else 0
    Ci, Ca = scf.C[:, nfz:scf.ndocc], scf.C[:, scf.ndocc:]
    ref = A.ccsd_blocks_ao(scf.eri, scf.F, Ci, Ca)
    o, v = Ci.shape[1], Ca.shape[1]
    tp = case["precision"] * case["precision_scale_for_current_rule"]
    print(f"{case['basis']} fc={case['frozen_core']} o={o} v={v}", flush=True)
    with CCSD(o, v, occ_block_size=2, unocc_block_size=3) as c:
        c.transform_ao(scf.eri, scf.F, Ci, Ca)
        for name, x in ref.items():
            e = _rel(c.get_integral(name, layout=layout), x, name)
            print(f"  {name}: {e:.2e}")
            assert e < 1e-12
        c.set_orbital_energies(np.diag(ref["Fij"]), np.diag(ref["Fab"]))
        E = c.compute_ccsd(target_precision=tp)
        rows = c.trace()
    e_tol = 2e-12 if case["basis"] == "6-31G" else 2e-11
    check_trace(rows, case, e_tol)
    assert abs(E - case["e_ccsd"]) < e_tol


def _hip_runtime():
    """The HIP runtime the library is bound to (same soname), for raw device buffers."""
    import ctypes
    from mpqc_amd import _lib as L
    L.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_transform_ao_from_device_buffers():
    import ctypes
    from mpqc_amd import _lib as L
    from mpqc_amd.ccsd import CCSD
    from mpqc_amd.chem.scf import rhf
    scf = rhf(_mol(), "6-31G")
    Ci, Ca = scf.C[:, 1:scf.ndocc], scf.C[:, scf.ndocc:]
    ref = A.ccsd_blocks_ao(scf.eri, scf.F, Ci, Ca)
    hip = _hip_runtime()
    host = [np.ascontiguousarray(a) for a in (scf.eri, scf.F, Ci, Ca)]
    dev = []
    try:
        for a in host:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            dev.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        with CCSD(Ci.shape[1], Ca.shape[1]) as c:
            c._check(c.lib.mpqc_ccsd_transform_ao(c.ctx, scf.F.shape[0], *dev, L.MEM_DEVICE))
            for name, x in ref.items():
                assert _rel(c.get_integral(name), x, name) < 1e-12
    finally:
        for p in dev:
            hip.hipFree(p)


@pytest.mark.parametrize("reduced", [True, False])
def test_transform_df_matches_oracle_and_golden(reduced):
    from mpqc_amd.ccsd import CCSD
    from mpqc_amd.chem.basis import BasisSet
    from mpqc_amd.chem.df import inv_sqrt, ri_rhf
    case = FIX["df_cases"][0]
    mol = _mol()
    d = ri_rhf(mol, case["basis"], case["df_basis"])
    obs, dfbs = BasisSet(case["basis"], mol), BasisSet(case["df_basis"], mol)
    B, Lm = obs.eri3(dfbs), inv_sqrt(dfbs.eri2())
    s = d.scf
    Ci, Ca = s.C[:, 1:s.ndocc], s.C[:, s.ndocc:]
    ref = A.ccsd_blocks_df(B, Lm, s.F, Ci, Ca, reduced_abcd=reduced)
    naux = B.shape[0]
    with CCSD(Ci.shape[1], Ca.shape[1], occ_block_size=2, unocc_block_size=2) as c:
        c.set_df_basis(naux, aux_block_size=5)
        if not reduced:
            c.set_option("reduced_abcd_memory", 0)
        c.transform_df(B, Lm, s.F, Ci, Ca)
        for name, x in ref.items():
            e = _rel(c.get_integral(name, layout="ta"), x, name)
            print(f"  {name}: {e:.2e}")
            assert e < 1e-12
        c.set_orbital_energies(np.diag(ref["Fij"]), np.diag(ref["Fab"]))
        E = c.compute_ccsd(target_precision=case["precision"])
        rows = c.trace()
    print(f"GPU-transformed DF-CCSD {E:.15f} reference {case['e_ccsd']:.15f}")
    check_trace(rows, case, 2e-12)
    assert abs(E - case["e_ccsd"]) < 2e-12
    # already-fitted factors (metric None) give the same blocks
    Xf = np.einsum("LK,Lmn->Kmn", Lm, B)
    with CCSD(Ci.shape[1], Ca.shape[1]) as c:
        c.set_df_basis(naux)
        c.transform_df(Xf, None, s.F, Ci, Ca)
        assert _rel(c.get_integral("Gijab"), ref["Gijab"]) < 1e-12
