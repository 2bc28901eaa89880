"""(T) on the GPU (mpqc_ccsd_triples; CCSD_T::compute_ccsd_t_straight,
ccsd_t.h:1127-1167) against the oracle (oracle/ccsd_t.py) at arbitrary
amplitudes, and the reference's golden (T) energy of the DF water.

Bar: 1e-12 relative to |E(T)| at FP64 (the sum runs in a different order).
"""
import numpy as np
import pytest

from oracle import ccsd as O
from oracle import ccsd_t as T
from spinorb_ccsd import mpqc_blocks, random_hamiltonian
from test_golden import FIX

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("no,nv,block_bytes", [(1, 4, 0), (2, 7, 0), (3, 6, 0), (5, 13, 0), (6, 20, 1 << 20)])
def test_triples_match_oracle_at_random_amplitudes(no, nv, block_bytes):
    from mpqc_amd.ccsd import CCSD, Integrals
    f, g = random_hamiltonian(no, nv, seed=no * 13 + nv, eri_scale=0.05)
    b = mpqc_blocks(f, g, no)
    rng = np.random.default_rng(no + 7 * nv)
    t1 = rng.normal(0, 0.05, (nv, no))
    t2 = rng.normal(0, 0.05, (nv, nv, no, no))   # no permutational symmetry assumed
    eo, ev = np.diag(f)[:no], np.diag(f)[no:]
    ref = T.ccsd_t_straight(t1, t2, b["Gijab"], b["Giabc"], b["Gijka"], eo, ev)
    print(f"o={no} v={nv} block_bytes={block_bytes}", flush=True)
    with CCSD(no, nv) as c:
        c.set_integrals(Integrals(**b))
        c.set_orbital_energies(eo, ev)
        if block_bytes:
            c.set_option("triples_block_bytes", block_bytes)
        c.set_amplitudes(t1, t2)
        et, sec = c.triples()
    print(f"  GPU {et:.15e} oracle {ref:.15e} ({sec * 1e3:.2f} ms)")
    if no == 1:  # two electrons: no triples (closed form E(T) = 0)
        assert abs(et) < 1e-16 and abs(ref) < 1e-16
    else:
        assert abs(et - ref) < 1e-12 * abs(ref)


def test_triples_of_converged_water_match_oracle():
    from mpqc_amd.ccsd import CCSD, Integrals
    from test_golden import _inputs
    case = FIX["cases"][1]  # cc-pVDZ, all electrons
    _, inp = _inputs(case)
    with CCSD(inp.o, inp.v) as c:
        c.set_integrals(Integrals(**inp.blocks))
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        c.compute_ccsd(target_precision=1e-11)
        t1, t2 = c.get_amplitudes()
        et, _ = c.triples()
    b = inp.blocks
    ref = T.ccsd_t_straight(t1, t2, b["Gijab"], b["Giabc"], b["Gijka"], inp.eps_occ, inp.eps_vir)
    print(f"H2O cc-pVDZ (T): GPU {et:.15e} oracle {ref:.15e}")
    assert abs(et - ref) < 1e-12 * abs(ref)


def test_gpu_df_triples_match_reference():
    """DF-CCSD then (T) with <ia|bc> built on the device from the factors."""
    from mpqc_amd.ccsd import CCSD
    from test_oracle_df import _df_inputs
    case = FIX["df_cases"][0]
    tcase = FIX["triples_cases"][0]
    _, inp = _df_inputs(case)
    with CCSD(inp.o, inp.v, occ_block_size=2, unocc_block_size=2) as c:
        c.set_df_basis(inp.naux)
        c.set_integrals(inp.df_blocks())
        c.set_orbital_energies(inp.base.eps_occ, inp.base.eps_vir)
        E = c.compute_ccsd(target_precision=case["precision"])
        et, _ = c.triples()
        giabc = c.get_integral("Giabc")
    print(f"GPU DF-CCSD {E:.15f}, (T) {et:.15e}; reference (T) {tcase['e_triples']:.15e}")
    assert np.abs(giabc - inp.base.blocks["Giabc"]).max() < 1e-13
    assert abs(et - tcase["e_triples"]) < 2e-12


@pytest.mark.parametrize("prec,tol", [("f16x3", 5e-8), ("bf16x3", 1e-6)])  # measured <= 5.6e-9 / 7.7e-8
def test_triples_split_precision_forms(prec, tol):
    """The W GEMMs on 16-bit planes: E(T) within the form's precision of the
    FP64 oracle (random amplitudes, several i blocks) and of FP64 at converged
    water amplitudes."""
    from mpqc_amd.ccsd import CCSD, Integrals
    from test_golden import _inputs
    no, nv = 6, 20
    f, g = random_hamiltonian(no, nv, seed=77, eri_scale=0.05)
    b = mpqc_blocks(f, g, no)
    rng = np.random.default_rng(78)
    t1 = rng.normal(0, 0.05, (nv, no))
    t2 = rng.normal(0, 0.05, (nv, nv, no, no))
    eo, ev = np.diag(f)[:no], np.diag(f)[no:]
    ref = T.ccsd_t_straight(t1, t2, b["Gijab"], b["Giabc"], b["Gijka"], eo, ev)
    with CCSD(no, nv, precision=prec) as c:
        c.set_integrals(Integrals(**b))
        c.set_orbital_energies(eo, ev)
        c.set_option("triples_block_bytes", 1 << 20)
        c.set_amplitudes(t1, t2)
        et, _ = c.triples()
    print(f"{prec} random: {et:.12e} oracle {ref:.12e} rel {abs(et - ref) / abs(ref):.2e}")
    assert abs(et - ref) < tol * abs(ref)
    case = FIX["cases"][1]
    _, inp = _inputs(case)
    with CCSD(inp.o, inp.v) as c:
        c.set_integrals(Integrals(**inp.blocks))
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        c.compute_ccsd(target_precision=1e-10)
        e64, _ = c.triples()
        c.set_option("precision", {"f16x3": 3, "bf16x3": 2}[prec])
        esp, _ = c.triples()
    print(f"{prec} water: {esp:.12e} f64 {e64:.12e} diff {esp - e64:.2e}")
    assert abs(esp - e64) < tol * abs(e64)
