"""GPU CCSD solve (C-ABI, MPQC TA tile layout) against the reference's golden
iteration tables and the oracle, on real molecules from the host front-end."""
import json
import os

import numpy as np
import pytest

from oracle import ccsd as O
from test_golden import FIX, _inputs, check_trace

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("idx", range(len(FIX["cases"])))
def test_gpu_ccsd_trace_matches_reference(idx):
    from mpqc_amd.ccsd import CCSD, Integrals
    case = FIX["cases"][idx]
    _, inp = _inputs(case)
    tp = case["precision"] * case["precision_scale_for_current_rule"]
    with CCSD(inp.o, inp.v, occ_block_size=4, unocc_block_size=4) as c:
        c.set_integrals(Integrals(**inp.blocks), layout="ta")
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        E = c.compute_ccsd(target_precision=tp)
        rows = c.trace()
    e_tol = 2e-12 if case["basis"] == "6-31G" else 2e-11
    print(f"{case['basis']} fc={case['frozen_core']}: E_gpu={E:.15f} E_ref={case['e_ccsd']:.15f} "
          f"rows {len(rows)}/{len(case['trace'])}")
    check_trace(rows, case, e_tol)
    assert len(rows) == len(case["trace"])
    assert abs(E - case["e_ccsd"]) < e_tol


def test_gpu_h_chain_lockstep_with_oracle():
    """H10 chain, cc-pVDZ, all electrons (o=5, v=45), default 24-wide tiles:
    every iteration's energy and amplitudes agree with the oracle."""
    from mpqc_amd.ccsd import CCSD, Integrals
    from mpqc_amd.chem.molecule import h_chain
    from mpqc_amd.chem.scf import rhf
    from mpqc_amd.chem.mo import ccsd_integrals
    scf = rhf(h_chain(10, 1.8), "cc-pVDZ")
    inp = ccsd_integrals(scf, frozen_core=False)
    ref = []
    E_ref, t1r, t2r, rows_ref = O.compute_ccsd(
        O.Integrals(**inp.blocks), inp.eps_occ, inp.eps_vir, target_precision=1e-10,
        callback=lambda it, t1, t2, r1, r2, E: ref.append((t1.copy(), t2.copy(), E)))
    with CCSD(inp.o, inp.v) as c:
        c.set_integrals(Integrals(**inp.blocks), layout="ta")
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        c.init_guess()
        worst = [0.0, 0.0]
        for it, (t1x, t2x, Ex) in enumerate(ref):
            E, err = c.energy_error()
            t1g, t2g = c.get_amplitudes(layout="ta")
            worst[0] = max(worst[0], abs(E - Ex))
            worst[1] = max(worst[1], np.abs(t1g - t1x).max(), np.abs(t2g - t2x).max())
            c.update()
            c.residual()
    print(f"H10 lockstep: {len(ref)} iterations, max |dE| {worst[0]:.2e}, max |dt| {worst[1]:.2e}")
    assert worst[0] < 1e-10 and worst[1] < 1e-9
