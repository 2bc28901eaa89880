"""The adjoint configuration at its full size: (H2O)6 on a 3x2x1 lattice (seed 5),
cc-pVDZ, all electrons (o=30, v=114); dE_CCSD/dF for a uniform field along z
at F = 1e-4 a.u. (f + F mu in the fixed MO basis) from the Lambda equations,
against a four-point central difference of the converged FP64 energy.

Forms: FP64 throughout, and the mixed forms: amplitudes converged on the f16x3
plane contractions, or on the bf16x3 ones to 1e-8 and then refined by FP64
iterations from those amplitudes (bf16x3 alone stalls at its residual floor,
about 1e-6 relative in the gradient), with the Lambda equations and the gradient
sweep on the FP64 tape at the final amplitudes.  Bar: 1e-6 relative (the
reference has no Lambda equations; finite differences are the pin).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F0 = 1e-4
H = 5e-4
TOL = 1e-6


@pytest.fixture(scope="module")
def h2o6_field():
    from mpqc_amd.adjoint import mo_dipole_blocks
    from mpqc_amd.ccsd import CCSD, Integrals
    from mpqc_amd.chem.mo import ccsd_integrals
    from mpqc_amd.chem.molecule import water_cluster
    from mpqc_amd.chem.scf import rhf
    print("(H2O)6: RHF/cc-pVDZ, MO integrals and dipole blocks on the host", flush=True)
    scf = rhf(water_cluster((3, 2, 1), seed=5), "cc-pVDZ")
    inp = ccsd_integrals(scf, frozen_core=False)
    assert (inp.o, inp.v) == (30, 114)
    mu = mo_dipole_blocks(scf, axis=2)

    def blocks_at(Fv):
        b = dict(inp.blocks)
        for n in ("Fia", "Fij", "Fab"):
            b[n] = inp.blocks[n] + Fv * mu[n]
        return b

    def e_at(Fv):
        with CCSD(inp.o, inp.v, max_iter=200) as c:
            c.set_integrals(Integrals(**blocks_at(Fv)))
            c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
            return c.compute_ccsd(target_precision=1e-14)

    fd = (8 * (e_at(F0 + H) - e_at(F0 - H)) - (e_at(F0 + 2 * H) - e_at(F0 - 2 * H))) / (12 * H)
    print(f"(H2O)6 dE/dF central difference (FP64 energies) {fd:.12e}", flush=True)
    return inp, mu, fd


@pytest.mark.parametrize("prec,target,refine", [("f64", 1e-14, None), ("f16x3", 1e-11, None),
                                                ("bf16x3", 1e-13, 1e-8)])
def test_h2o6_field_gradient_against_central_difference(h2o6_field, prec, target, refine):
    import torch
    from mpqc_amd.adjoint import ccsd_energy, last_info
    inp, mu, fd = h2o6_field
    F = torch.tensor(F0, dtype=torch.float64, requires_grad=True)
    tb = {n: torch.from_numpy(np.array(a)) for n, a in inp.blocks.items()}
    for n in ("Fia", "Fij", "Fab"):
        tb[n] = tb[n] + F * torch.from_numpy(mu[n])
    E = ccsd_energy(tb, inp.eps_occ, inp.eps_vir, precision=prec, target_precision=target,
                    lambda_precision=1e-16, max_iter=200, refine=refine)
    E.backward()
    g = F.grad.item()
    rel = abs(g - fd) / abs(fd)
    print(f"{prec}: E_corr {E.item():.12f} dE/dF adjoint {g:.12e} difference {fd:.12e} rel {rel:.2e} "
          f"({last_info.get('ccsd_iterations')} CCSD {prec} + {last_info.get('refine_iterations', 0)} FP64 "
          f"refinement, {last_info.get('lambda_iterations')} Lambda iterations)")
    assert rel < TOL
