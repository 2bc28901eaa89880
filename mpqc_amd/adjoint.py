"""Derivatives of the converged CCSD correlation energy through the adjoint
(Lambda) kernels of the HIP library, as a ``torch.autograd.Function``.

This is a verification binding: the forward pass runs the GPU solve
(``CCSD::compute_ccsd``, ccsd.h:268-477), the backward pass solves the Lambda
equations (dR/dt)^T lambda = -dE/dt at the converged amplitudes and returns
dE/d(block) for the nine integral blocks of ``cc::Integrals``
(ccsd_r1_r2.h:24-56).  Anything that builds the blocks from torch tensors
(e.g. a field-dependent Fock matrix f + F mu) gets dE/dF by the chain rule.

The reference has no Lambda equations or CCSD gradients; the convention is the
one of the converged energy of ccsd.h:376-380 at R = 0,
E = 2 <f_ai, t1> + <g_abij, 2 tau(a,b,i,j) - tau(b,a,i,j)>, with the blocks
treated as independent inputs (no orbital relaxation, fixed MO basis).
"""
from __future__ import annotations

from typing import Dict

import numpy as np
import torch

from . import _lib as L
from .ccsd import CCSD, PRECISIONS, Integrals

NAMES = list(L.INTEGRAL_IDS)

# iterations of the last forward solve and Lambda solve (diagnostics)
last_info = {}


class CCSDEnergy(torch.autograd.Function):
    """E_corr(Fia, Fij, Fab, Gijab, Gijkl, Giajb, Giabc, Gijka, Gabcd)."""

    @staticmethod
    def forward(ctx, eps_occ, eps_vir, opts, *blocks):
        o, v = blocks[1].shape[0], blocks[2].shape[0]
        c = CCSD(o, v, precision=opts.get("precision", "f64"), max_iter=opts.get("max_iter", 100))
        try:
            c.set_integrals(Integrals(**{n: b.detach().cpu().numpy() for n, b in zip(NAMES, blocks)}))
            c.set_orbital_energies(np.asarray(eps_occ, dtype=np.float64),
                                   np.asarray(eps_vir, dtype=np.float64))
            target = opts.get("target_precision", 1e-11)
            refine = opts.get("refine")
            last_info.pop("refine_iterations", None)
            if refine and opts.get("precision", "f64") != "f64":
                # iterative refinement: the reduced-precision form converges to `refine`,
                # then FP64 iterations continue from its amplitudes to `target`
                c.compute_ccsd(target_precision=refine)
                last_info["ccsd_iterations"] = c.n_iter
                c.set_option("precision", PRECISIONS["f64"])
                c.set_option("solve_restart", 2)   # project t2 onto its exact symmetry first
                e = c.compute_ccsd(target_precision=target)
                last_info["refine_iterations"] = c.n_iter
            else:
                e = c.compute_ccsd(target_precision=target)
                last_info["ccsd_iterations"] = c.n_iter
        except Exception:
            c.close()
            raise
        ctx.c = c
        ctx.opts = opts
        ctx.devices = [b.device for b in blocks]
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        c = ctx.c
        try:
            # the error measure divides |r| by the element count, so the default is tight
            last_info["lambda_iterations"] = c.lambda_solve(
                target_precision=ctx.opts.get("lambda_precision", 1e-15),
                max_iter=ctx.opts.get("lambda_max_iter", 200))
            g = c.energy_gradient()
        finally:
            c.close()
        grads = [grad_e * torch.from_numpy(g[n]).to(dev) for n, dev in zip(NAMES, ctx.devices)]
        return (None, None, None, *grads)


def ccsd_energy(blocks: Dict[str, torch.Tensor], eps_occ, eps_vir, **opts) -> torch.Tensor:
    """Converged CCSD correlation energy as a differentiable function of the
    integral blocks (float64 tensors, reference index orders)."""
    return CCSDEnergy.apply(eps_occ, eps_vir, opts, *[blocks[n] for n in NAMES])


def mo_dipole_blocks(scf, axis: int = 2, frozen_core: bool = False) -> Dict[str, np.ndarray]:
    """Blocks of the electric dipole operator mu_axis = -r_axis in the SCF MO
    basis (occupied active / virtual), shaped like Fia, Fij, Fab.  A uniform
    field F along `axis` adds F * mu to the Fock operator (fixed orbitals)."""
    _, _, _, D = scf.basis.one_electron(dipole=True)
    mu = -(scf.C.T @ D[axis] @ scf.C)
    nocc = scf.ndocc
    nfz = scf.basis.mol.core_electrons // 2 if frozen_core else 0
    return {"Fia": mu[nfz:nocc, nocc:], "Fij": mu[nfz:nocc, nfz:nocc], "Fab": mu[nocc:, nocc:]}
