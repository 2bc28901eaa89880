"""Adjoint (Lambda) kernels: the transposed-Jacobian product of the residual
against central differences of the oracle residual, and the energy gradients
(dE/dF for a uniform field, dE/d(integral block)) against central differences
of the converged FP64 GPU energy.  The reference has no Lambda equations, so
finite differences are the pin."""
import numpy as np
import pytest
import torch

from test_golden import FIX, _inputs

pytestmark = pytest.mark.gpu


def _random_case(no, nv, seed):
    from spinorb_ccsd import mpqc_blocks, random_hamiltonian
    f, g = random_hamiltonian(no, nv, seed=seed, eri_scale=0.03)
    return f, mpqc_blocks(f, g, no)


@pytest.mark.parametrize("symmetric", [True, False])
def test_jacobian_transpose_matches_oracle_differences(symmetric):
    from oracle import ccsd as O
    from mpqc_amd.ccsd import CCSD, Integrals
    no, nv = 4, 13
    f, b = _random_case(no, nv, seed=31)
    rng = np.random.default_rng(7)
    t1 = rng.normal(0, 0.02, (nv, no))
    t2 = rng.normal(0, 0.02, (nv, nv, no, no))
    if symmetric:
        t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))

    def R(a1, a2):
        tau = O.make_tau(a1, a2)
        ints = O.Integrals(**b)  # r1 hands FIJ / FAB to r2 through ints
        r1 = O.cs_ccsd_r1(a1, a2, tau, ints)
        return r1, O.cs_ccsd_r2(a1, a2, tau, ints)

    y1 = rng.normal(size=t1.shape)
    y2 = rng.normal(size=t2.shape)
    with CCSD(no, nv, precision="f64") as c:
        c.set_integrals(Integrals(**b))
        c.set_orbital_energies(np.diag(f)[:no], np.diag(f)[no:])
        c.set_amplitudes(t1, t2)
        x1, x2 = c.jacobian_transpose(y1, y2)
    h = 1e-4
    for k in range(3):
        d1 = rng.normal(size=t1.shape)
        d2 = rng.normal(size=t2.shape)
        p1, p2 = R(t1 + h * d1, t2 + h * d2)
        m1, m2 = R(t1 - h * d1, t2 - h * d2)
        fd = (np.vdot(y1, p1 - m1) + np.vdot(y2, p2 - m2)) / (2 * h)
        an = np.vdot(d1, x1) + np.vdot(d2, x2)
        print(f"direction {k}: adjoint {an:.12e} central difference {fd:.12e}")
        assert abs(an - fd) <= 1e-7 * max(abs(fd), 1.0)


def _central_difference(e_of, x0, h):
    """Four-point central difference (truncation O(h^4))."""
    return (8 * (e_of(x0 + h) - e_of(x0 - h)) - (e_of(x0 + 2 * h) - e_of(x0 - 2 * h))) / (12 * h)


# The solver's error measure is |r| / (size1 + size2), so 1e-14 leaves |r| ~ 1e-10 and
# energies accurate well below the difference quotients' resolution.
PREC = 1e-14


def _energy(blocks, eo, ev, prec=PREC):
    from mpqc_amd.ccsd import CCSD, Integrals
    o, v = blocks["Fij"].shape[0], blocks["Fab"].shape[0]
    with CCSD(o, v, precision="f64", max_iter=200) as c:
        c.set_integrals(Integrals(**blocks))
        c.set_orbital_energies(eo, ev)
        return c.compute_ccsd(target_precision=prec)


@pytest.mark.parametrize("idx", [0, 1])
def test_field_gradient_matches_central_differences(idx):
    from mpqc_amd.adjoint import ccsd_energy, mo_dipole_blocks
    case = FIX["cases"][idx]
    scf, inp = _inputs(case)
    mu = mo_dipole_blocks(scf, axis=2, frozen_core=case["frozen_core"])
    F = torch.tensor(1e-4, dtype=torch.float64, requires_grad=True)
    blocks = {n: torch.from_numpy(np.array(a)) for n, a in inp.blocks.items()}
    for n in ("Fia", "Fij", "Fab"):
        blocks[n] = blocks[n] + F * torch.from_numpy(mu[n])
    E = ccsd_energy(blocks, inp.eps_occ, inp.eps_vir, target_precision=PREC, lambda_precision=1e-14,
                    max_iter=200)
    E.backward()
    dEdF = F.grad.item()
    h = 5e-4

    def e_at(Fv):
        b = dict(inp.blocks)
        for n in ("Fia", "Fij", "Fab"):
            b[n] = inp.blocks[n] + Fv * mu[n]
        return _energy(b, inp.eps_occ, inp.eps_vir)

    fd = _central_difference(e_at, 1e-4, h)
    print(f"{case['basis']} fc={case['frozen_core']}: dE/dF adjoint {dEdF:.10e} "
          f"central difference {fd:.10e} rel {abs(dEdF - fd) / abs(fd):.2e}")
    assert abs(dEdF - fd) <= 1e-6 * abs(fd)


def _symmetric_mo_perturbation(o, v, rng):
    """Blocks of a random perturbation of the MO two-electron integrals with the
    8-fold permutational symmetry of real orbitals, (pq|rs) in chemists'
    notation, cut into the reference's physicist blocks <pq|rs> = (pr|qs), plus
    symmetric Fock blocks."""
    n = o + v
    X = rng.normal(size=(n, n, n, n))
    X = X + X.transpose(1, 0, 2, 3)
    X = X + X.transpose(0, 1, 3, 2)
    X = X + X.transpose(2, 3, 0, 1)
    occ, vir = slice(0, o), slice(o, n)

    def phys(A, B, C, D):
        return np.ascontiguousarray(X[A, C, B, D].transpose(0, 2, 1, 3))

    F = rng.normal(size=(n, n))
    F = F + F.T
    return {"Gijab": phys(occ, occ, vir, vir), "Gijkl": phys(occ, occ, occ, occ),
            "Giajb": phys(occ, vir, occ, vir), "Giabc": phys(occ, vir, vir, vir),
            "Gijka": phys(occ, occ, occ, vir), "Gabcd": phys(vir, vir, vir, vir),
            "Fia": F[occ, vir], "Fij": F[occ, occ], "Fab": F[vir, vir]}


def _gradient(inp, symmetric=True):
    from mpqc_amd.ccsd import CCSD, Integrals
    with CCSD(inp.o, inp.v, precision="f64", max_iter=200) as c:
        c.set_option("lambda_symmetric", 1 if symmetric else 0)
        c.set_integrals(Integrals(**inp.blocks))
        c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
        c.compute_ccsd(target_precision=PREC)
        n = c.lambda_solve(target_precision=1e-14, max_iter=400)
        return c.energy_gradient(), n


def test_integral_gradients_match_symmetric_perturbations():
    """All nine blocks at once along a perturbation of the MO integrals that keeps
    their permutational symmetry (the physical case: fields, orbital changes)."""
    case = FIX["cases"][0]
    _, inp = _inputs(case)
    grad, n = _gradient(inp)
    rng = np.random.default_rng(5)
    dX = _symmetric_mo_perturbation(inp.o, inp.v, rng)
    norm = np.sqrt(sum(np.vdot(x, x) for x in dX.values()))
    dX = {k: x / norm for k, x in dX.items()}

    def e_at(s):
        return _energy({k: inp.blocks[k] + s * dX[k] for k in inp.blocks}, inp.eps_occ, inp.eps_vir)

    fd = _central_difference(e_at, 0.0, 5e-4)
    an = sum(float(np.vdot(grad[k], dX[k])) for k in dX)
    print(f"symmetric perturbation ({n} Lambda iterations): adjoint {an:.10e} central difference {fd:.10e}")
    for k in dX:
        print(f"  {k}: <grad, dX> = {float(np.vdot(grad[k], dX[k])):.6e}")
    assert abs(an - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("name", ["Gabcd", "Gijkl", "Gijab", "Giabc"])
def test_full_space_gradients_match_arbitrary_perturbations(name):
    """With the full-space Lambda solve the block gradients hold along arbitrary
    (symmetry-breaking) directions too."""
    case = FIX["cases"][0]
    _, inp = _inputs(case)
    grad, n = _gradient(inp, symmetric=False)
    rng = np.random.default_rng(6)
    g = grad[name]
    r = rng.normal(size=g.shape)
    X = (g / np.linalg.norm(g) + r / np.linalg.norm(r)) / np.sqrt(2.0)

    def e_at(s):
        b = dict(inp.blocks)
        b[name] = inp.blocks[name] + s * X
        return _energy(b, inp.eps_occ, inp.eps_vir)

    fd = _central_difference(e_at, 0.0, 5e-4)
    an = float(np.vdot(g, X))
    print(f"{name} ({n} Lambda iterations): adjoint {an:.10e} central difference {fd:.10e}")
    assert abs(an - fd) <= 1e-6 * abs(fd)
