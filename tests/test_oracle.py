"""Pins for the CPU oracle (oracle/ccsd.py, oracle/diis.py).

Each check is chosen so that a plausible mistake in a restated term (dropped
term, wrong sign, swapped index, transposed operand) fails it:
* independent spin-orbital CCSD (textbook equations) reaches the same energy
  and the oracle's residual vanishes at the spin-orbital amplitudes;
* two-electron systems: CCSD == FCI;
* MP2 closed form for the first Jacobi step;
* invariance of E_corr under occupied-occupied / virtual-virtual rotations
  (exercises every off-diagonal Fock term);
* size consistency for non-interacting fragments;
* t = 0 residuals equal the integrals (ccsd.h:351-352).
"""
import numpy as np
import pytest

from oracle import ccsd as O
from oracle.diis import DIIS, colpiv_qr_solve
from spinorb_ccsd import (fci_two_electron, mpqc_blocks, random_hamiltonian, spin_to_spatial,
                          spinorb_ccsd)


def _solve(f, g, no, tp=1e-12, **kw):
    ints = O.Integrals(**mpqc_blocks(f, g, no))
    return O.compute_ccsd(ints, np.diag(f)[:no], np.diag(f)[no:], max_iter=200,
                          target_precision=tp, **kw)


@pytest.mark.parametrize("no,nv,seed", [(2, 4, 1), (3, 5, 2), (4, 7, 3)])
def test_matches_spin_orbital_ccsd(no, nv, seed):
    f, g = random_hamiltonian(no, nv, seed=seed)
    E, t1, t2, _ = _solve(f, g, no)
    Es, t1s, t2s, _ = spinorb_ccsd(f, g, no)
    assert abs(E - Es) < 1e-11
    t1x, t2x = spin_to_spatial(t1s, t2s, no, nv)
    ints = O.Integrals(**mpqc_blocks(f, g, no))
    tau = O.make_tau(t1x, t2x)
    r1 = O.cs_ccsd_r1(t1x, t2x, tau, ints)
    r2 = O.cs_ccsd_r2(t1x, t2x, tau, ints)
    assert np.abs(r1).max() < 1e-10 and np.abs(r2).max() < 1e-10
    assert np.abs(t1 - t1x).max() < 1e-9 and np.abs(t2 - t2x).max() < 1e-9


def test_nonzero_fov_reference_sign_documented():
    """With f_ov != 0 the exact form (fov_sign=-1) equals spin-orbital CCSD;
    the reference as written (+1) differs only through that term."""
    no, nv = 3, 5
    f, g = random_hamiltonian(no, nv, seed=3, f_ov=0.03)
    Es, t1s, t2s, _ = spinorb_ccsd(f, g, no)
    E_exact, *_ = _solve(f, g, no, fov_sign=-1.0)
    E_ref, *_ = _solve(f, g, no, fov_sign=1.0)
    assert abs(E_exact - Es) < 1e-11
    assert abs(E_ref - Es) > 1e-7
    # r1 does not involve F_VV: it vanishes at the exact amplitudes either way
    t1x, t2x = spin_to_spatial(t1s, t2s, no, nv)
    ints = O.Integrals(**mpqc_blocks(f, g, no))
    r1 = O.cs_ccsd_r1(t1x, t2x, O.make_tau(t1x, t2x), ints)
    assert np.abs(r1).max() < 1e-10


@pytest.mark.parametrize("nv,f_ov,seed", [(5, 0.0, 4), (8, 0.0, 5), (6, 0.05, 6)])
def test_two_electron_ccsd_equals_fci(nv, f_ov, seed):
    f, g = random_hamiltonian(1, nv, seed=seed, f_ov=f_ov, eri_scale=0.08)
    E, *_ = _solve(f, g, 1, fov_sign=-1.0 if f_ov else 1.0)
    assert abs(E - fci_two_electron(f, g)) < 1e-11


def test_first_step_is_mp2():
    no, nv = 3, 6
    f, g = random_hamiltonian(no, nv, seed=7)
    ints = O.Integrals(**mpqc_blocks(f, g, no))
    eo, ev = np.diag(f)[:no], np.diag(f)[no:]
    t1 = np.zeros((nv, no))
    t2 = np.zeros((nv, nv, no, no))
    tau = O.make_tau(t1, t2)
    r1 = O.cs_ccsd_r1(t1, t2, tau, ints)
    r2 = O.cs_ccsd_r2(t1, t2, tau, ints)
    assert np.array_equal(r1, ints.Fia.T)
    assert np.allclose(r2, ints.Gijab.transpose(2, 3, 0, 1), atol=0, rtol=0)
    t2n = O.jacobi_update_t2(r2, eo, ev)
    e_mp2 = O.energy(ints.Fia, ints.Gijab, np.zeros_like(t1), t2n, 0 * r1, 0 * r2)
    # textbook closed-shell MP2: sum (ia|jb)[2(ia|jb) - (ib|ja)] / (e_i+e_j-e_a-e_b)
    o, v = slice(0, no), slice(no, no + nv)
    iajb = g[o, v, o, v]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    e_ref = float(np.sum(iajb * (2 * iajb - iajb.transpose(0, 3, 2, 1)) / D))
    assert abs(e_mp2 - e_ref) < 1e-14


def test_invariance_under_occ_vir_rotations():
    no, nv = 3, 6
    f, g = random_hamiltonian(no, nv, seed=8, fock_offdiag=0.0)
    E0, *_ = _solve(f, g, no)
    rng = np.random.default_rng(1)
    U = np.eye(no + nv)
    U[:no, :no] = np.linalg.qr(rng.normal(size=(no, no)))[0]
    U[no:, no:] = np.linalg.qr(rng.normal(size=(nv, nv)))[0]
    fr = U.T @ f @ U
    gr = np.einsum("pqrs,pa,qb,rc,sd->abcd", g, U, U, U, U, optimize=True)
    E1, *_ = _solve(fr, gr, no)
    assert abs(E0 - E1) < 1e-11


def test_size_consistency():
    fa, ga = random_hamiltonian(2, 4, seed=9)
    fb, gb = random_hamiltonian(3, 5, seed=10)
    Ea, *_ = _solve(fa, ga, 2)
    Eb, *_ = _solve(fb, gb, 3)
    na, nb = 6, 8
    n = na + nb
    # combined orbital order: occ(A) occ(B) vir(A) vir(B)
    idx_a = list(range(0, 2)) + list(range(5, 9))
    idx_b = list(range(2, 5)) + list(range(9, 14))
    f = np.zeros((n, n))
    g = np.zeros((n, n, n, n))
    f[np.ix_(idx_a, idx_a)] = fa
    f[np.ix_(idx_b, idx_b)] = fb
    g[np.ix_(idx_a, idx_a, idx_a, idx_a)] = ga
    g[np.ix_(idx_b, idx_b, idx_b, idx_b)] = gb
    Eab, *_ = _solve(f, g, 5)
    assert abs(Eab - (Ea + Eb)) < 1e-11


def test_diis_minimises_extrapolated_error():
    """DIIS coefficients sum to 1 and minimise |sum c_k e_k| (Pulay)."""
    rng = np.random.default_rng(2)
    errs = [[rng.normal(size=5), rng.normal(size=(3, 3))] for _ in range(4)]
    d = DIIS(start=1, ndiis=8)
    for e in errs:
        x = [np.zeros(5), np.zeros((3, 3))]
        d.extrapolate(x, e)
    nskip, c = d.last_coefficients
    assert nskip == 0
    assert abs(c.sum() - 1) < 1e-13
    E = np.array([np.concatenate([a.ravel(), b.ravel()]) for a, b in errs])
    # brute force: minimise |E^T c|^2 s.t. sum c = 1 via least squares on the null space
    n = len(errs)
    Z = np.linalg.svd(np.ones((1, n)))[2][1:].T
    c0 = np.ones(n) / n
    y = np.linalg.lstsq(E.T @ Z, -E.T @ c0, rcond=None)[0]
    assert np.abs((c0 + Z @ y) - c).max() < 1e-10


def test_colpiv_qr_solve_matches_lstsq():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(6, 6))
    b = rng.normal(size=6)
    x, det = colpiv_qr_solve(A, b)
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-12)
    assert abs(det - abs(np.linalg.det(A))) < 1e-10 * abs(det)


def test_ccsd_flop_count_matches_published_table():
    # BASELINE: C1 (10,90) 2.338e10, C4 (40,152) 4.749e12, C5 (30,114) 8.517e11
    assert abs(O.ccsd_flops(10, 90) / 2.338e10 - 1) < 1e-3
    assert abs(O.ccsd_flops(40, 152) / 4.749e12 - 1) < 1e-3
    assert abs(O.ccsd_flops(30, 114) / 8.517e11 - 1) < 1e-3
