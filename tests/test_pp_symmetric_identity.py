"""The pair-symmetric form of the particle-particle ladder (mpqc_amd/csrc/ccsd_ppsym.hip)
against the full contraction b(a,b,i,j) = sum_cd Gabcd(a,b,c,d) tau(c,d,i,j)
(ccsd_r1_r2.h:275), in numpy with the kernels' pair numbering and packing: integrals
with the (ab)(cd) symmetry of (ac|bd), amplitudes without any symmetry."""
import numpy as np
import pytest


def _sym_pair(p, q, n):
    return p * n - p * (p - 1) // 2 + (q - p)


def _anti_pair(p, q, n):
    return p * (n - 1) - p * (p - 1) // 2 + (q - p - 1)


def _pair_symmetric_ladder(G, tau):
    v, o = G.shape[0], tau.shape[2]
    sym = [(p, q) for p in range(v) for q in range(p, v)]
    anti = [(p, q) for p in range(v) for q in range(p + 1, v)]
    assert all(_sym_pair(p, q, v) == k for k, (p, q) in enumerate(sym))
    assert all(_anti_pair(p, q, v) == k for k, (p, q) in enumerate(anti))
    t = tau.reshape(v, v, o * o)
    Vp = np.array([[G[a, b, c, d] + G[a, b, d, c] for (c, d) in sym] for (a, b) in sym]).reshape(len(sym), len(sym))
    S = np.array([(0.5 if c == d else 1.0) * (t[c, d] + t[d, c]) for (c, d) in sym]).reshape(len(sym), o * o)
    P = Vp @ S
    if anti:
        Vm = np.array([[G[a, b, c, d] - G[a, b, d, c] for (c, d) in anti] for (a, b) in anti])
        A = np.array([t[c, d] - t[d, c] for (c, d) in anti])
        M = Vm @ A
    out = np.zeros((v, v, o * o))
    for a in range(v):
        for b in range(v):
            p, q = min(a, b), max(a, b)
            x = P[_sym_pair(p, q, v)].copy()
            if a != b:
                m = M[_anti_pair(p, q, v)]
                x += m if a < b else -m
            out[a, b] = 0.5 * x
    return out.reshape(v, v, o, o)


@pytest.mark.parametrize("o,v", [(3, 5), (1, 4), (2, 1), (4, 7)])
def test_pair_symmetric_ladder_equals_full_contraction(o, v):
    rng = np.random.default_rng(o * 10 + v)
    G = rng.normal(size=(v, v, v, v))
    G = G + G.transpose(1, 0, 3, 2)            # (ac|bd) = (bd|ac)
    tau = rng.normal(size=(v, v, o, o))        # no symmetry assumed
    ref = np.einsum("abcd,cdij->abij", G, tau)
    got = _pair_symmetric_ladder(G, tau)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_pair_symmetric_ladder_needs_the_integral_symmetry():
    # without Gabcd(b,a,d,c) = Gabcd(a,b,c,d) the form differs: the symmetry is what it uses
    rng = np.random.default_rng(1)
    G = rng.normal(size=(4, 4, 4, 4))
    tau = rng.normal(size=(4, 4, 2, 2))
    ref = np.einsum("abcd,cdij->abij", G, tau)
    assert np.abs(_pair_symmetric_ladder(G, tau) - ref).max() > 1e-3
