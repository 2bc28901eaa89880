"""Parity at the full workload sizes, through properties that do not depend on
an oracle run at that size:

* (H2O)8 cc-pVDZ (o=40, v=152, all electrons): the split-precision (f16x3)
  residual against the FP64 residual at MP2 amplitudes, within the bound the
  split form meets against the oracle at small sizes (2e-6 relative at o=6,
  v=33) scaled by sqrt(K) of the longest contraction (the FP32 chunk sums and
  the dropped lo*lo products add up as a random walk over K = v^2), and the
  (ab)(ij) symmetry of r2 that the symmetrisation and the symmetric ladders
  give for symmetric amplitudes (relative to the integrals' scale: the host
  transform's integrals are symmetric only to round-off);
* the particle-particle ladder kernel at o=32, v=256 (V = 34 GB): checksums of
  checksums, w^T R = (w^T V) tau and R u = V (tau u) for random w, u, with the
  right-hand sides formed by torch's own reductions and products.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mp2_amplitudes(inp):
    b = inp.blocks
    eo, ev = inp.eps_occ, inp.eps_vir
    d1 = eo[None, :] - ev[:, None]
    d2 = (eo[None, None, :, None] + eo[None, None, None, :] - ev[:, None, None, None] - ev[None, :, None, None])
    return b["Fia"].T / d1, b["Gijab"].transpose(2, 3, 0, 1) / d2


def test_h2o8_split_residual_within_bound_of_fp64():
    from mpqc_amd.ccsd import CCSD, Integrals
    from mpqc_amd.chem.mo import ccsd_integrals
    from mpqc_amd.chem.molecule import water_cluster
    from mpqc_amd.chem.scf import rhf
    inp = ccsd_integrals(rhf(water_cluster((2, 2, 2), seed=4), "cc-pVDZ"), frozen_core=False)
    assert (inp.o, inp.v) == (40, 152)
    t1, t2 = _mp2_amplitudes(inp)
    ints = Integrals(**inp.blocks)
    res = {}
    for prec in ("f64", "f16x3"):
        print(f"(H2O)8 residual, {prec}", flush=True)
        with CCSD(inp.o, inp.v, precision=prec) as c:
            c.set_integrals(ints)
            c.set_orbital_energies(inp.eps_occ, inp.eps_vir)
            res[prec] = c.compute_residuals(t1, t2)
    r1, r2 = res["f64"]
    s1, s2 = res["f16x3"]
    e1 = np.abs(s1 - r1).max() / np.abs(r1).max()
    e2 = np.abs(s2 - r2).max() / np.abs(r2).max()
    bound = 2e-6 * np.sqrt((inp.v / 33.0) ** 2)     # small-size bound x sqrt(K / K_small), K = v^2
    sym = np.abs(r2 - r2.transpose(1, 0, 3, 2)).max() / np.abs(inp.blocks["Gijab"]).max()
    print(f"f16x3 vs f64: r1 {e1:.2e} r2 {e2:.2e} (bound {bound:.1e}); r2 (ab)(ij) asymmetry {sym:.2e}")
    assert e1 < bound and e2 < bound
    assert max(e1, e2) > 1e-12          # the split form really ran
    assert sym < 1e-12                  # round-off of long sums; an index error is O(1)


def test_pp_ladder_full_size_checksums():
    torch = pytest.importorskip("torch")
    from mpqc_amd.ccsd import pp_ladder
    o, v = 32, 256
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    print(f"pp-ladder o={o} v={v}: V {v ** 4 * 8 / 2 ** 30:.1f} GiB", flush=True)
    V = torch.empty((v * v, v * v), dtype=torch.float64, device="cuda")
    V.normal_(0.0, 0.05, generator=g)
    tau = torch.empty((v * v, o * o), dtype=torch.float64, device="cuda")
    tau.normal_(0.0, 1e-2, generator=g)
    R = torch.zeros((v * v, o * o), dtype=torch.float64, device="cuda")
    pp_ladder(V.data_ptr(), tau.data_ptr(), R.data_ptr(), o, v)
    torch.cuda.synchronize()
    w = torch.empty(v * v, dtype=torch.float64, device="cuda")
    w.normal_(generator=g)
    u = torch.empty(o * o, dtype=torch.float64, device="cuda")
    u.normal_(generator=g)
    left = w @ R                      # w^T R
    left_ref = (w @ V) @ tau          # (w^T V) tau
    right = R @ u                     # R u
    right_ref = V @ (tau @ u)         # V (tau u)
    el = ((left - left_ref).abs().max() / left_ref.abs().max()).item()
    er = ((right - right_ref).abs().max() / right_ref.abs().max()).item()
    print(f"w^T R {el:.2e}  R u {er:.2e}")
    assert el < 1e-11 and er < 1e-11
