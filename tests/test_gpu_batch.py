"""Batched residuals: many molecules' cc::compute_cs_ccsd_r1/r2 in grouped
launches (one hipGraph), checked molecule by molecule against the oracle and
against the single-molecule GPU path, before and after autotuning and across
graph replays with new amplitudes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_residual(blocks, t1, t2):
    from oracle import ccsd as O
    ints = O.Integrals(**blocks)
    tau = O.make_tau(t1, t2)
    r1 = O.cs_ccsd_r1(t1, t2, tau, ints)
    return r1, O.cs_ccsd_r2(t1, t2, tau, ints)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _random_batch(sizes, seed):
    from spinorb_ccsd import mpqc_blocks, random_hamiltonian
    rng = np.random.default_rng(seed)
    mols = []
    for k, (no, nv) in enumerate(sizes):
        f, g = random_hamiltonian(no, nv, seed=seed + k, eri_scale=0.03)
        mols.append(mpqc_blocks(f, g, no))
    return mols, rng


def _amplitudes(rng, no, nv, scale=0.02):
    t1 = rng.normal(0, scale, (nv, no))
    t2 = rng.normal(0, scale, (nv, nv, no, no))
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_batched_residuals_match_oracle(prec, tol):
    from mpqc_amd.batch import CCSDBatch
    sizes = [(2, 7), (3, 11), (4, 16), (5, 9), (3, 25), (6, 19), (2, 30), (4, 13), (5, 22), (1, 5)]
    mols, rng = _random_batch(sizes, seed=40)
    with CCSDBatch([s[0] for s in sizes], [s[1] for s in sizes], precision=prec) as B:
        for k, b in enumerate(mols):
            B.set_integrals(k, b)
        for rnd in range(3):  # first run (plan + capture), a replay, a replay after tuning
            amps = [_amplitudes(rng, no, nv) for no, nv in sizes]
            for k, (t1, t2) in enumerate(amps):
                B.set_amplitudes(k, t1, t2)
            if rnd == 2:
                B.tune(reps=1)
            B.residual()
            for k, (t1, t2) in enumerate(amps):
                r1, r2 = B.get_residuals(k)
                o1, o2 = _oracle_residual(mols[k], t1, t2)
                e1, e2 = _rel(r1, o1), _rel(r2, o2)
                print(f"round {rnd} molecule {k} {sizes[k]}: r1 {e1:.2e} r2 {e2:.2e}")
                assert e1 < tol and e2 < tol


def test_batched_hchains_match_single_molecule_path():
    from mpqc_amd.batch import CCSDBatch
    from mpqc_amd.ccsd import CCSD, Integrals
    from mpqc_amd.chem.molecule import h_chain
    from mpqc_amd.chem.mo import ccsd_integrals
    from mpqc_amd.chem.scf import rhf
    from oracle import ccsd as O
    chains = [(4, 1.6), (6, 1.8), (8, 2.0), (4, 1.9)]
    inps = [ccsd_integrals(rhf(h_chain(n, s), "cc-pVDZ")) for n, s in chains]
    with CCSDBatch([i.o for i in inps], [i.v for i in inps]) as B:
        amps = []
        for k, inp in enumerate(inps):
            B.set_integrals(k, inp.blocks)
            # first-iteration amplitudes t = R(0) / D
            ints = O.Integrals(**inp.blocks)
            t1 = O.jacobi_update_t1(ints.Fia.T, inp.eps_occ, inp.eps_vir)
            t2 = O.jacobi_update_t2(ints.Gijab.transpose(2, 3, 0, 1), inp.eps_occ, inp.eps_vir)
            B.set_amplitudes(k, t1, t2)
            amps.append((t1, t2))
        B.residual()
        for k, inp in enumerate(inps):
            r1, r2 = B.get_residuals(k)
            with CCSD(inp.o, inp.v) as c:
                c.set_integrals(Integrals(**inp.blocks))
                s1, s2 = c.compute_residuals(*amps[k])
            o1, o2 = _oracle_residual(inp.blocks, *amps[k])
            print(f"H{chains[k][0]} o={inp.o} v={inp.v}: vs single {_rel(r1, s1):.2e} {_rel(r2, s2):.2e}"
                  f" vs oracle {_rel(r1, o1):.2e} {_rel(r2, o2):.2e}")
            # (the single-molecule context runs the pair-symmetric pp ladder: sums of a
            # different order than the batched full contraction)
            assert _rel(r1, s1) < 1e-12 and _rel(r2, s2) < 1e-12
            assert _rel(r1, o1) < 1e-12 and _rel(r2, o2) < 1e-12
        st = B.stages()
        assert sum(s["kind"] == "contraction" for s in st) > 30
