"""Writes tests/golden/abi_residual_o2_v5.txt: one CCSD residual and one
converged CCSD energy of the oracle (oracle/ccsd.py) on a small random
closed-shell Hamiltonian, for the C++ program tests/abi_c_check.cpp, which
drives the C-ABI the way a C++ host (MPQC's CCSD::compute_ccsd) would.

Format: whitespace-separated text.
  o v
  n_occ_tiles occ_hashmarks...      (uneven TiledRange1s)
  n_vir_tiles vir_hashmarks...
  then records "name count values..." for Fia Fij Fab Gijab Gijkl Giajb Giabc
  Gijka Gabcd (dense, the reference's index orders), eps_occ, eps_vir, t1, t2
  (dense "a,i" / "a,b,i,j"), r1, r2 (the oracle's residual at t1, t2) and
  E_ccsd (the oracle's energy converged to 1e-12).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import ccsd as O  # noqa: E402
from spinorb_ccsd import mpqc_blocks, random_hamiltonian  # noqa: E402

NAMES = ["Fia", "Fij", "Fab", "Gijab", "Gijkl", "Giajb", "Giabc", "Gijka", "Gabcd"]


def main():
    no, nv = 2, 5
    f, g = random_hamiltonian(no, nv, seed=77, eri_scale=0.05)
    b = mpqc_blocks(f, g, no)
    eo, ev = np.diag(f)[:no].copy(), np.diag(f)[no:].copy()
    rng = np.random.default_rng(77)
    t1 = rng.normal(0, 0.02, (nv, no))
    t2 = rng.normal(0, 0.02, (nv, nv, no, no))
    ints = O.Integrals(**b)
    tau = O.make_tau(t1, t2)
    r1 = O.cs_ccsd_r1(t1, t2, tau, ints)
    r2 = O.cs_ccsd_r2(t1, t2, tau, ints)
    E, _, _, _ = O.compute_ccsd(O.Integrals(**b), eo, ev, max_iter=100, target_precision=1e-12)
    out = [f"{no} {nv}", "2 0 1 2", "2 0 2 5"]

    def rec(name, a):
        a = np.asarray(a, dtype=np.float64).ravel()
        out.append(f"{name} {a.size} " + " ".join("%.17g" % x for x in a))

    for n in NAMES:
        rec(n, b[n])
    rec("eps_occ", eo)
    rec("eps_vir", ev)
    rec("t1", t1)
    rec("t2", t2)
    rec("r1", r1)
    rec("r2", r2)
    rec("E_ccsd", [E])
    path = os.path.join(HERE, "abi_residual_o2_v5.txt")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    print("wrote", path, f"E = {E:.14f}")


if __name__ == "__main__":
    main()
