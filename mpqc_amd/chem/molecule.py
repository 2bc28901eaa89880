"""Molecules and seeded geometry generators.

Units: positions are stored in bohr.  Å -> bohr uses the 2010 CODATA bohr
radius 0.52917721092 Å (reference: src/mpqc/chemistry/units/units.h:103; the
validation inputs select "units": "2010CODATA").
"""
from __future__ import annotations

import dataclasses
from typing import List, Tuple

import numpy as np

from .basis_data import CORE_ELECTRONS, Z

BOHR_2010 = 0.52917721092  # Å


@dataclasses.dataclass
class Molecule:
    symbols: List[str]
    xyz: np.ndarray  # (natom, 3) bohr
    charge: int = 0

    @classmethod
    def from_angstrom(cls, atoms: List[Tuple[str, float, float, float]], charge=0):
        sym = [a[0] for a in atoms]
        xyz = np.array([[a[1], a[2], a[3]] for a in atoms], dtype=np.float64) / BOHR_2010
        return cls(sym, xyz, charge)

    @classmethod
    def from_xyz_text(cls, text: str):
        lines = [l for l in text.strip().splitlines()]
        n = int(lines[0].split()[0])
        atoms = []
        for l in lines[2:2 + n]:
            p = l.split()
            atoms.append((p[0], float(p[1]), float(p[2]), float(p[3])))
        return cls.from_angstrom(atoms)

    @property
    def Z(self) -> np.ndarray:
        return np.array([Z[s] for s in self.symbols], dtype=np.float64)

    @property
    def n_electrons(self) -> int:
        return int(self.Z.sum()) - self.charge

    @property
    def core_electrons(self) -> int:
        """Number of core electrons (reference: molecule.cpp:190-204)."""
        return sum(CORE_ELECTRONS[s] for s in self.symbols)

    def nuclear_repulsion(self) -> float:
        e = 0.0
        Zs = self.Z
        for i in range(len(Zs)):
            for j in range(i):
                e += Zs[i] * Zs[j] / np.linalg.norm(self.xyz[i] - self.xyz[j])
        return float(e)


# ---------------------------------------------------------------- generators
def h_chain(n: int, spacing_bohr: float) -> Molecule:
    """Linear chain of n H atoms along z with uniform spacing (bohr)."""
    xyz = np.zeros((n, 3))
    xyz[:, 2] = spacing_bohr * np.arange(n)
    return Molecule(["H"] * n, xyz)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(dims=(2, 2, 2), seed=4, oo=2.8, jitter=0.1, r_oh=0.9572,
                  angle_deg=104.52) -> Molecule:
    """Water molecules with O on a cubic lattice of edge `oo` Å, each O
    coordinate jittered by U[-jitter, jitter] Å, monomer geometry r_OH, HOH
    angle, and a uniformly random orientation per monomer (numpy PCG64 seed)."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(angle_deg)
    mono = np.array([[0.0, 0.0, 0.0],
                     [r_oh * np.sin(th / 2), 0.0, r_oh * np.cos(th / 2)],
                     [-r_oh * np.sin(th / 2), 0.0, r_oh * np.cos(th / 2)]])
    atoms = []
    for ix in range(dims[0]):
        for iy in range(dims[1]):
            for iz in range(dims[2]):
                O = np.array([ix, iy, iz], dtype=float) * oo + rng.uniform(-jitter, jitter, 3)
                R = _random_rotation(rng)
                pos = O + mono @ R.T
                for s, p in zip(["O", "H", "H"], pos):
                    atoms.append((s, *p))
    return Molecule.from_angstrom(atoms)


def h_chain_batch(count=256, seed=3, nmin=8, nmax=32, smin=1.6, smax=2.0):
    """Seeded batch of H chains: n uniform over even {nmin..nmax}, spacing
    U[smin, smax] bohr."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.arange(nmin, nmax + 1, 2), size=count)
    sp = rng.uniform(smin, smax, size=count)
    return [h_chain(int(n), float(s)) for n, s in zip(ns, sp)]
