"""Host chemistry front-end: molecules, Gaussian basis sets, integrals, RHF and
the AO->MO transform that fills the CCSD integral blocks.

This is the part MPQC keeps on the host (libint2 integrals, SCF, LCAOFactory's
4-index transform); it feeds the GPU CCSD path through the C-ABI.
"""
from .molecule import Molecule, BOHR_2010  # noqa: F401
from .basis import BasisSet  # noqa: F401
from .scf import rhf  # noqa: F401
from .mo import ccsd_integrals  # noqa: F401
