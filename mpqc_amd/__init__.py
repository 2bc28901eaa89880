"""mpqc_amd: MI355X-native (gfx950) closed-shell CCSD residual for MPQC.

The compute path is ``mpqc_amd/lib/libmpqc_ccsd.so`` (HIP, C-ABI in
``include/mpqc_ccsd.h``); ``mpqc_amd.ccsd`` mirrors the reference's solver
interface on top of it.  ``mpqc_amd.chem`` is the host front-end that
produces MO integrals (integrals stay on the host).
"""
from .ccsd import CCSD, Integrals, compute_trange1, pp_ladder  # noqa: F401
