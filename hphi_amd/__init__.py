"""hphi_amd — MI355X-native matrix-free H·v engine for HPhi.

The product is the C ABI library ``libhphi_amd.so`` (include/hphi_amd.h):
device basis tables (sz), diagonal (diagonalcalc), gather-form H·v
(mltply) and device-resident solvers.  This package builds it
(``hphi_amd.build``), binds it (``hphi_amd.capi``) and generates expert-mode
inputs (``hphi_amd.lattice``).
"""

from . import lattice  # noqa: F401

__all__ = ["lattice", "capi", "build"]
