"""MI355X (gfx950) implementation of EFDC+ 8.5's external-mode solve
(CALPUV2C + CONGRAD) and constituent transport (CALCONC + CALTRAN).

``efdc_amd.lib.Context`` is the host-side mirror of the C-ABI declared in
include/efdc_hip.h; ``efdc_amd.grids`` builds synthetic cases in the
var_global layout.
"""
from . import abi, grids  # noqa: F401

__all__ = ["abi", "grids", "lib", "build"]
