"""SurfaceGrowth per-step MD engine for AMD MI355X (gfx950).

The product is the C-ABI library ``lib/libsurfacegrowth.so`` (include/surfacegrowth.h: the
reference's wrapper API; include/sgk.h: the stage-level engine) and the console driver
``lib/surfacegrowth``.  This package only loads and drives them.
"""
from . import _native
from .engine import Batch, Engine, System, make_options, make_system, record_dict
from .configs import CONFIGS, config_system

__all__ = ["Batch", "Engine", "System", "make_system", "make_options", "record_dict", "CONFIGS", "config_system",
           "_native"]
