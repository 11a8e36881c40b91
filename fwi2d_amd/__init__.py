"""fwi2d_amd: MI355X-native elastic wave-propagation and FWI-gradient engine.

The compute path is libfwi2d_hip.so (hand-written HIP for gfx950) behind the C ABI
of include/fwi2d_hip.h; this package only binds it (fwi2d_amd.lib) and provides
synthetic workloads (fwi2d_amd.workloads).
"""
from .lib import AcousticEngine, Engine, Fwi2dError, Nlls, WeOp, filter_apply, filter_prepare, grid_from, lbfgs, load  # noqa: F401

__all__ = ["AcousticEngine", "Engine", "WeOp", "Nlls", "lbfgs", "filter_prepare", "filter_apply", "Fwi2dError", "grid_from", "load"]
