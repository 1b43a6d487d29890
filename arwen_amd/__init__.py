"""arwen_amd: MI355X-native (gfx950 HIP) WRF-ARW-class dycore step.

A drop-in for the reference's ``gpuwm.core.dycore.step`` hot path: the RK3
large step with flux-form 5th/3rd-order advection and split-explicit
acoustic substeps with the vertically implicit w/phi solve, on FP32
PyTorch-ROCm tensors in the reference state layout.
"""
from .config import RunConfig, validate_for_step
from .grid import (BaseState, VerticalCoord, make_base_state,
                   make_vertical_coord, rebalance_hydrostatic)

__version__ = "0.1.0"

__all__ = ["RunConfig", "validate_for_step", "BaseState", "VerticalCoord",
           "make_base_state", "make_vertical_coord", "rebalance_hydrostatic"]
