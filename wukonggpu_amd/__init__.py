"""wukonggpu_amd: MI355X-native graph exploration for Wukong's GPU engine path.

Host C++ (gstore builder, engine) over a C ABI to hand-written HIP kernels for
gfx950. See include/wukong_gstore.h, include/wukong_gpu.h and
include/wukong_engine.h for the native interfaces.
"""
from .gstore import GStore, Triples, VID_BASE, TYPE_ID, PREDICATE_ID, IN, OUT  # noqa: F401
from .engine import Device, Table, GPUEngine, Batch, DeviceError, pack_queries, plan_of  # noqa: F401
from ._native import STEP_I2U, STEP_C2U, STEP_K2U, STEP_K2K, STEP_K2C  # noqa: F401
