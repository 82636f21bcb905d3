"""perix_amd -- MI355X-native PDDO hot path of PeriX.

The compute lives in libpxg.so (hand-written HIP for gfx950, C ABI declared
in include/pxg.h).  This package is a thin ctypes mirror of that ABI plus
host helpers that follow the reference's driver logic (structured point
cloud, BC lists, the explicit step schedule).  There is no CPU fallback:
importing the device API without the built library raises.
"""
from .pxg import (Context, PxgError, lib_path, load, FP64, FP32C, FracParams, BC,
                  available_devices)
from .mesh import StructuredGrid, build_structured_cloud, horizon_radius, WALLS
from . import problems

__all__ = ["Context", "PxgError", "lib_path", "load", "FP64", "FP32C", "FracParams", "BC",
           "available_devices", "StructuredGrid", "build_structured_cloud", "horizon_radius",
           "WALLS", "problems"]
