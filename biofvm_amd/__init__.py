"""biofvm_amd — MI355X-native diffusion-decay engine for BioFVM's hot path.

The compute path is ``libbiofvm_amd.so``: hand-written gfx950 HIP kernels for
the LOD Thomas sweeps, Dirichlet nodes, bulk and agent sources/sinks, behind
the C ABI in ``include/biofvm_amd.h``.  This package is a thin ctypes layer
over that ABI (``device``) plus seeded synthetic inputs (``workloads``).
"""
from ._capi import ALL, BULK, CELLS, DIFFUSION, FP32, FP64, BfvmError, LIB_PATH, load
from .device import DeviceMicroenvironment, Ensemble, run_case, setup_case

__all__ = ["ALL", "BULK", "CELLS", "DIFFUSION", "FP32", "FP64", "BfvmError", "LIB_PATH", "load",
           "DeviceMicroenvironment", "Ensemble", "run_case", "setup_case"]
