"""icflow_amd -- MI355X (gfx950) implementation of icFlow's per-Newton-iteration assembly path.

The device library is libicflow_hip.so (C ABI: include/icflow_hip.h, include/icflow_host.h);
this package holds its build script, its ctypes binding and a Python mirror of the reference's
GPU_Functionality host class for driving it with the reference's buffers.
"""
from ._lib import ICF_FP64, ICF_MIXED, load  # noqa: F401
