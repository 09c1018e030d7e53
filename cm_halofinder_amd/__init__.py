"""MI355X (gfx950) Friends-of-Friends linking engine for CM_Halofinder.

The compute path is the HIP library ``libcmfof.so`` (C ABI in
``include/fof_gpu.h``); :mod:`cm_halofinder_amd.host` mirrors the
reference's FOF() interface on top of it and :mod:`cm_halofinder_amd.workloads`
builds the synthetic inputs.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib", "host", "workloads"]
