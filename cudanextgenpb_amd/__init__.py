"""MI355X (gfx950) linearized Poisson-Boltzmann hot path of NextGenPB.

Native pieces: ``lib/libngpb.so`` (HIP kernels + C ABI, include/ngpb.h) and
``lib/libngpb_workload.so`` (host-side synthetic inputs). Build them in-tree with
``python -m cudanextgenpb_amd.build``.
"""
from . import _paths  # noqa: F401

__all__ = ["native", "pb", "workload", "build"]
