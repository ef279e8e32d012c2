"""vampire_amd — MI355X-native stochastic LLG-Heun / Metropolis hot path for VAMPIRE.

The compute path is libvmx.so (HIP kernels for gfx950 behind the C ABI of include/vmx.h).
Python here builds systems in the reference's data model and marshals arrays; it never
computes spin dynamics itself.
"""
from . import system  # noqa: F401

__all__ = ["system"]
