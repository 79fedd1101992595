"""MI355X (gfx950) simulation core for NeuralSim: HIP kernels behind a C ABI (include/nsx.h).

The compute path is libnsx.so only; importing this package does not load it, the first
engine call does (and raises NativeLibraryMissing if it was not built).
"""
from . import abi, configs  # noqa: F401
from ._lib import NativeLibraryMissing, NsxError, device_count, load  # noqa: F401
from .engine import BatchedPopulation, GpuBackend, SimulationEngine  # noqa: F401

__all__ = ["abi", "configs", "load", "device_count", "SimulationEngine", "GpuBackend",
           "BatchedPopulation",
           "NativeLibraryMissing", "NsxError"]
