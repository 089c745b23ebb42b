"""MI355X-native (gfx950) sparse-convolution GPIS ray-flight engine.

Drop-in for the flight path behind Tungsten's ``sparse_conv_noise`` medium:
the HIP kernels live in ``csrc/`` and are reached through the C ABI declared in
``include/gpis_hip.h`` (``libgpis_hip.so``); ``medium`` mirrors the medium's
JSON parameters and sampleDistance/transmittance semantics on top of it.
"""
from . import _capi  # noqa: F401
from .medium import DenseMeanGrid, GpisDevice, MediumConfigError, RayBatch, SparseConvolutionNoiseMedium  # noqa: F401

__all__ = ["GpisDevice", "RayBatch", "SparseConvolutionNoiseMedium"]
