"""MI355X exchange-correlation quadrature for pDynamo3's QCModelDFT (hand-written HIP kernels behind a C-ABI).

The compute library (libxcquad.so) is loaded lazily by pdynamo3_amd.integrator; importing this package alone needs no GPU.
"""

__all__ = ["definitions", "basis", "molecules"]
