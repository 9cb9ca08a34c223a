"""Builds libxcquad.so (the HIP library) in place with hipcc for gfx950.

    python -m pdynamo3_amd.build        (or build() from Python)

The library is written next to this file; it is git-ignored and travels with the source tree.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBRARY = os.path.join(HERE, "libxcquad.so")
SOURCES = ["xq_alloc.hip", "xq_api.hip", "xq_grid.hip", "xq_basis.hip", "xq_integrate.hip", "xq_f64.hip", "xq_mixed.hip",
           "xq_gradients.hip", "xq_uks.hip", "xq_noalloc.hip", "qs_integrals.hip", "qs_dense.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
         # Packed FP32 instructions stay off: hipcc would otherwise form v_pk_* FP32 ops from plain float code.
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]
# rocBLAS and rocSOLVER for the SCF's dense steps (qs_dense.hip).
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build(verbose=False):
    """Compile every HIP source into libxcquad.so; raises CalledProcessError on a compiler error."""
    cmd = [HIPCC] + FLAGS + ["-o", LIBRARY + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES] + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIBRARY + ".tmp", LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIBRARY)
