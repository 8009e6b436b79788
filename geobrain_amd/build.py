"""Build recipe for the HIP library (gfx950 only).

``python -m geobrain_amd.build`` compiles ``csrc/gbw_acoustic2d.hip`` (2D
propagator, batched forward, shared entry points) and ``csrc/gbw_acoustic3d.hip``
(3D propagator) each to an object, then links them into ``libgbw_acoustic2d.so``
next to this file.  The flags:

* ``--offload-arch=gfx950``: MI355X (CDNA4) code objects only.
* ``-ffp-contract=off``: no fused multiply-add contraction, so every
  floating-point operation rounds once, in the order the reference's FP32
  PyTorch ops use.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math.
* 3D source only: ``-mllvm -amdgpu-sched-strategy=iterative-ilp``, the
  backend's instruction-level-parallelism scheduler (3D adjoint p kernel 901 ->
  864 us per step on 8 shots of 240^3; the 2D kernels measured about as fast
  or slower with it, so they keep the default scheduler).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, "csrc", "gbw_acoustic2d.hip"),
        os.path.join(HERE, "csrc", "gbw_acoustic3d.hip")]
LIB = os.path.join(HERE, "libgbw_acoustic2d.so")

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per-source additions
SRC_FLAGS = {
    "gbw_acoustic3d.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build geobrain_amd")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)


def build(verbose: bool = False) -> str:
    """Compile the library; returns its path.  Always recompiles."""
    tmp = LIB + ".tmp"
    objs = []
    try:
        for src in SRCS:
            obj = os.path.join(HERE, os.path.basename(src) + ".o.tmp")
            _run([hipcc(), *HIPCC_FLAGS, *SRC_FLAGS.get(os.path.basename(src), []), "-I",
                  os.path.join(ROOT, "include"), "-c", src, "-o", obj], verbose)
            objs.append(obj)
        _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    finally:
        for obj in objs:
            if os.path.exists(obj):
                os.remove(obj)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
