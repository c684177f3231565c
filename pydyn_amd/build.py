"""Build libtm5hip.so in-tree with hipcc for gfx950.

    python -m pydyn_amd.build

FP64 everywhere with no contraction (-ffp-contract=off): every add and multiply rounds on its
own, as NumPy's do, which is what makes the kernels bit-identical to the reference.  Packed
FP32 instructions are disabled (they are not used: the library has no FP32 arithmetic).
"""

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libtm5hip.so"
SOURCES = ["advect.hip", "convec.hip", "column.hip", "lapack.hip", "physics.hip", "zoom.hip", "tree.hip", "adjoint.hip",
           "api.hip"]
ARCH = os.environ.get("TM5_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the column kernel's fully unrolled LU and substitutions (register-indexed arrays) exceed
    # the default size limit for an unroll pragma; unrolled they need no scratch memory
    "-mllvm", "-pragma-unroll-threshold=200000",
    "-Wall",
    "-Wno-unused-function",
]  # fmt: skip


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(exe).exists():
        raise RuntimeError("hipcc not found: ROCm is needed to build libtm5hip")
    return exe


def build(verbose=False):
    """Compile every source into pydyn_amd/libtm5hip.so (always rebuilds)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", *(str(CSRC / s) for s in SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
