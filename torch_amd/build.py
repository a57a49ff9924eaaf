"""Build libtgrav.so in-tree with hipcc for gfx950.

    python -m torch_amd.build        # or torch_amd.build.build()

The library is plain HIP C++ behind a C ABI (include/tgrav.h); no torch headers.
Packed FP32 instructions are disabled (-packed-fp32-ops): every FP32 operation is
issued as a scalar-per-lane VALU op.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "tgrav.hip")
DEPS = [os.path.join(HERE, "csrc", f) for f in ("kernels.hpp", "batch.hpp", "bidir.hpp", "grid.hpp", "flash.hpp", "pm.hpp")] + \
       [os.path.join(ROOT, "include", "tgrav.h")]
OUT = os.path.join(HERE, "libtgrav.so")
ARCH = os.environ.get("TGRAV_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compile every time it is called (the does-it-build check)."""
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-o", tmp, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    msgs = [l for l in (res.stdout + res.stderr).splitlines() if noise not in l]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs), file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
