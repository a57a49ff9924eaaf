"""VALU throughput of the instruction kinds the direct-sum kernels are made of (valu_rates.hip):

    python -m torch_amd.probe.valu_rates          # builds with hipcc for gfx950, then runs

The rates are the measured ceilings for instruction mixes: the mixed kernels are made of FP32
FMA / mul / add plus v_rsq_f32, the FP64 kernels of FP64 FMA / mul / add plus conversions and the
FP32 seed.  Per kind it prints the wave64 instructions issued per second per SIMD (one per 4
cycles at the published 2.4 GHz would be 0.6 G/s, the rate a one-operation-per-lane-per-clock
peak assumes) and the flop rate of the whole device.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "valu_rates.hip")
EXE = os.path.join(HERE, "valu_rates")


def build():
    from torch_amd.build import hipcc, ARCH
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-o", EXE, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    return EXE


def main():
    if not os.path.exists(EXE) or "--build" in sys.argv:
        build()
    if "--build-only" in sys.argv:
        return 0
    return subprocess.call([EXE])


if __name__ == "__main__":
    sys.exit(main())
