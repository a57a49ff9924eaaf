"""The C ABI from a C program (tests/c/abi_drive.c): include/tgrav.h compiles as C99 and links
against libtgrav.so (CPU), and the program's host-pointer evaluations match the oracle (GPU).

Bars as in test_gpu_parity: FP64 within 1e-12 normwise relative error per target; mixed
<= 1e-6 median and <= 1e-4 max relative acceleration error; the kick from v = 0 is dt * a,
bit for bit.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from tests import _util
from torch_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "abi_drive.c")
LIBDIR = os.path.join(ROOT, "torch_amd")
G = W.G_PC_MSUN_KMS


def build(out_dir):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc is not installed")
    if not os.path.exists(os.path.join(LIBDIR, "libtgrav.so")):
        pytest.skip("libtgrav.so is not built (python -m torch_amd.build)")
    exe = os.path.join(str(out_dir), "abi_drive")
    cmd = [gcc, "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           SRC, "-L", LIBDIR, "-ltgrav", "-Wl,-rpath," + LIBDIR, "-Wl,-rpath-link,/opt/rocm/lib",
           "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def test_header_compiles_as_c_and_links(tmp_path):
    exe = build(tmp_path)
    assert os.access(exe, os.X_OK)


def parse(text):
    rows = {}
    for line in text.splitlines():
        tag, *vals = line.split()
        rows.setdefault(tag, []).append([float(v) for v in vals])
    return {k: np.array(v) for k, v in rows.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_c_program_matches_oracle(tmp_path, mode):
    exe = build(tmp_path)
    print("abi_drive", mode, flush=True)
    res = subprocess.run([exe, "0" if mode == "fp64" else "1"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    d = parse(res.stdout)
    S, T, K, B, C = d["S"], d["T"], d["K"], d["B"], d["C"]
    ref = oracle.eval_plummer(S[:, 0], S[:, 1], S[:, 2], S[:, 3], S[:, 4], T[:, 0], T[:, 1], T[:, 2], None, G,
                              mode=oracle.NEUMAIER)
    got = tuple(T[:, 3 + q] for q in range(4))
    s = _util.summary(got, ref)
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
    else:
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, s
    for q in range(3):
        np.testing.assert_array_equal(K[:, q], 0.25 * T[:, 3 + q])
    bx, by, bz = oracle.block_centres(B[:, 0], B[:, 1], B[:, 2], B[:, 3])
    refc = oracle.eval_plummer(S[:, 0], S[:, 1], S[:, 2], S[:, 3], S[:, 4], bx, by, bz, None, G,
                               mode=oracle.NEUMAIER)
    s = _util.summary(tuple(C[:, q] for q in range(4)), refc)
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
    else:
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, s
