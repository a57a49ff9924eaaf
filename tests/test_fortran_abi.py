"""The C ABI from Fortran (tests/fortran/): the interface module of the integration guide
compiles with the ROCm Fortran compiler and links against libtgrav.so (CPU), and a program
that calls the library as FLASH's get_accel_gas_on_particles would matches the oracle (GPU).

Bars as in test_c_abi: FP64 within 1e-12 normwise relative error per target; mixed <= 1e-6
median and <= 1e-4 max relative acceleration error; the kick from v = 0 is dt * a, bit for bit.
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
SRC = [os.path.join(ROOT, "tests", "fortran", f) for f in ("tgrav_c.f90", "reverse_drive.f90")]
LIBDIR = os.path.join(ROOT, "torch_amd")
G = W.G_PC_MSUN_KMS


def compiler():
    for cand in ("/opt/rocm/bin/amdflang", shutil.which("amdflang") or "", shutil.which("flang") or ""):
        if cand and os.path.exists(cand):
            return cand
    return None


def build(out_dir):
    fc = compiler()
    if fc is None:
        pytest.skip("no Fortran compiler (amdflang) in this image")
    if not os.path.exists(os.path.join(LIBDIR, "libtgrav.so")):
        pytest.skip("libtgrav.so is not built (python -m torch_amd.build)")
    exe = os.path.join(str(out_dir), "reverse_drive")
    cmd = [fc, "-O1", "-J", str(out_dir), *SRC, "-L", LIBDIR, "-ltgrav", "-Wl,-rpath," + LIBDIR,
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    return exe


def test_interface_module_compiles_and_links(tmp_path):
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
def test_fortran_program_matches_oracle(tmp_path, mode):
    exe = build(tmp_path)
    print("reverse_drive", mode, flush=True)
    res = subprocess.run([exe, "0" if mode == "fp64" else "1"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "STOP" not in res.stderr, res.stderr
    d = parse(res.stdout)
    S, T, K = d["S"], d["T"], d["K"]
    assert len(S) == 600 and len(T) == 257
    ref = oracle.eval_plummer(S[:, 0], S[:, 1], S[:, 2], S[:, 3], S[:, 4], T[:, 0], T[:, 1], T[:, 2], None, G,
                              mode=oracle.NEUMAIER)
    s = _util.summary(tuple(T[:, 3 + q] for q in range(4)), ref)
    print(mode, s, flush=True)
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
    else:
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, s
    for q in range(3):
        np.testing.assert_array_equal(K[:, q], 0.25 * T[:, 3 + q])
