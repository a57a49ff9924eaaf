"""Writes the committed golden fixtures under tests/golden/ (run from the repo root):

    python tests/golden/make_golden.py

plummer_kat.npz  -- known-answer cases of the Plummer direct sum computed by the 60-digit
                    decimal brute force (oracle/brute.py), not by any floating-point code:
                    inputs plus expected (ax, ay, az, phi) rounded once to FP64.
c1_oracle.npz    -- the c1 workload (seed 1001): its inputs and the compensated FP64
                    oracle's forward (stars -> 16^3 grid) and reverse (cells -> stars)
                    outputs, the reference CPU result on the same seeded inputs.

The reference ships no golden vectors for this path (its gravity tests are an empty
class, src/amuse/torch_amuse_flash/tests/test_flash.py:1-6), so these fixtures are the
project's own, pinned by the brute force and by closed forms in tests/test_oracle.py.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import brute  # noqa: E402
from torch_amd import workloads as W  # noqa: E402

G = W.G_PC_MSUN_KMS
OUT = os.path.dirname(os.path.abspath(__file__))


def kat_cases():
    rng = np.random.Generator(np.random.PCG64(7))
    cases = []
    # 1: one source, one target (closed form)
    cases.append(dict(s=np.array([[0.3, -0.2, 0.1]]), m=np.array([2.5]), se=np.array([0.01]),
                      t=np.array([[1.0, 0.5, -0.7]]), te=np.array([0.0])))
    # 2: 16 random sources x 32 random targets, mixed softening, incl. eps = 0
    s = rng.uniform(-1, 1, (16, 3))
    t = rng.uniform(-1, 1, (32, 3))
    se = rng.uniform(0, 0.05, 16)
    se[:4] = 0.0
    te = rng.uniform(0, 0.05, 32)
    te[::3] = 0.0
    t[0] = s[0]            # coincident with eps_s = eps_t = 0: self-exclusion
    te[0] = 0.0
    t[1] = s[5]            # coincident but softened: finite, contributes phi only
    cases.append(dict(s=s, m=W.kroupa_masses(rng, 16), se=se, t=t, te=te))
    # 3: a tight binary (1 AU) far from the origin, targets near and far (cancellation)
    au = W.AU_PC
    c = np.array([0.8, -0.6, 0.4])
    s = np.array([c + [0.5 * au, 0, 0], c - [0.5 * au, 0, 0], [0.0, 0.0, 0.0]])
    t = np.array([c + [0, 3 * au, 0], c + [2e-3, 0, 0], [5.0, 5.0, 5.0], c])
    cases.append(dict(s=s, m=np.array([10.0, 10.0, 1.0]), se=np.zeros(3), t=t, te=np.zeros(4)))
    # 4: two equal masses, target at the midpoint (zero acceleration by symmetry)
    cases.append(dict(s=np.array([[-1.0, 0, 0], [1.0, 0, 0]]), m=np.array([3.0, 3.0]),
                      se=np.array([0.1, 0.1]), t=np.array([[0.0, 0.0, 0.0], [0.0, 2.0, 0.0]]),
                      te=np.zeros(2)))
    return cases


def main():
    arrays = {}
    for i, c in enumerate(kat_cases()):
        s, t = c["s"], c["t"]
        res = brute.eval_plummer(s[:, 0], s[:, 1], s[:, 2], c["m"], c["se"],
                                 t[:, 0], t[:, 1], t[:, 2], c["te"], G)
        exp = np.array([[float(v) for v in r] for r in res])
        arrays.update({f"c{i}_s": s, f"c{i}_m": c["m"], f"c{i}_se": c["se"], f"c{i}_t": t,
                       f"c{i}_te": c["te"], f"c{i}_expected": exp})
    arrays["n_cases"] = np.array(len(kat_cases()))
    arrays["G"] = np.array(G)
    np.savez_compressed(os.path.join(OUT, "plummer_kat.npz"), **arrays)

    w = W.c1(1001)
    st, ce = w["stars"], w["cells"]
    fwd = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"],
                              ce["x"], ce["y"], ce["z"], None, G, mode=oracle.NEUMAIER, threads=8)
    rev = oracle.eval_plummer(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"],
                              st["x"], st["y"], st["z"], None, G, mode=oracle.NEUMAIER, threads=8)
    np.savez_compressed(
        os.path.join(OUT, "c1_oracle.npz"), seed=np.array(1001), G=np.array(G),
        star_x=st["x"], star_y=st["y"], star_z=st["z"], star_m=st["m"], star_eps=st["eps"],
        cell_x=ce["x"], cell_y=ce["y"], cell_z=ce["z"], cell_m=ce["m"], cell_eps=ce["eps"],
        fwd=np.stack(fwd), rev=np.stack(rev))
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
