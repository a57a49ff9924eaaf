"""Synthetic workload generators (CPU): determinism and the distributions they promise."""
import numpy as np

from torch_amd import workloads as W
from torch_amd.blocks import block_centres


def test_seeded_and_reproducible():
    a = W.c1(1001)
    b = W.c1(1001)
    for k in ("x", "y", "z", "m", "vx"):
        assert np.array_equal(a["stars"][k], b["stars"][k])
    assert np.array_equal(a["cells"]["m"], b["cells"]["m"])
    c = W.c1(1002)
    assert not np.array_equal(a["stars"]["x"], c["stars"]["x"])


def test_kroupa_imf_segments():
    rng = W.rng_for(5)
    m = W.kroupa_masses(rng, 400_000)
    assert m.min() >= 0.08 and m.max() <= 100.0
    # log-log slopes of the two segments: dN/dm ~ m^-1.3 and m^-2.3 (src/imf_sample.py:144-155)
    for lo, hi, alpha in ((0.1, 0.4, 1.3), (1.0, 20.0, 2.3)):
        edges = np.geomspace(lo, hi, 9)
        cnt, _ = np.histogram(m, edges)
        dens = cnt / np.diff(edges)
        mid = np.sqrt(edges[1:] * edges[:-1])
        slope = np.polyfit(np.log(mid), np.log(dens), 1)[0]
        assert abs(slope + alpha) < 0.08, (alpha, slope)


def test_plummer_half_mass_radius():
    rng = W.rng_for(6)
    pos, b = W.plummer_positions(rng, 200_000, r_vir=1.0)
    r = np.linalg.norm(pos, axis=1)
    assert abs(b - 3 * np.pi / 16) < 1e-15
    # half-mass radius of a Plummer sphere: b / sqrt(2^(2/3) - 1) ~ 1.305 b (of the truncated
    # 0.999 mass fraction: median of r at mass fraction 0.4995)
    rh = b / np.sqrt(0.4995 ** (-2 / 3) - 1)
    assert abs(np.median(r) / rh - 1) < 0.01


def test_binaries():
    rng = W.rng_for(7)
    st = W.star_cluster(rng, 10_000, binary_frac=0.2)
    assert (st["binary"] >= 0).sum() == 2000
    pairs = st["binary"][st["binary"] >= 0]
    assert np.all(np.bincount(pairs) == 2)
    mem = np.where(st["binary"] >= 0)[0]
    order = mem[np.argsort(st["binary"][mem], kind="stable")]
    p1, p2 = order[0::2], order[1::2]
    sep = np.sqrt((st["x"][p1] - st["x"][p2]) ** 2 + (st["y"][p1] - st["y"][p2]) ** 2 +
                  (st["z"][p1] - st["z"][p2]) ** 2) / W.AU_PC
    assert sep.min() >= 0.0 and sep.max() <= 200.0 + 1e-9     # r <= a (1+e) <= 2 a_max
    assert np.median(sep) > 3 and np.median(sep) < 60


def test_c2_shapes_and_block_layout():
    w = W.c2(1002)
    st, ce = w["stars"], w["cells"]
    assert len(st["x"]) == 10_000 and len(ce["x"]) == 1_024_000
    assert ce["block_origin"].shape == (2000, 3)
    assert set(np.unique(ce["block_size"])) <= {0.5 / 2 ** k for k in range(5)}
    x, y, z, dx = block_centres(ce["block_origin"], ce["block_size"] / 8)
    assert np.array_equal(x, ce["x"]) and np.array_equal(z, ce["z"])
    assert np.all(ce["m"] > 0) and np.array_equal(ce["eps"], ce["dx"])
    # cells lie inside the 4 pc box and inside their own blocks
    assert np.all(np.abs(ce["x"]) < 2) and np.all(np.abs(ce["y"]) < 2)


def test_c4_sizes():
    probs = W.c4(1004)
    assert len(probs) == 64
    for p in probs:
        ns, nt = len(p["sources"]["x"]), len(p["targets"]["x"])
        assert 100 <= ns <= 50_000 and 1000 <= nt <= 1_000_448 and nt % 512 == 0
        assert np.all((p["sources"]["eps"] >= 0.005) & (p["sources"]["eps"] <= 0.1))
