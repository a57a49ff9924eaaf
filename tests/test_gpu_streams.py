"""Stream semantics with device tensors (needs an MI355X): calls are ordered on torch's current
stream, so inputs produced and outputs consumed on a side stream need no synchronisation, and
two contexts on two streams run independently.  Results are compared bit for bit with the
same evaluation from host arrays (same kernels and launch shapes)."""
import numpy as np
import pytest

from torch_amd import GravityField
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    rng = W.rng_for(71)
    st = W.star_cluster(rng, 3000)
    t = rng.uniform(-2.0, 2.0, (3, 50_000))
    return st, t


def busy(stream, n=40):
    """Queue a few milliseconds of elementwise work on `stream` (so a mis-ordered call would
    read inputs that are not written yet)."""
    import torch
    with torch.cuda.stream(stream):
        a = torch.ones(1 << 24, device="cuda")
        for _ in range(n):
            a.mul_(1.0000001).add_(1e-9)
    return a


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_side_stream_orders_inputs_and_outputs(problem, mode):
    import torch
    st, t = problem
    S = [st[k] for k in ("x", "y", "z", "m", "eps")]
    with GravityField(mode) as f:
        f.set_sources(*S)
        ref = f.evaluate(None, t[0], t[1], t[2])
        base = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in list(t) + S]
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        keep = busy(s)
        with torch.cuda.stream(s):
            xs = [b.clone() for b in base]            # written on s, after the busy work
            f.set_sources(*xs[3:])
            out = f.evaluate(None, xs[0], xs[1], xs[2])
            got = [o.clone() for o in out]            # read on s, before any synchronisation
        s.synchronize()
        del keep
    for p, q in zip(got, ref):
        np.testing.assert_array_equal(p.cpu().numpy(), q)


def test_two_contexts_on_two_streams(problem):
    import torch
    st, t = problem
    S = [st[k] for k in ("x", "y", "z", "m", "eps")]
    half = t.shape[1] // 2
    with GravityField("fp64") as f, GravityField("mixed") as g:
        f.set_sources(*S)
        g.set_sources(*S)
        ref_f = f.evaluate(None, t[0][:half], t[1][:half], t[2][:half])
        ref_g = g.evaluate(None, t[0][half:], t[1][half:], t[2][half:])
        d = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in t]
        dS = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in S]
        torch.cuda.synchronize()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        keep = busy(s1, 10)
        with torch.cuda.stream(s1):
            f.set_sources(*dS)
            of = f.evaluate(None, d[0][:half], d[1][:half], d[2][:half])
        with torch.cuda.stream(s2):
            g.set_sources(*dS)
            og = g.evaluate(None, d[0][half:], d[1][half:], d[2][half:])
        torch.cuda.synchronize()
        del keep
    for p, q in zip(of, ref_f):
        np.testing.assert_array_equal(p.cpu().numpy(), q)
    for p, q in zip(og, ref_g):
        np.testing.assert_array_equal(p.cpu().numpy(), q)


def test_contexts_on_two_host_threads(problem):
    """Two host threads, each with its own context, evaluating at the same time (the library
    keeps no mutable global state; ctypes releases the GIL during the calls): every result is
    bit-identical to the same evaluation run alone."""
    import threading
    st, t = problem
    S = [st[k] for k in ("x", "y", "z", "m", "eps")]
    half = t.shape[1] // 2
    parts = {"fp64": (t[0][:half], t[1][:half], t[2][:half]),
             "mixed": (t[0][half:], t[1][half:], t[2][half:])}
    ref = {}
    for mode, (x, y, z) in parts.items():
        with GravityField(mode) as f:
            f.set_sources(*S)
            ref[mode] = f.evaluate(None, x, y, z)
    got, errors = {mode: [] for mode in parts}, []

    def work(mode):
        try:
            x, y, z = parts[mode]
            with GravityField(mode) as f:
                for _ in range(6):
                    f.set_sources(*S)
                    got[mode].append(f.evaluate(None, x, y, z))
        except Exception as e:          # surfaced below, in the test's thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(mode,)) for mode in parts]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for mode in parts:
        assert len(got[mode]) == 6
        for res in got[mode]:
            for p, q in zip(res, ref[mode]):
                np.testing.assert_array_equal(p, q)
