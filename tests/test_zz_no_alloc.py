"""The allocation-free entry point tgrav_eval_static against tgrav_eval (needs an MI355X).

tgrav_eval_static takes the source set with the call and keeps the packed sources and the chunk
partial sums in a fixed per-device array instead of allocated buffers; it runs tgrav_eval's
generic kernels with the same launch shape, so on the same inputs the two give the same bits (a
stronger bar than the 1e-12 FP64 / mixed-gate tolerance, checked as well).  Two calls in a row,
the second smaller, show that the shared array carries nothing stale from the first.  Sizes
beyond the array are refused.  Called through the C ABI (ctypes), device pointers.
"""
import ctypes

import numpy as np
import pytest

from tests import _util
from torch_amd import _lib
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


@pytest.mark.parametrize("mode", [_lib.PLAIN_FP64, _lib.MIXED])
def test_static_entry_point_matches_eval(mode):
    import torch
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    _lib.check(lib.tgrav_create(ctypes.byref(ctx), 0, None), "tgrav_create")
    try:
        rng = W.rng_for(90 + mode)
        for ns, nt in ((5000, 30000), (700, 1234)):
            print(f"no-alloc: mode {mode}, {ns} sources -> {nt} targets", flush=True)
            st = W.star_cluster(rng, ns)
            d = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")  # noqa
            S = [d(st[k]) for k in ("x", "y", "z", "m")] + [d(rng.uniform(0.005, 0.05, ns))]
            t = rng.uniform(-2.0, 2.0, (3, nt))
            T = [d(t[0]), d(t[1]), d(t[2])]
            et = d(rng.uniform(0.0, 0.01, nt))
            ref = [torch.empty(nt, dtype=torch.float64, device="cuda") for _ in range(4)]
            got = [torch.full((nt,), np.nan, dtype=torch.float64, device="cuda") for _ in range(4)]
            torch.cuda.synchronize()
            _lib.check(lib.tgrav_set_sources(ctx, ns, *[_p(a) for a in S], _lib.DEVICE), "tgrav_set_sources")
            _lib.check(lib.tgrav_eval(ctx, nt, _p(et), *[_p(a) for a in T], *[_p(a) for a in ref], mode,
                                      _lib.DEVICE), "tgrav_eval")
            _lib.check(lib.tgrav_eval_static(ctx, ns, *[_p(a) for a in S], nt, _p(et), *[_p(a) for a in T],
                                             *[_p(a) for a in got], mode), "tgrav_eval_static")
            _lib.check(lib.tgrav_synchronize(ctx), "tgrav_synchronize")
            g = tuple(a.cpu().numpy() for a in got)
            r = tuple(a.cpu().numpy() for a in ref)
            s = _util.summary(g, r)
            if mode == _lib.PLAIN_FP64:
                assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
            else:
                assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, s
            for q in range(4):
                assert np.array_equal(g[q], r[q]), (ns, nt, q)
        # beyond the fixed array: refused before any device work
        big = 2_000_000
        B = [torch.zeros(big, dtype=torch.float64, device="cuda") for _ in range(4)]
        rc = lib.tgrav_eval_static(ctx, big, *[_p(a) for a in B], None, 10, None, *[_p(a) for a in B[:3]],
                                   *[_p(a) for a in B], mode)
        assert rc == _lib.ELIMIT
        assert b"scratch" in lib.tgrav_last_error()
    finally:
        lib.tgrav_destroy(ctx)
