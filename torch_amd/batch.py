"""Batched uneven evaluations: many independent (sources, targets) problems in one
persistent launch, optionally replayed from a captured hipGraph.

This is the batching form of the field evaluation: differing star subsets or AMR
leaf-block target sets (the per-block calls of .deprecated/bridge_multiples.py:550-596,
thousands per bridge kick in the reference) become one plan that is built once and then
run (or graph-replayed) every kick.  Problems are given as dicts of float64 device
tensors; numpy arrays are copied to the device once, at plan creation.  Outputs are
device tensors owned by the plan.  With dt != 0 the plan also applies the FP64 kick
v += dt * a to each problem's velocity tensors in place.

The plan keeps pointers to its inputs: after changing positions/masses in place, run()
again (the graph re-reads them); do not reallocate them.
"""
import ctypes

import numpy as np

from . import _lib
from .field import GravityField, MODES


def _dev(a, device):
    import torch
    if a is None:
        return None
    if not torch.is_tensor(a):
        a = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
    if a.dtype != torch.float64 or not a.is_contiguous() or a.device != torch.device(device):
        raise _lib.TgravError(_lib.EINVAL, "BatchPlan",
                              "inputs must be contiguous float64 tensors on the plan's device")
    return a


class BatchPlan:
    """problems: list of dict(sources=dict(x,y,z,m[,eps]), targets=dict(x,y,z[,eps])
    [, velocities=(vx,vy,vz)]).  Targets with block_origin[n, 3] and block_dx[n] (optional
    block_eps[n]) are 8^3 leaf blocks: the plan evaluates them with the column kernel, outputs
    in the blocks' C order; their x, y, z are not read.  Sources with block_origin[n, 3],
    block_dx[n] (optional block_eps[n]) and m[512 n] (cell masses rho dV in the blocks' C order)
    are 8^3 leaf blocks, walked block by block and column by column; their x, y, z are not read.
    Explicit targets in the blocks' C order (each aligned 8 share x and y, each aligned 512 share
    eps) take the column kernel too: the plan tests that order on the device at every run, so
    targets updated in place between runs always get a kernel that fits them
    (field.set_block_kernel(False) before creating the plan turns the column kernels off).
    field: a GravityField whose context (device, stream, G, kernel configuration) the plan uses."""

    def __init__(self, field: GravityField, problems, dt=0.0, potential=True):
        import torch
        self.field = field
        self.lib = field.lib
        self.device = torch.device("cuda", field.device)
        self.keep = []
        self.outputs = []
        arr = (_lib.Problem * max(len(problems), 1))()
        for i, p in enumerate(problems):
            S, T = p["sources"], p["targets"]
            sblk = S.get("block_origin")
            if sblk is not None:     # 8^3 source blocks: corners, cell widths, eps per block, cell masses
                so = _dev(sblk, self.device).reshape(-1, 3)
                sb = [so[:, 0].contiguous(), so[:, 1].contiguous(), so[:, 2].contiguous(),
                      _dev(S["block_dx"], self.device), _dev(S.get("block_eps"), self.device)]
                s = [None, None, None, _dev(S["m"], self.device), None]
                ns = 512 * int(sb[3].numel())
            else:
                sb = [None] * 5
                s = [_dev(S[k], self.device) for k in ("x", "y", "z", "m")] + [_dev(S.get("eps"), self.device)]
                ns = int(s[0].numel())
            blk = T.get("block_origin")
            if blk is not None:      # 8^3 leaf blocks: corners, cell widths, optional eps per block
                o = _dev(blk, self.device).reshape(-1, 3)
                tb = [o[:, 0].contiguous(), o[:, 1].contiguous(), o[:, 2].contiguous(),
                      _dev(T["block_dx"], self.device), _dev(T.get("block_eps"), self.device)]
                t = [None, None, None, None]
                nt = 512 * int(tb[3].numel())
            else:
                tb = [None] * 5
                t = [_dev(T[k], self.device) for k in ("x", "y", "z")] + [_dev(T.get("eps"), self.device)]
                nt = int(t[0].numel())
            out = [torch.empty(nt, dtype=torch.float64, device=self.device) for _ in range(3)]
            out.append(torch.empty(nt, dtype=torch.float64, device=self.device) if potential else None)
            v = [None, None, None]
            if dt != 0.0 and p.get("velocities") is not None:
                v = [_dev(q, self.device) for q in p["velocities"]]
            self.keep += s + sb + t + tb + out + v
            self.outputs.append(tuple(out))
            P = arr[i]
            P.n_sources = ns
            P.sx, P.sy, P.sz, P.sm, P.se = [(a.data_ptr() if a is not None else None) for a in s]
            P.n_targets = nt
            P.tx, P.ty, P.tz, P.te = [(a.data_ptr() if a is not None else None) for a in t]
            P.n_target_blocks = nt // 512 if blk is not None else 0
            P.tbx, P.tby, P.tbz, P.tbd, P.tbe = [(a.data_ptr() if a is not None else None) for a in tb]
            P.n_source_blocks = ns // 512 if sblk is not None else 0
            P.sbx, P.sby, P.sbz, P.sbd, P.sbe = [(a.data_ptr() if a is not None else None) for a in sb]
            P.ax, P.ay, P.az, P.phi = [(a.data_ptr() if a is not None else None) for a in out]
            P.vx, P.vy, P.vz = [(a.data_ptr() if a is not None else None) for a in v]
        self.velocities = [p.get("velocities") for p in problems] if dt != 0.0 else None
        torch.cuda.synchronize(self.device)
        h = ctypes.c_void_p()
        rc = self.lib.tgrav_batch_create(field.handle, len(problems), arr, field.mode, float(dt),
                                         ctypes.byref(h))
        _lib.check(rc, "tgrav_batch_create")
        self._h = h
        field._plans.add(self)   # the field closes its plans before its context
        n_items, inter = ctypes.c_int64(), ctypes.c_double()
        _lib.check(self.lib.tgrav_batch_info(h, ctypes.byref(n_items), ctypes.byref(inter)),
                   "tgrav_batch_info")
        self.n_items = n_items.value
        self.interactions = inter.value
        pbytes = ctypes.c_int64()
        _lib.check(self.lib.tgrav_batch_partials(h, ctypes.byref(pbytes)), "tgrav_batch_partials")
        self.partial_bytes = pbytes.value   # chunk partial sums written per run

    def run(self, graph=True):
        """Enqueue the plan on torch's current stream (graph=True: captured once, then
        replayed).  Returns the per-problem (ax, ay, az, phi) device tensors."""
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib.tgrav_set_stream(self.field.handle, ctypes.c_void_p(s or None)),
                   "tgrav_set_stream")
        _lib.check(self.lib.tgrav_batch_run(self._h, int(bool(graph))), "tgrav_batch_run")
        return self.outputs

    def close(self):
        if getattr(self, "_h", None):
            self.lib.tgrav_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def problems_from_workload(probs, device="cuda"):
    """Moves a list of workload problems (numpy dicts) to device tensors."""
    import torch
    out = []
    for p in probs:
        d = {}
        for side in ("sources", "targets"):
            d[side] = {k: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=device)
                       for k, v in p[side].items()}
        out.append(d)
    return out


__all__ = ["BatchPlan", "problems_from_workload", "MODES"]
