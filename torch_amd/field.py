"""Host-side mirror of the reference's field-evaluation interface.

The two sides of Torch's gravity bridge expose the direct-sum field under AMUSE
legacy-function names (all FP64 arrays of one length, ``ERROR_CODE`` on failure):

* the stellar side (the AMUSE N-body code, src/torch_mainloop.py:128-140):
  ``get_gravity_at_point(eps, x, y, z) -> (ax, ay, az)`` and
  ``get_potential_at_point(eps, x, y, z) -> phi``
  (spec shape of src/amuse/torch_amuse_flash/interface.py:221-243, 1436-1446);
  sources are the stars -> :class:`StellarGravityField`;
* the gas side (the FLASH worker): ``get_accel_gas_on_particles(eps, x, y, z)`` and
  ``get_potential_at_point(eps, x, y, z)`` (interface.py:221-231, 548-559;
  interface.F90:1073-1216); sources are the leaf-block cells as point masses
  m = rho dV -> :class:`GasGravityField`.

Every evaluation runs on the GPU through libtgrav.so.  Inputs may be numpy arrays (host
pointers; the call is synchronous and returns numpy arrays) or float64 torch tensors on a
ROCm device (device pointers; work is enqueued on torch's current stream and torch
tensors are returned).  Scalars are broadcast.  A non-zero library return code raises
:class:`TgravError`, the way AMUSE raises on a non-zero ERROR_CODE.

Softening: eps2_ij = eps_source_j^2 + eps_i^2 where eps_i is the per-point ``eps``
argument (in the reference's FLASH exports ``eps`` is accepted but unused and a global
radius applies, Gravity_getAccelAtPoint.F90:89-145; here it is honoured).  FLASH's own
spline / linear kernels with that global radius, and its periodic image sum, are
available too (:meth:`GravityField.set_softening`, :meth:`GravityField.set_periodic`).
Units are those of the inputs; the default G = 4.300917270e-3 pc (km/s)^2 / Msun.  AMUSE
quantities (anything with ``value_in``) are accepted too, as the reference's legacy functions
take them (interface.py:1436-1446, 1616-1620): they are read in the field's unit system
(``units=``, :mod:`torch_amd.units`; pc / Msun / km/s unless given) and the results of a call
that received quantities are quantities: accelerations in speed^2 / length, potentials in
speed^2.
"""
import ctypes
import os
import weakref

import numpy as np

from . import _lib
from . import units as _units
from ._lib import DEVICE, HOST, MIXED, PLAIN_FP64, TgravError
from .units import G_PC_MSUN_KMS

MODES = {"fp64": PLAIN_FP64, "plain_fp64": PLAIN_FP64, "mixed": MIXED}


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class _Arrays:
    """Normalises a group of same-length inputs to one backend (numpy host / torch device)."""

    def __init__(self, n, device_like=None):
        self.n = n
        self.dev = device_like
        self.keep = []

    def inp(self, a, name, allow_none=False):
        if a is None:
            if allow_none:
                return None
            raise TgravError(_lib.EINVAL, "argument", f"{name} is required")
        if self.dev is not None:
            import torch
            if not torch.is_tensor(a):
                a = torch.as_tensor(np.asarray(a, dtype=np.float64), device=self.dev)
            a = a.to(device=self.dev, dtype=torch.float64)
            if a.dim() == 0:
                a = a.expand(self.n)
            a = a.contiguous().reshape(-1)
        else:
            a = np.asarray(a, dtype=np.float64)
            if a.ndim == 0:
                a = np.full(self.n, float(a))
            a = np.ascontiguousarray(a).reshape(-1)
        if a.shape[0] != self.n:
            raise TgravError(_lib.EINVAL, "argument",
                             f"{name} has length {a.shape[0]}, expected {self.n}")
        self.keep.append(a)
        return a

    def out(self):
        if self.dev is not None:
            import torch
            a = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        else:
            a = np.empty(self.n, dtype=np.float64)
        self.keep.append(a)
        return a

    @staticmethod
    def ptr(a):
        if a is None:
            return None
        if _is_torch(a):
            return ctypes.c_void_p(a.data_ptr())
        return ctypes.c_void_p(a.ctypes.data)

    @property
    def kind(self):
        return DEVICE if self.dev is not None else HOST


def _length(*arrs):
    for a in arrs:
        if a is not None and np.ndim(a) > 0:
            return int(a.shape[0]) if hasattr(a, "shape") else len(a)
    return 1


def _host(a):
    """Quantities wrap host arrays: device results are copied back for them."""
    return a.cpu().numpy() if a is not None and _is_torch(a) else a


def _torch_device(*arrs):
    for a in arrs:
        if a is not None and _is_torch(a) and a.is_cuda:
            return a.device
    return None


class GravityField:
    """A source set on one GPU and direct-sum evaluations against it.

    mode: "fp64" (plain FP64, matches the FP64 oracle to <= 1e-12) or "mixed"
    (FP64 displacements about a per-tile anchor, FP32 rsqrt/FMA bulk, two-sum
    compensated sums, FP64 result and kick).
    """

    def __init__(self, mode="fp64", device=0, G=None, validate=True, tune_cache="default", units=None):
        """tune_cache: path of an autotuner cache the library applies by itself (launch shapes
        per device and size bucket; shapes without an entry take the built-in heuristics);
        "default" is the package's tune_cache.json (or $TGRAV_TUNE_CACHE); None uses the
        heuristics only.  An explicit set_config / set_bidir_config wins over the cache.
        units: a torch_amd.units.UnitSystem for AMUSE quantities (e.g. units.cgs() for FLASH's
        cm / g / s); G defaults to its G, else to G in pc, Msun, km/s."""
        if mode not in MODES:
            raise ValueError(f"mode must be one of {sorted(MODES)}")
        self.lib = _lib.load()
        self.mode = MODES[mode]
        self.device = device
        ctx = ctypes.c_void_p()
        if isinstance(tune_cache, str) and tune_cache == "default":
            tune_cache = os.environ.get("TGRAV_TUNE_CACHE", os.path.join(os.path.dirname(__file__), "tune_cache.json"))
            if not os.path.exists(tune_cache):
                tune_cache = None
        path = None if tune_cache is None else os.fsencode(tune_cache)
        _lib.check(self.lib.tgrav_create(ctypes.byref(ctx), int(device), path), "tgrav_create")
        self._ctx = ctx
        self.units = units
        if G is None:
            G = units.G if units is not None else G_PC_MSUN_KMS
        self.G = G
        _lib.check(self.lib.tgrav_set_constant(ctx, float(G)), "tgrav_set_constant")
        _lib.check(self.lib.tgrav_set_validation(ctx, int(bool(validate))), "tgrav_set_validation")
        self.n_sources = None
        self._plans = weakref.WeakSet()   # batch plans on this context: closed before it

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        for plan in list(getattr(self, "_plans", ())):
            plan.close()
        if getattr(self, "_ctx", None):
            self.lib.tgrav_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._ctx

    # -- AMUSE quantities -----------------------------------------------------------
    def _unit_system(self):
        if self.units is None:
            if self.G != G_PC_MSUN_KMS:
                raise TgravError(_lib.EINVAL, "argument", "quantities with a custom G need units= "
                                 "(a torch_amd.units.UnitSystem with that G)")
            self.units = _units.pc_msun_kms()
        return self.units

    def _plain(self, *pairs):
        """[(arg, kind)] -> (the args as plain numbers in the field's units, whether any was a
        quantity)."""
        if not any(_units.is_quantity(a) for a, _ in pairs):
            return [a for a, _ in pairs], False
        us = self._unit_system()
        return [us.value(a, kind) for a, kind in pairs], True

    def _quantities(self, outs, kinds, wrap):
        if not wrap:
            return tuple(outs)
        us = self._unit_system()
        return tuple(us.quantity(_host(o), k) for o, k in zip(outs, kinds))

    def _bind_stream(self, arrays):
        """Device tensors: run on torch's current stream (0 is the null stream) so work is
        ordered with the caller's.  Host arrays: the context's own stream (calls sync)."""
        if arrays.kind == DEVICE:
            import torch
            s = torch.cuda.current_stream(arrays.dev).cuda_stream
            _lib.check(self.lib.tgrav_set_stream(self._ctx, ctypes.c_void_p(s or None)),
                       "tgrav_set_stream")
        else:
            _lib.check(self.lib.tgrav_use_own_stream(self._ctx), "tgrav_use_own_stream")

    def set_config(self, variant=-1, nsplit=-1, mode=None):
        m = self.mode if mode is None else MODES[mode]
        _lib.check(self.lib.tgrav_set_config(self._ctx, m, int(variant), int(nsplit)),
                   "tgrav_set_config")

    # -- sources ----------------------------------------------------------------
    def set_sources(self, x, y, z, mass, eps=None):
        (x, y, z, mass, eps), _ = self._plain((x, "length"), (y, "length"), (z, "length"), (mass, "mass"),
                                              (eps, "length"))
        n = _length(x, y, z, mass)
        A = _Arrays(n, _torch_device(x, y, z, mass, eps))
        xs = [A.inp(x, "x"), A.inp(y, "y"), A.inp(z, "z"), A.inp(mass, "mass"),
              A.inp(eps, "eps", allow_none=True)]
        self._bind_stream(A)
        rc = self.lib.tgrav_set_sources(self._ctx, n, *[A.ptr(a) for a in xs], A.kind)
        _lib.check(rc, "tgrav_set_sources")
        self.n_sources = n

    def set_source_blocks(self, origin, dx, mass, nb=8, eps=None):
        """Leaf blocks as the source set: lower corners origin[n,3], cell width dx[n], cell
        masses mass[n*512] (rho dV) in C order, softening eps[n] per block (default: the cell
        width).  Evaluations use the column kernel for 8^3 blocks."""
        (origin, dx, mass, eps), _ = self._plain((origin, "length"), (dx, "length"), (mass, "mass"),
                                                 (eps, "length"))
        nblk = int(origin.shape[0])
        dev = _torch_device(origin, dx, mass, eps)
        A, B = _Arrays(nblk, dev), _Arrays(nblk * nb ** 3, dev)
        o = [A.inp(origin[:, q], "origin") for q in range(3)]
        d = A.inp(dx, "dx")
        e = A.inp(dx if eps is None else eps, "eps")
        m = B.inp(mass, "mass")
        self._bind_stream(A)
        rc = self.lib.tgrav_set_source_blocks(self._ctx, nblk, int(nb), *[A.ptr(q) for q in o],
                                              A.ptr(d), B.ptr(m), A.ptr(e), A.kind)
        _lib.check(rc, "tgrav_set_source_blocks")
        self.n_sources = nblk * nb ** 3

    # -- evaluation ---------------------------------------------------------------
    def evaluate(self, eps, x, y, z, acceleration=True, potential=True):
        """(ax, ay, az, phi) at the points; unrequested outputs are None (quantities when any
        argument was one)."""
        (eps, x, y, z), wrap = self._plain((eps, "length"), (x, "length"), (y, "length"), (z, "length"))
        n = _length(x, y, z, eps)
        A = _Arrays(n, _torch_device(x, y, z, eps))
        e = A.inp(eps, "eps", allow_none=True)
        xs = [A.inp(x, "x"), A.inp(y, "y"), A.inp(z, "z")]
        outs = [A.out() if acceleration else None for _ in range(3)] + \
               [A.out() if potential else None]
        self._bind_stream(A)
        rc = self.lib.tgrav_eval(self._ctx, n, A.ptr(e), *[A.ptr(a) for a in xs],
                                 *[A.ptr(o) for o in outs], self.mode, A.kind)
        _lib.check(rc, "tgrav_eval")
        return self._quantities(outs, ("acceleration",) * 3 + ("potential",), wrap)

    def evaluate_and_kick(self, eps, x, y, z, dt, vx, vy, vz, potential=False):
        """v += dt * a(x) in FP64, fused with the evaluation; returns (ax, ay, az, phi) and
        updates vx/vy/vz in place (numpy float64 arrays or float64 device tensors, in the
        field's speed unit when the other arguments are quantities)."""
        (eps, x, y, z, dt), wrap = self._plain((eps, "length"), (x, "length"), (y, "length"), (z, "length"),
                                               (dt, "time"))
        n = _length(x, y, z, eps)
        A = _Arrays(n, _torch_device(x, y, z, eps, vx))
        e = A.inp(eps, "eps", allow_none=True)
        xs = [A.inp(x, "x"), A.inp(y, "y"), A.inp(z, "z")]
        for v, name in ((vx, "vx"), (vy, "vy"), (vz, "vz")):
            self._check_inplace(v, n, name, A)
        outs = [A.out() for _ in range(3)] + [A.out() if potential else None]
        self._bind_stream(A)
        rc = self.lib.tgrav_eval_kick(self._ctx, n, A.ptr(e), *[A.ptr(a) for a in xs], float(dt),
                                      A.ptr(vx), A.ptr(vy), A.ptr(vz), *[A.ptr(o) for o in outs],
                                      self.mode, A.kind)
        _lib.check(rc, "tgrav_eval_kick")
        return self._quantities(outs, ("acceleration",) * 3 + ("potential",), wrap)

    def evaluate_blocks(self, origin, dx, nb=8, eps=None, potential=True, dt=0.0, vel=None):
        """Field at the nb^3 cell centres of leaf blocks (lower corners origin[n,3], cell
        width dx[n]); outputs in C order b*nb^3 + (i*nb + j)*nb + k.  vel=(vx,vy,vz)
        per-cell arrays are kicked by dt in place."""
        (origin, dx, eps, dt), wrap = self._plain((origin, "length"), (dx, "length"), (eps, "length"),
                                                  (dt, "time"))
        nblk = int(origin.shape[0])
        ncell = nblk * nb ** 3
        A = _Arrays(nblk, _torch_device(origin, dx, eps))
        ox, oy, oz = (A.inp(origin[:, q], "origin") for q in range(3))
        d = A.inp(dx, "dx")
        e = A.inp(eps, "eps", allow_none=True)
        B = _Arrays(ncell, A.dev)
        outs = [B.out() for _ in range(3)] + [B.out() if potential else None]
        v = [None, None, None]
        if vel is not None:
            for q, name in enumerate(("vx", "vy", "vz")):
                self._check_inplace(vel[q], ncell, name, B)
                v[q] = vel[q]
        self._bind_stream(A)
        rc = self.lib.tgrav_eval_blocks(self._ctx, nblk, int(nb), A.ptr(ox), A.ptr(oy), A.ptr(oz),
                                        A.ptr(d), A.ptr(e), *[B.ptr(o) for o in outs], float(dt),
                                        *[B.ptr(q) for q in v], self.mode, A.kind)
        _lib.check(rc, "tgrav_eval_blocks")
        return self._quantities(outs, ("acceleration",) * 3 + ("potential",), wrap)

    def bidirectional(self, cx, cy, cz, cm, ce, px, py, pz, pm, pe, potential=True):
        """One fused pass over all (cell, particle) pairs with eps2 = ce^2 + pe^2
        (Particles_sinkAccelGasOnSinksAndSinksOnGas.F90:204-316): returns
        ((cax, cay, caz, cphi) of the cells from the particles,
         (pax, pay, paz, pphi) of the particles from the cells)."""
        L, M = "length", "mass"
        (cx, cy, cz, cm, ce, px, py, pz, pm, pe), wrap = self._plain(
            (cx, L), (cy, L), (cz, L), (cm, M), (ce, L), (px, L), (py, L), (pz, L), (pm, M), (pe, L))
        nc, np_ = _length(cx, cy, cz, cm), _length(px, py, pz, pm)
        dev = _torch_device(cx, cy, cz, cm, ce, px, py, pz, pm, pe)
        A, B = _Arrays(nc, dev), _Arrays(np_, dev)
        ca = [A.inp(cx, "cx"), A.inp(cy, "cy"), A.inp(cz, "cz"), A.inp(cm, "cm"),
              A.inp(ce, "ce", allow_none=True)]
        pa = [B.inp(px, "px"), B.inp(py, "py"), B.inp(pz, "pz"), B.inp(pm, "pm"),
              B.inp(pe, "pe", allow_none=True)]
        co = [A.out() for _ in range(3)] + [A.out() if potential else None]
        po = [B.out() for _ in range(3)] + [B.out() if potential else None]
        self._bind_stream(A)
        rc = self.lib.tgrav_bidirectional(self._ctx, nc, *[A.ptr(a) for a in ca], np_,
                                          *[B.ptr(a) for a in pa], *[A.ptr(o) for o in co],
                                          *[B.ptr(o) for o in po], self.mode, A.kind)
        _lib.check(rc, "tgrav_bidirectional")
        kinds = ("acceleration",) * 3 + ("potential",)
        return self._quantities(co, kinds, wrap), self._quantities(po, kinds, wrap)

    def bidirectional_blocks(self, origin, dx, cm, ceps, px, py, pz, pm, pe, potential=True):
        """The fused pass with the gas cells as 8^3 leaf blocks: lower corners origin[n, 3], cell
        widths dx[n], cell masses cm[n*512] (C order), softening ceps[n] per block (None = 0).
        Returns the same as :meth:`bidirectional`, the cell outputs in the blocks' C order."""
        L, M = "length", "mass"
        (origin, dx, cm, ceps, px, py, pz, pm, pe), wrap = self._plain(
            (origin, L), (dx, L), (cm, M), (ceps, L), (px, L), (py, L), (pz, L), (pm, M), (pe, L))
        nblk = int(origin.shape[0])
        np_ = _length(px, py, pz, pm)
        dev = _torch_device(origin, dx, cm, ceps, px, py, pz, pm, pe)
        A, M, B = _Arrays(nblk, dev), _Arrays(nblk * 512, dev), _Arrays(np_, dev)
        ba = [A.inp(origin[:, q], "origin") for q in range(3)] + [A.inp(dx, "dx")]
        m = M.inp(cm, "cm")
        e = A.inp(ceps, "ceps", allow_none=True)
        pa = [B.inp(px, "px"), B.inp(py, "py"), B.inp(pz, "pz"), B.inp(pm, "pm"),
              B.inp(pe, "pe", allow_none=True)]
        co = [M.out() for _ in range(3)] + [M.out() if potential else None]
        po = [B.out() for _ in range(3)] + [B.out() if potential else None]
        self._bind_stream(A)
        rc = self.lib.tgrav_bidirectional_blocks(self._ctx, nblk, 8, *[A.ptr(a) for a in ba], M.ptr(m),
                                                 A.ptr(e), np_, *[B.ptr(a) for a in pa],
                                                 *[M.ptr(o) for o in co], *[B.ptr(o) for o in po],
                                                 self.mode, A.kind)
        _lib.check(rc, "tgrav_bidirectional_blocks")
        kinds = ("acceleration",) * 3 + ("potential",)
        return self._quantities(co, kinds, wrap), self._quantities(po, kinds, wrap)

    SOFTENING = {"plummer": 0, "spline": 1, "linear": 2}

    def set_softening(self, kernel="plummer", rsoft=0.0, quirks=False):
        """Pair kernel of later evaluations: "plummer" (default; per-point softening) or
        FLASH's "spline" / "linear" kernels with the global radius `rsoft`
        (sink_softening_radius, Gravity_getAccelAtPoint.F90:88-100,140-145), FP64 mode only.
        quirks=True reproduces the reference's potential inside rsoft as written."""
        if kernel not in self.SOFTENING:
            raise ValueError(f"kernel must be one of {sorted(self.SOFTENING)}")
        (rsoft,), _ = self._plain((rsoft, "length"))
        _lib.check(self.lib.tgrav_set_softening(self._ctx, self.SOFTENING[kernel], float(rsoft),
                                                int(bool(quirks))), "tgrav_set_softening")

    def set_periodic(self, box=None, nrep=6, per_target=False):
        """Periodic image sum of the FLASH kernels over n in [-nrep, nrep]^3 within
        nrep * min(box) (Gravity_getAccelAtPoint.F90:250-312); box=None: isolated.
        per_target=True keeps the per-target image walk instead of the pairs grouped by
        separation (the default where nrep <= 10 and there are at most 2^25 cells; the targets
        go in batches of at most 2^25 pairs)."""
        if box is None:
            _lib.check(self.lib.tgrav_set_periodic(self._ctx, 0, 0.0, 0.0, 0.0, 0), "tgrav_set_periodic")
            return
        lx, ly, lz = (float(self._plain((b, "length"))[0][0]) for b in box)
        _lib.check(self.lib.tgrav_set_periodic(self._ctx, 2 if per_target else 1, lx, ly, lz, int(nrep)),
                   "tgrav_set_periodic")

    def set_block_kernel(self, on=True):
        """Use (default) or bypass the 8^3 column kernels: block-structured targets and sources,
        and explicit arrays recognised as leaf blocks in C order."""
        _lib.check(self.lib.tgrav_set_block_kernel(self._ctx, int(bool(on))), "tgrav_set_block_kernel")

    def set_column_waves(self, waves):
        """Waves per workgroup of the 8^3 column kernels (1, 2 or 4; default 4): the waves of a workgroup
        share one leaf block and split its sources, so fewer source chunks (and partial sums in
        memory) give the same parallelism.  Batch plans use the value set when they are created."""
        _lib.check(self.lib.tgrav_set_column_waves(self._ctx, int(waves)), "tgrav_set_column_waves")

    def column_waves(self):
        return int(self.lib.tgrav_get_column_waves(self._ctx))

    def set_bidir_config(self, variant=-1, nsplit=-1):
        _lib.check(self.lib.tgrav_set_bidir_config(self._ctx, self.mode, int(variant), int(nsplit)),
                   "tgrav_set_bidir_config")

    def kick(self, dt, ax, ay, az, vx, vy, vz):
        """v += dt * a in FP64, in place (the bridge passes dt/2)."""
        n = _length(ax)
        A = _Arrays(n, _torch_device(ax, vx))
        a = [A.inp(ax, "ax"), A.inp(ay, "ay"), A.inp(az, "az")]
        for v, name in ((vx, "vx"), (vy, "vy"), (vz, "vz")):
            self._check_inplace(v, n, name, A)
        self._bind_stream(A)
        rc = self.lib.tgrav_kick(self._ctx, n, float(dt), *[A.ptr(q) for q in a],
                                 A.ptr(vx), A.ptr(vy), A.ptr(vz), A.kind)
        _lib.check(rc, "tgrav_kick")

    @staticmethod
    def _check_inplace(v, n, name, A):
        if A.kind == DEVICE:
            import torch
            ok = torch.is_tensor(v) and v.dtype == torch.float64 and v.is_contiguous() \
                and v.device == A.dev and v.numel() == n
        else:
            ok = isinstance(v, np.ndarray) and v.dtype == np.float64 and \
                v.flags.c_contiguous and v.size == n and v.flags.writeable
        if not ok:
            raise TgravError(_lib.EINVAL, "argument",
                             f"{name} must be a writable contiguous float64 array of length {n} "
                             "on the same device as the positions")

    def set_timing(self, on=True):
        """Record HIP events around every evaluation kernel launch (instrumentation)."""
        _lib.check(self.lib.tgrav_set_timing(self._ctx, int(bool(on))), "tgrav_set_timing")

    def kernel_times(self, max_records=256):
        """[(ms, n_targets, n_sources)] of the launches timed since set_timing/last read."""
        ms = (ctypes.c_double * max_records)()
        nt = (ctypes.c_int64 * max_records)()
        ns = (ctypes.c_int64 * max_records)()
        n = self.lib.tgrav_timing_read(self._ctx, max_records, ms, nt, ns)
        if n < 0:
            _lib.check(n, "tgrav_timing_read")
        return [(ms[i], nt[i], ns[i]) for i in range(min(n, max_records))]

    def synchronize(self):
        _lib.check(self.lib.tgrav_synchronize(self._ctx), "tgrav_synchronize")


class StellarGravityField(GravityField):
    """Stars as sources: the stellar side's AMUSE GravityFieldInterface methods."""

    def set_stars(self, x, y, z, mass, eps=None):
        self.set_sources(x, y, z, mass, eps)

    def get_gravity_at_point(self, eps, x, y, z):
        ax, ay, az, _ = self.evaluate(eps, x, y, z, potential=False)
        return ax, ay, az

    def get_potential_at_point(self, eps, x, y, z):
        return self.evaluate(eps, x, y, z, acceleration=False)[3]


class GasGravityField(GravityField):
    """Gas cells as point masses (m = rho dV, Gravity_getAccelAtPoint.F90:188-204) as
    sources: the FLASH worker's exports."""

    def set_cells(self, x, y, z, rho, dx, eps=None):
        """Cells with centre (x,y,z), density rho and width dx; softening eps (default:
        the cell width)."""
        (x, y, z, rho, dx, eps), _ = self._plain((x, "length"), (y, "length"), (z, "length"),
                                                 (rho, "density"), (dx, "length"), (eps, "length"))
        mass = rho * dx * dx * dx
        self.set_sources(x, y, z, mass, dx if eps is None else eps)

    def set_blocks(self, origin, dx, mass, nb=8, eps=None):
        """Leaf blocks as the source set (see GravityField.set_source_blocks); softening
        defaults to the cell width, as for set_cells."""
        self.set_source_blocks(origin, dx, mass, nb, eps)

    def get_accel_gas_on_particles(self, eps, x, y, z):
        ax, ay, az, _ = self.evaluate(eps, x, y, z, potential=False)
        return ax, ay, az

    def get_potential_at_point(self, eps, x, y, z):
        return self.evaluate(eps, x, y, z, acceleration=False)[3]
