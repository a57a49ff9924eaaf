"""Particle-mesh pieces beside the bridge kick, on FLASH leaf blocks with guard cells.

The reference's production bridge kick takes the particle-mesh route
(src/amuse/torch_amuse_flash/interface.F90:1547-1907 and :1231-1533): particle mass is
mapped to the grid (Particles_updateGridVar -> Grid_mapParticlesToMesh, called at
Gravity_potentialListOfBlocksPDENonly.F90:260-266), a Poisson solve gives the potential,
finite differences give the acceleration (Gravity_accelListOfBlocks, interface.F90:1767-1769),
the gas is kicked with it (interface.F90:1777-1830) and the acceleration is mapped back to
the particles (Particles_longRangeForce.F90:70-97).  :class:`ParticleMesh` runs those pieces
(not the Poisson solve) on the GPU through libtgrav.so, with the reference's method names:

    update_grid_var          CIC mass deposit as density (deterministic, fixed-point sums)
    accel_list_of_blocks     finite-difference acceleration, optionally fused with the kick
    map_mesh_to_particles    CIC interpolation of a block field to particles
    long_range_force         accel_list_of_blocks + map_mesh_to_particles

Blocks (:class:`LeafBlocks`): nblk cubic leaf blocks with interior lower corners
origin[nblk, 3], physical sizes size[nblk], nb interior cells and ng guard layers per axis.
A block variable is an array of nblk * (nb + 2 ng)^3 values, block-major, x fastest
(solnData(VAR, i, j, k) of one variable).  Particles carry their block index blk (0-based).
Inputs may be numpy arrays (synchronous) or ROCm torch tensors (torch's current stream).
"""
import ctypes
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import DEVICE, HOST
from .field import GravityField, _is_torch


class LeafBlocks:
    def __init__(self, origin, size, nb=8, ng=4):
        self.origin = origin
        self.size = size
        self.nb = int(nb)
        self.ng = int(ng)
        self.nblk = int(size.shape[0])
        self.nbg = self.nb + 2 * self.ng

    @property
    def ncell(self):
        return self.nblk * self.nbg ** 3

    def centres(self):
        """Guarded cell centres (x, y, z) in block layout (numpy; for tests and set-up)."""
        o = np.asarray(_host(self.origin), dtype=np.float64).reshape(-1, 3)
        d = np.asarray(_host(self.size), dtype=np.float64) / self.nb
        idx = np.arange(self.nbg) - self.ng + 0.5
        out = []
        for a in range(3):
            shape = [1, 1, 1]
            shape[2 - a] = self.nbg                     # x is the last (fastest) axis
            c = o[:, a][:, None, None, None] + idx.reshape(shape)[None] * d[:, None, None, None]
            out.append(np.broadcast_to(c, (self.nblk, self.nbg, self.nbg, self.nbg)).ravel())
        return tuple(out)


_Where = namedtuple("_Where", "kind dev")   # what GravityField._bind_stream reads


def _host(a):
    return a.detach().cpu().numpy() if _is_torch(a) else a


class ParticleMesh(GravityField):
    """Particle-mesh kernels on one GPU (FP64)."""

    def __init__(self, device=0, validate=True):
        super().__init__("fp64", device=device, validate=validate)

    # -- array plumbing ---------------------------------------------------------------
    def _prep(self, dev, arrs, n, name, dtype=np.float64):
        out = []
        for a, nm in zip(arrs, name):
            if a is None:
                out.append(None)
                continue
            if dev is not None:
                import torch
                t = a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a), device=dev)
                t = t.to(device=dev, dtype=torch.int64 if dtype == np.int64 else torch.float64)
                t = t.contiguous().reshape(-1)
            else:
                t = np.ascontiguousarray(np.asarray(a, dtype=dtype)).reshape(-1)
            if t.shape[0] != n:
                raise _lib.TgravError(_lib.EINVAL, "argument", f"{nm} has length {t.shape[0]}, expected {n}")
            out.append(t)
        return out

    @staticmethod
    def _dev(*arrs):
        for a in arrs:
            if a is not None and _is_torch(a) and a.is_cuda:
                return a.device
        return None

    @staticmethod
    def _ptr(a):
        if a is None:
            return None
        return ctypes.c_void_p(a.data_ptr() if _is_torch(a) else a.ctypes.data)

    def _empty(self, dev, n):
        if dev is not None:
            import torch
            return torch.empty(n, dtype=torch.float64, device=dev)
        return np.empty(n, dtype=np.float64)

    def _stream(self, dev):
        """Bind the context to torch's current stream (device arrays) or its own stream (host
        arrays); returns the pointer kind for the call."""
        where = _Where(DEVICE if dev is not None else HOST, dev)
        self._bind_stream(where)
        return where.kind

    def _geom(self, blocks, dev):
        o = blocks.origin
        o = [o[:, q] for q in range(3)]
        return self._prep(dev, o + [blocks.size], blocks.nblk, ["origin x", "origin y", "origin z", "size"])

    # -- the reference's operations ---------------------------------------------------
    def update_grid_var(self, x, y, z, m, blk, blocks, rho=None, accumulate=False):
        """CIC density rho = sum m w / dV on the guarded blocks; with accumulate=True added to
        `rho` in place, otherwise a new array (rho, if given, is overwritten)."""
        n = len(x) if not _is_torch(x) else int(x.shape[0])
        dev = self._dev(x, y, z, m, blk, blocks.origin, blocks.size, rho)
        xs = self._prep(dev, [x, y, z, m], n, ["x", "y", "z", "m"])
        b = self._prep(dev, [blk], n, ["blk"], dtype=np.int64)[0]
        g = self._geom(blocks, dev)
        if rho is None:
            rho = self._empty(dev, blocks.ncell)
            accumulate = False
        else:
            self._check_out(rho, blocks.ncell, dev, "rho")
        kind = self._stream(dev)
        rc = self.lib.tgrav_pm_deposit(self._ctx, n, *[self._ptr(a) for a in xs], self._ptr(b), blocks.nblk,
                                       blocks.nb, blocks.ng, *[self._ptr(a) for a in g], self._ptr(rho),
                                       int(bool(accumulate)), kind)
        _lib.check(rc, "tgrav_pm_deposit")
        return rho

    def accel_list_of_blocks(self, phi, blocks, dt=0.0, vel=None, store=True):
        """g = -grad(phi) by FLASH's row rule on every guarded cell; vel=(vx, vy, vz) (block
        arrays, updated in place) also gets the grid kick v += g dt in the same pass.
        Returns (gx, gy, gz), or None with store=False."""
        dev = self._dev(phi, blocks.size, *(vel or ()))
        f = self._prep(dev, [phi], blocks.ncell, ["phi"])[0]
        bs = self._prep(dev, [blocks.size], blocks.nblk, ["size"])[0]
        gs = [self._empty(dev, blocks.ncell) for _ in range(3)] if store else [None] * 3
        vs = [None] * 3
        if vel is not None:
            for v, nm in zip(vel, ("vx", "vy", "vz")):
                self._check_out(v, blocks.ncell, dev, nm)
            vs = list(vel)
        kind = self._stream(dev)
        rc = self.lib.tgrav_pm_gradient(self._ctx, blocks.nblk, blocks.nb, blocks.ng, self._ptr(bs), self._ptr(f),
                                        *[self._ptr(a) for a in gs], float(dt), *[self._ptr(a) for a in vs], kind)
        _lib.check(rc, "tgrav_pm_gradient")
        return tuple(gs) if store else None

    def map_mesh_to_particles(self, x, y, z, blk, blocks, gx, gy, gz):
        """CIC interpolation of the block field (gx, gy, gz) to the particles."""
        n = len(x) if not _is_torch(x) else int(x.shape[0])
        dev = self._dev(x, y, z, blk, gx, blocks.origin)
        xs = self._prep(dev, [x, y, z], n, ["x", "y", "z"])
        b = self._prep(dev, [blk], n, ["blk"], dtype=np.int64)[0]
        g = self._geom(blocks, dev)
        fs = self._prep(dev, [gx, gy, gz], blocks.ncell, ["gx", "gy", "gz"])
        outs = [self._empty(dev, n) for _ in range(3)]
        kind = self._stream(dev)
        rc = self.lib.tgrav_pm_interpolate(self._ctx, n, *[self._ptr(a) for a in xs], self._ptr(b), blocks.nblk,
                                           blocks.nb, blocks.ng, *[self._ptr(a) for a in g],
                                           *[self._ptr(a) for a in fs], *[self._ptr(a) for a in outs], kind)
        _lib.check(rc, "tgrav_pm_interpolate")
        return tuple(outs)

    def long_range_force(self, x, y, z, blk, blocks, phi):
        """Particle accelerations from the block potential: finite differences, then CIC
        (Particles_longRangeForce.F90:70-97)."""
        g = self.accel_list_of_blocks(phi, blocks)
        return self.map_mesh_to_particles(x, y, z, blk, blocks, *g)

    @staticmethod
    def _check_out(a, n, dev, name):
        if _is_torch(a):
            import torch
            if a.dtype != torch.float64 or not a.is_contiguous() or a.numel() != n or a.device != dev:
                raise _lib.TgravError(_lib.EINVAL, "argument",
                                      f"{name} must be a contiguous float64 tensor of {n} on {dev}")
        else:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.size == n) or dev is not None:
                raise _lib.TgravError(_lib.EINVAL, "argument",
                                      f"{name} must be a contiguous float64 numpy array of {n}")
