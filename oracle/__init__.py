"""CPU oracle for the softened direct-sum gravity field -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results against (or as
the timed CPU baseline).  The product (``torch_amd``) never imports it.

``liboracle.so`` is built from ``tgrav_oracle.c`` by ``oracle/Makefile``; each C
function cites the reference file:line it restates.  ``oracle.brute`` holds a
50-digit ``decimal`` brute force that pins the C restatement independently.

Parity status of each function (what pins it is in tests/test_oracle.py):
  eval_plummer    pinned: decimal brute force, closed forms, invariants
  eval_flash      pinned: closed forms (r >= rsoft Newtonian, continuity at rsoft,
                  r -> 0 limit, a = -grad phi in clean mode); the reference's quirk
                  potential inside rsoft by quirk + clean = hinv q^4 / 5 (q < 1), 0 (1 <= q < 2),
                  with planted coefficient mistakes failing it; periodic branch: equals the
                  isolated kernel over explicitly replicated, cutoff-filtered images; zero
                  field at lattice symmetry points
  bidirectional   pinned: equals two eval_plummer calls; Newton's third law
  kick            pinned: exact arithmetic on representable inputs
  block_centres   pinned: numpy meshgrid('ij').ravel('C') of the same centres
  pm_deposit      pinned: closed forms (a particle on a cell centre, split weights), mass
                  conservation, linearity; adjoint identity with pm_interpolate
  pm_gradient     pinned: exact on quadratic potentials (central differences), end rule
  pm_interpolate  pinned: exact on linear fields, value at cell centres, adjoint identity
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SEQ, NEUMAIER, LONGDOUBLE = 0, 1, 2
_D = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = load(path)
    return _LIB


def load(path):
    """A library built from tgrav_oracle.c at `path`, with its argument types declared (the
    tests load copies with planted mistakes this way)."""
    L = ctypes.CDLL(path)
    i64, dbl, i32 = ctypes.c_int64, ctypes.c_double, ctypes.c_int
    L.oracle_eval_plummer.argtypes = [i64] + [_D] * 5 + [i64] + [_D] * 4 + [dbl] + [_D] * 4 + [i32, i32]
    L.oracle_eval_flash.argtypes = ([i64] + [_D] * 4 + [i64] + [_D] * 3 + [dbl, dbl, i32, i32, i32, dbl, dbl, dbl]
                                    + [_D] * 4 + [i32])
    L.oracle_bidirectional.argtypes = [i64] + [_D] * 5 + [i64] + [_D] * 5 + [dbl] + [_D] * 8
    L.oracle_kick.argtypes = [i64, dbl] + [_D] * 6
    L.oracle_block_centres.argtypes = [i64, i32] + [_D] * 7
    _I = ctypes.POINTER(ctypes.c_int64)
    L.oracle_pm_deposit.argtypes = [i64] + [_D] * 4 + [_I, i64, i32, i32] + [_D] * 5
    L.oracle_pm_gradient.argtypes = [i64, i32, i32] + [_D] * 5
    L.oracle_pm_interpolate.argtypes = [i64] + [_D] * 3 + [_I, i64, i32, i32] + [_D] * 10
    for f in (L.oracle_eval_plummer, L.oracle_eval_flash, L.oracle_bidirectional,
              L.oracle_kick, L.oracle_block_centres, L.oracle_pm_deposit,
              L.oracle_pm_gradient, L.oracle_pm_interpolate):
        f.restype = ctypes.c_int
    return L


def _in(a, n=None):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.shape != (n,):
        raise ValueError(f"expected shape ({n},), got {a.shape}")
    return a, a.ctypes.data_as(_D)


def _out(n):
    a = np.zeros(n, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def eval_plummer(sx, sy, sz, sm, se, tx, ty, tz, te, G, mode=SEQ, threads=1):
    """(ax, ay, az, phi) at the targets from the sources (Plummer kernel)."""
    ns, nt = len(sx), len(tx)
    keep = [_in(a, ns) for a in (sx, sy, sz, sm, se)] + [_in(a, nt) for a in (tx, ty, tz, te)]
    outs = [_out(nt) for _ in range(4)]
    rc = lib().oracle_eval_plummer(ns, *[k[1] for k in keep[:5]], nt, *[k[1] for k in keep[5:]],
                                   float(G), *[o[1] for o in outs], int(mode), int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle_eval_plummer failed ({rc})")
    return tuple(o[0] for o in outs)


def eval_flash(sx, sy, sz, sm, tx, ty, tz, G, rsoft, kernel="spline", quirks=False, threads=1,
               box=None, nrep=6):
    """(ax, ay, az, phi) with FLASH's softening kernels; box=(Lx, Ly, Lz) selects the periodic
    image sum over n in [-nrep, nrep]^3 within nrep * min(box)."""
    ns, nt = len(sx), len(tx)
    L = (0.0, 0.0, 0.0) if box is None else tuple(float(v) for v in box)
    nr = 0 if box is None else int(nrep)
    keep = [_in(a, ns) for a in (sx, sy, sz, sm)] + [_in(a, nt) for a in (tx, ty, tz)]
    outs = [_out(nt) for _ in range(4)]
    typ = {"spline": 1, "linear": 2}[kernel]
    rc = lib().oracle_eval_flash(ns, *[k[1] for k in keep[:4]], nt, *[k[1] for k in keep[4:]],
                                 float(G), float(rsoft), typ, int(bool(quirks)), nr, *L,
                                 *[o[1] for o in outs], int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle_eval_flash failed ({rc})")
    return tuple(o[0] for o in outs)


def bidirectional(cx, cy, cz, cm, ce, px, py, pz, pm, pe, G):
    """Returns ((pax,pay,paz,pphi) on particles, (cax,cay,caz,cphi) on cells)."""
    nc, np_ = len(cx), len(px)
    keep = [_in(a, nc) for a in (cx, cy, cz, cm, ce)] + [_in(a, np_) for a in (px, py, pz, pm, pe)]
    po = [_out(np_) for _ in range(4)]
    co = [_out(nc) for _ in range(4)]
    rc = lib().oracle_bidirectional(nc, *[k[1] for k in keep[:5]], np_, *[k[1] for k in keep[5:]],
                                    float(G), *[o[1] for o in po], *[o[1] for o in co])
    if rc != 0:
        raise RuntimeError(f"oracle_bidirectional failed ({rc})")
    return tuple(o[0] for o in po), tuple(o[0] for o in co)


def kick(dt, ax, ay, az, vx, vy, vz):
    """Returns new (vx, vy, vz) = v + dt*a."""
    n = len(ax)
    a = [_in(v, n) for v in (ax, ay, az)]
    v = [np.array(q, dtype=np.float64, copy=True) for q in (vx, vy, vz)]
    rc = lib().oracle_kick(n, float(dt), *[q[1] for q in a], *[q.ctypes.data_as(_D) for q in v])
    if rc != 0:
        raise RuntimeError("oracle_kick failed")
    return tuple(v)


def block_centres(ox, oy, oz, dx, nb=8):
    nblk = len(ox)
    keep = [_in(a, nblk) for a in (ox, oy, oz, dx)]
    outs = [_out(nblk * nb ** 3) for _ in range(3)]
    rc = lib().oracle_block_centres(nblk, nb, *[k[1] for k in keep], *[o[1] for o in outs])
    if rc != 0:
        raise RuntimeError("oracle_block_centres failed")
    return tuple(o[0] for o in outs)


def _blocks(blocks):
    """blocks = dict(origin=(nblk, 3), size=(nblk,), nb=8, ng=...)."""
    o = np.ascontiguousarray(np.asarray(blocks["origin"], dtype=np.float64).reshape(-1, 3))
    bs = np.ascontiguousarray(np.asarray(blocks["size"], dtype=np.float64).reshape(-1))
    ox, oy, oz = (np.ascontiguousarray(o[:, a]) for a in range(3))
    return len(bs), int(blocks.get("nb", 8)), int(blocks["ng"]), ox, oy, oz, bs


def _i64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def pm_deposit(x, y, z, m, blk, blocks, rho=None):
    """CIC density of the particles on guarded blocks (added to rho if given)."""
    nblk, nb, ng, ox, oy, oz, bs = _blocks(blocks)
    nbg = nb + 2 * ng
    n = len(x)
    out = np.zeros(nblk * nbg ** 3) if rho is None else np.array(rho, dtype=np.float64, copy=True)
    keep = [_in(a, n) for a in (x, y, z, m)]
    kb = _i64(blk)
    g = [_in(a, nblk) for a in (ox, oy, oz, bs)]
    rc = lib().oracle_pm_deposit(n, *[k[1] for k in keep], kb[1], nblk, nb, ng, *[q[1] for q in g],
                                 out.ctypes.data_as(_D))
    if rc != 0:
        raise RuntimeError(f"oracle_pm_deposit failed ({rc})")
    return out


def pm_gradient(phi, blocks):
    nblk, nb, ng, ox, oy, oz, bs = _blocks(blocks)
    nbg = nb + 2 * ng
    f = _in(phi, nblk * nbg ** 3)
    outs = [_out(nblk * nbg ** 3) for _ in range(3)]
    rc = lib().oracle_pm_gradient(nblk, nb, ng, _in(bs, nblk)[1], f[1], *[o[1] for o in outs])
    if rc != 0:
        raise RuntimeError(f"oracle_pm_gradient failed ({rc})")
    return tuple(o[0] for o in outs)


def pm_interpolate(x, y, z, blk, blocks, gx, gy, gz):
    nblk, nb, ng, ox, oy, oz, bs = _blocks(blocks)
    nbg = nb + 2 * ng
    n = len(x)
    keep = [_in(a, n) for a in (x, y, z)]
    kb = _i64(blk)
    g = [_in(a, nblk) for a in (ox, oy, oz, bs)]
    fields = [_in(a, nblk * nbg ** 3) for a in (gx, gy, gz)]
    outs = [_out(n) for _ in range(3)]
    rc = lib().oracle_pm_interpolate(n, *[k[1] for k in keep], kb[1], nblk, nb, ng, *[q[1] for q in g],
                                     *[f[1] for f in fields], *[o[1] for o in outs])
    if rc != 0:
        raise RuntimeError(f"oracle_pm_interpolate failed ({rc})")
    return tuple(o[0] for o in outs)
