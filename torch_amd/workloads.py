"""Seeded synthetic inputs for the direct-sum gravity benchmarks (CPU, numpy only).

Units: pc, Msun, km/s (time unit pc/(km/s) ~ 0.9778 Myr).  Every generator draws
from ``numpy.random.Generator(PCG64(seed))`` so a seed reproduces a workload bit for
bit.  Nothing here touches the GPU.

Workloads (all FP64):
  c1  1,024 Plummer stars (r_vir = 1 pc), Kroupa 0.08-100 Msun, eps_s = 0.01 pc ->
      4,096 centres of a uniform 16^3 grid on [-2,2]^3 pc (eps_t = 0).
      Reverse: those 4,096 cells as point masses (m = rho dx^3 from a sigma = 1.5
      lognormal field, eps_s = dx) -> the 1,024 stars (eps_t = 0).
  c2  10,000 stars (Plummer r_vir = 1 pc, Kroupa, 20% of stars in 1,000 hard binaries,
      a log-uniform 1-100 AU, thermal e), eps_s = 0.01 pc -> 2,000 leaf blocks of 8^3
      cells (1,024,000 centres) from density-weighted refinement of a sigma = 1.5
      lognormal field on a 4 pc box (base 8^3 blocks of 0.5 pc, minimum 0.03125 pc).
      Reverse: the cells (m = rho dx^3, eps_s = dx) -> the stars (eps_t = 0).
  c4  64 independent problems, N_s = floor(10^U[2, log10 5e4]) sources (Plummer/Kroupa
      stars for even problems, gas cells for odd ones), N_t = 512 ceil(10^U[3,6]/512)
      block-cell targets, per-source eps ~ U[0.005, 0.1] pc.
  c5  100,000 stars (Plummer r_vir = 2 pc, 30% binary members) -> 8,192 leaf blocks
      (4,194,304 centres) of a sigma = 2 lognormal field on a 10 pc box (base 16^3
      blocks of 0.625 pc, minimum 10/256 pc); reverse cells -> stars as in c2.

The eps_t = 0 choice for stars in the reverse direction follows the reference's
own call ``hydro.get_accel_gas_on_particles(0.0 | units.m, ...)``
(.deprecated/bridge_multiples.py:605).  Kroupa slopes follow src/imf_sample.py:144-155
(alpha = 1.3 below 0.5 Msun, 2.3 above).  The gas field has P(k) ~ k^-4 (Burgers-like);
its mean density is 100 Msun/pc^3 (c1, c2) or 10 Msun/pc^3 (c5).
"""
import heapq
import math

import numpy as np

G_PC_MSUN_KMS = 4.300917270e-3     # G in pc (km/s)^2 / Msun
AU_PC = 4.84813681e-6              # 1 AU in pc


def rng_for(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ---------------------------------------------------------------- stars ----
def kroupa_masses(rng, n, mmin=0.08, mbreak=0.5, mmax=100.0, a1=1.3, a2=2.3):
    """Inverse-CDF draw from the two-segment Kroupa IMF, continuous at mbreak."""
    e1, e2 = 1.0 - a1, 1.0 - a2
    i1 = (mbreak ** e1 - mmin ** e1) / e1
    c2 = mbreak ** (a2 - a1)
    i2 = c2 * (mmax ** e2 - mbreak ** e2) / e2
    u = rng.random(n) * (i1 + i2)
    lo = (mmin ** e1 + u * e1) ** (1.0 / e1)
    hi = (mbreak ** e2 + (u - i1) * e2 / c2) ** (1.0 / e2)
    return np.where(u < i1, lo, hi)


def _isotropic(rng, n):
    cz = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    sz = np.sqrt(1.0 - cz * cz)
    return np.stack([sz * np.cos(ph), sz * np.sin(ph), cz], axis=1)


def plummer_positions(rng, n, r_vir=1.0, mfrac=0.999):
    """Aarseth-Henon-Wielen (1974) radii, truncated at the mfrac mass fraction."""
    b = 3.0 * np.pi / 16.0 * r_vir
    x1 = rng.uniform(0.0, mfrac, n)
    x1 = np.maximum(x1, 1e-300)
    r = b / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    return r[:, None] * _isotropic(rng, n), b


def plummer_velocities(rng, pos, b, mtot, G=G_PC_MSUN_KMS):
    """AHW74 rejection sampling of q = v/v_esc from q^2 (1-q^2)^(7/2)."""
    n = len(pos)
    r2 = np.sum(pos * pos, axis=1)
    vesc = np.sqrt(2.0 * G * mtot / b) * (1.0 + r2 / (b * b)) ** (-0.25)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.random(todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    return (q * vesc)[:, None] * _isotropic(rng, n)


def _kepler_E(M, e, iters=40):
    E = np.where(e > 0.8, np.pi * np.ones_like(M), M)
    for _ in range(iters):
        E = E - (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
    return E


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


def star_cluster(rng, n, r_vir=1.0, binary_frac=0.0, a_min_au=1.0, a_max_au=100.0, eps=0.01,
                 G=G_PC_MSUN_KMS):
    """Plummer/Kroupa cluster; binary_frac of the n stars are members of hard binaries
    (n*binary_frac/2 pairs; a log-uniform in [a_min, a_max] AU, e = sqrt(U) thermal,
    uniform mean anomaly, random orientation; each pair's centre of mass takes a Plummer
    position/velocity).  Returns a dict of FP64 arrays plus `binary` (pair index or -1)."""
    npairs = int(round(n * binary_frac)) // 2
    nsys = n - npairs                      # singles + pair centres of mass
    m = kroupa_masses(rng, n)
    com, b = plummer_positions(rng, nsys, r_vir)
    vcom = plummer_velocities(rng, com, b, float(m.sum()), G)
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    binary = -np.ones(n, dtype=np.int64)
    nsing = nsys - npairs
    pos[:nsing] = com[:nsing]
    vel[:nsing] = vcom[:nsing]
    if npairs:
        m1 = m[nsing:nsing + npairs]
        m2 = m[nsing + npairs:]
        mt = m1 + m2
        a = AU_PC * 10.0 ** rng.uniform(np.log10(a_min_au), np.log10(a_max_au), npairs)
        e = np.sqrt(rng.random(npairs))
        M = rng.uniform(0.0, 2.0 * np.pi, npairs)
        E = _kepler_E(M, e)
        nmot = np.sqrt(G * mt / a ** 3)
        denom = 1.0 - e * np.cos(E)
        rel = np.stack([a * (np.cos(E) - e), a * np.sqrt(1 - e * e) * np.sin(E), np.zeros(npairs)], 1)
        vrel = np.stack([-a * nmot * np.sin(E) / denom,
                         a * nmot * np.sqrt(1 - e * e) * np.cos(E) / denom, np.zeros(npairs)], 1)
        R = _random_rotations(rng, npairs)
        rel = np.einsum("nij,nj->ni", R, rel)
        vrel = np.einsum("nij,nj->ni", R, vrel)
        c, vc = com[nsing:], vcom[nsing:]
        f1 = (m2 / mt)[:, None]
        f2 = (m1 / mt)[:, None]
        i1 = slice(nsing, nsing + npairs)
        i2 = slice(nsing + npairs, n)
        pos[i1] = c + f1 * rel
        pos[i2] = c - f2 * rel
        vel[i1] = vc + f1 * vrel
        vel[i2] = vc - f2 * vrel
        binary[i1] = np.arange(npairs)
        binary[i2] = np.arange(npairs)
    return dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
                vx=vel[:, 0].copy(), vy=vel[:, 1].copy(), vz=vel[:, 2].copy(),
                m=m, eps=np.full(n, float(eps)), binary=binary)


# ------------------------------------------------------------------ gas ----
def lognormal_field(rng, ngrid, sigma, slope=-4.0):
    """Periodic lognormal density on ngrid^3 with unit mean: g Gaussian with P(k) ~ k^slope,
    unit variance; rho = exp(sigma g - sigma^2/2)."""
    white = rng.normal(size=(ngrid, ngrid, ngrid))
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(ngrid)[:, None, None]
    ky = np.fft.fftfreq(ngrid)[None, :, None]
    kz = np.fft.rfftfreq(ngrid)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = np.inf
    fk *= k2 ** (slope / 4.0)
    g = np.fft.irfftn(fk, s=(ngrid, ngrid, ngrid), axes=(0, 1, 2))
    g = (g - g.mean()) / g.std()
    return np.exp(sigma * g - 0.5 * sigma * sigma)


def refine_blocks(field, box, base_n, min_size, n_leaves):
    """Density-weighted octree: split the most massive leaf (if larger than min_size) until
    there are >= n_leaves leaves, then drop the lightest leaves to exactly n_leaves.
    Returns (origin[n,3], size[n]) sorted by (size, origin) order of creation."""
    ng = field.shape[0]
    cell = box / ng
    sat = np.zeros((ng + 1, ng + 1, ng + 1))
    sat[1:, 1:, 1:] = field.cumsum(0).cumsum(1).cumsum(2)

    def mass(o, s):
        i0 = np.rint((o + box / 2) / cell).astype(int)
        i1 = i0 + int(round(s / cell))
        a, b = i0, i1
        tot = (sat[b[0], b[1], b[2]] - sat[a[0], b[1], b[2]] - sat[b[0], a[1], b[2]]
               - sat[b[0], b[1], a[2]] + sat[a[0], a[1], b[2]] + sat[a[0], b[1], a[2]]
               + sat[b[0], a[1], a[2]] - sat[a[0], a[1], a[2]])
        return tot * cell ** 3

    s0 = box / base_n
    heap = []
    uid = 0
    for i in range(base_n):
        for j in range(base_n):
            for k in range(base_n):
                o = np.array([i, j, k]) * s0 - box / 2
                heapq.heappush(heap, (-mass(o, s0), uid, tuple(o), s0))
                uid += 1
    frozen = []
    while len(heap) + len(frozen) < n_leaves:
        negm, _, o, s = heapq.heappop(heap)
        if s / 2 < min_size - 1e-12:
            frozen.append((negm, _, o, s))
            if not heap:
                break
            continue
        h = s / 2
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    oc = np.array(o) + np.array([di, dj, dk]) * h
                    heapq.heappush(heap, (-mass(oc, h), uid, tuple(oc), h))
                    uid += 1
    leaves = sorted(heap + frozen)[:n_leaves]           # heaviest first
    leaves.sort(key=lambda t: t[1])                     # creation order
    origin = np.array([t[2] for t in leaves], dtype=np.float64)
    size = np.array([t[3] for t in leaves], dtype=np.float64)
    return origin, size


def block_cell_centres(origin, size, nb=8):
    """Cell centres (x, y, z, dx) of blocks with lower corners `origin` and edge `size`."""
    from .blocks import block_centres
    return block_centres(origin, size / nb, nb)


def sample_field(field, box, x, y, z):
    """Trilinear interpolation of the periodic cell-centred field at points in [-box/2, box/2)."""
    ng = field.shape[0]
    u = [(c + box / 2) / box * ng - 0.5 for c in (x, y, z)]
    i0 = [np.floor(c).astype(np.int64) for c in u]
    f = [c - i for c, i in zip(u, i0)]
    out = np.zeros_like(x)
    for dx in (0, 1):
        wx = f[0] if dx else 1 - f[0]
        ix = (i0[0] + dx) % ng
        for dy in (0, 1):
            wy = f[1] if dy else 1 - f[1]
            iy = (i0[1] + dy) % ng
            for dz in (0, 1):
                wz = f[2] if dz else 1 - f[2]
                iz = (i0[2] + dz) % ng
                out += wx * wy * wz * field[ix, iy, iz]
    return out


def gas_cells(rng, box, base_n, n_leaves, sigma, ngrid, rho_mean, min_size=None, nb=8):
    field = lognormal_field(rng, ngrid, sigma) * rho_mean
    if min_size is None:
        min_size = box / ngrid
    origin, size = refine_blocks(field, box, base_n, min_size, n_leaves)
    x, y, z, dx = block_cell_centres(origin, size, nb)
    rho = sample_field(field, box, x, y, z)
    return dict(x=x, y=y, z=z, m=rho * dx ** 3, eps=dx.copy(), dx=dx, rho=rho,
                block_origin=origin, block_size=size, nb=nb)


def uniform_grid(n, box):
    dx = box / n
    c = -box / 2 + (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return X.ravel().copy(), Y.ravel().copy(), Z.ravel().copy(), dx


# -------------------------------------------------------------- configs ----
def c1(seed=1001):
    rng = rng_for(seed)
    stars = star_cluster(rng, 1024, r_vir=1.0, eps=0.01)
    x, y, z, dx = uniform_grid(16, 4.0)
    rho = lognormal_field(rng, 16, 1.5) * 100.0
    cells = dict(x=x, y=y, z=z, m=rho.ravel() * dx ** 3, eps=np.full(x.size, dx),
                 dx=np.full(x.size, dx), rho=rho.ravel())
    return dict(name="c1", seed=seed, stars=stars, cells=cells)


def c2(seed=1002):
    rng = rng_for(seed)
    stars = star_cluster(rng, 10_000, r_vir=1.0, binary_frac=0.2, eps=0.01)
    cells = gas_cells(rng, box=4.0, base_n=8, n_leaves=2000, sigma=1.5, ngrid=128,
                      rho_mean=100.0, min_size=0.03125)
    return dict(name="c2", seed=seed, stars=stars, cells=cells)


def c5(seed=1005):
    rng = rng_for(seed)
    stars = star_cluster(rng, 100_000, r_vir=2.0, binary_frac=0.3, eps=0.01)
    cells = gas_cells(rng, box=10.0, base_n=16, n_leaves=8192, sigma=2.0, ngrid=256,
                      rho_mean=10.0, min_size=10.0 / 256)
    return dict(name="c5", seed=seed, stars=stars, cells=cells)


def c4(seed=1004, n_problems=64):
    """64 uneven problems; each is dict(sources=..., targets=...)."""
    rng = rng_for(seed)
    probs = []
    for p in range(n_problems):
        ns = int(math.floor(10.0 ** rng.uniform(2.0, math.log10(5e4))))
        nt = 512 * int(math.ceil(10.0 ** rng.uniform(3.0, 6.0) / 512))
        if p % 2 == 0:
            st = star_cluster(rng, ns, r_vir=1.0)
            src = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"])
        else:
            nblk = -(-ns // 512)
            o = rng.uniform(-2.0, 1.5, (nblk, 3))
            s = 10.0 ** rng.uniform(np.log10(0.03125), np.log10(0.5), nblk)
            x, y, z, dx = block_cell_centres(o, s)
            rho = 100.0 * np.exp(1.5 * rng.normal(size=x.size) - 1.125)
            src = dict(x=x[:ns], y=y[:ns], z=z[:ns], m=(rho * dx ** 3)[:ns])
        src["eps"] = rng.uniform(0.005, 0.1, ns)
        nblk = nt // 512
        o = rng.uniform(-2.0, 1.5, (nblk, 3))
        s = 10.0 ** rng.uniform(np.log10(0.03125), np.log10(0.5), nblk)
        tx, ty, tz, _ = block_cell_centres(o, s)
        probs.append(dict(sources=src, targets=dict(x=tx, y=ty, z=tz, eps=np.zeros(nt), block_origin=o,
                                                    block_dx=s / 8)))
    return probs


def containing_block(x, y, z, origin, size):
    """Index of the leaf block (lower corner origin[b], edge size[b]) holding each point, -1
    outside all blocks (blocks are disjoint leaves of an octree)."""
    out = np.full(len(x), -1, dtype=np.int64)
    p = np.stack([x, y, z], 1)
    for lo in range(0, len(size), 256):
        o = origin[lo:lo + 256]
        s = size[lo:lo + 256]
        inside = np.all((p[:, None, :] >= o[None, :, :]) & (p[:, None, :] < o[None, :, :] + s[None, :, None]), axis=2)
        hit = inside.any(1)
        out[hit] = lo + inside[hit].argmax(1)
    return out
