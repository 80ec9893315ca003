"""Seeded synthetic volumes for the benchmark workloads and the tests.

Every generator draws from SplitMix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), so a (generator, seed) pair
always gives the same parameters.  Volumes are laid out [z][y][x] (x
fastest), the node order of kvs::StructuredVolumeObject.

  blobs(n, ...)        sum of isotropic Gaussian blobs plus uniform noise (float32)
  ct_like(n, ...)      3-D value noise on a 32^3 lattice plus constant ellipsoids (uint8)
  turbulence(n, ...)   sum of random-phase Fourier modes, |k|^(-5/3), in [-1, 1] (float32)

Large volumes are generated on the GPU with torch (plumbing only: the
benchmark needs a 4 GiB field in HBM); bytes are identical wherever the same
function produces them, and the CPU baseline always reads the very bytes the
GPU extracts from.
"""
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """SplitMix64 stream: state += golden; output = mix(state)."""

    def __init__(self, seed):
        self.state = int(seed) & 0xFFFFFFFFFFFFFFFF

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def bulk_uniform(self, n):
        """n uniforms in [0,1) from the next n outputs of the stream (vectorised)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _grid(dims):
    nx, ny, nz = dims
    x = np.arange(nx, dtype=np.float64) / max(nx - 1, 1)
    y = np.arange(ny, dtype=np.float64) / max(ny - 1, 1)
    z = np.arange(nz, dtype=np.float64) / max(nz - 1, 1)
    return x, y, z


def blob_params(seed, nblobs):
    rng = SplitMix64(seed)
    out = []
    for _ in range(nblobs):
        c = [rng.uniform(0.2, 0.8) for _ in range(3)]
        s = rng.uniform(0.08, 0.2)
        out.append((c, s))
    return rng, out


def blobs(dims, seed=0, nblobs=3, noise=0.01):
    """Sum of nblobs isotropic Gaussians (peak 1, centres U[0.2,0.8]^3 of the
    unit box, sigma U[0.08,0.2]) plus U[-noise, noise], float32 [z][y][x]."""
    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    rng, params = blob_params(seed, nblobs)
    x, y, z = _grid(dims)
    f = np.zeros((nz, ny, nx), dtype=np.float64)
    for (c, s) in params:
        gx = np.exp(-((x - c[0]) ** 2) / (2 * s * s))
        gy = np.exp(-((y - c[1]) ** 2) / (2 * s * s))
        gz = np.exp(-((z - c[2]) ** 2) / (2 * s * s))
        f += gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    if noise:
        f += (rng.bulk_uniform(nx * ny * nz).reshape(nz, ny, nx) * 2.0 - 1.0) * noise
    return f.astype(np.float32)


def ct_like(dims, seed=1, nellipsoids=20, dtype=np.uint8, lattice=32):
    """CT-like volume: value noise on a (lattice+1)^3 grid, trilinearly
    upsampled, plus nellipsoids randomly rotated ellipsoids of constant value
    U{60..220}/255 overriding the noise, rescaled to the dtype's 0..max."""
    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    rng = SplitMix64(seed)
    lat = rng.bulk_uniform((lattice + 1) ** 3).reshape(lattice + 1, lattice + 1, lattice + 1)
    x, y, z = _grid(dims)

    def axis_weights(t):
        u = t * lattice
        i0 = np.minimum(np.floor(u).astype(np.int64), lattice - 1)
        return i0, u - i0

    ix, wx = axis_weights(x)
    iy, wy = axis_weights(y)
    iz, wz = axis_weights(z)
    f = np.zeros((nz, ny, nx), dtype=np.float64)
    for dz in (0, 1):
        az = (wz if dz else 1 - wz)[:, None, None]
        for dy in (0, 1):
            ay = (wy if dy else 1 - wy)[None, :, None]
            for dx in (0, 1):
                ax = (wx if dx else 1 - wx)[None, None, :]
                f += az * ay * ax * lat[(iz + dz)[:, None, None], (iy + dy)[None, :, None], (ix + dx)[None, None, :]]
    zz = z[:, None, None]
    yy = y[None, :, None]
    xx = x[None, None, :]
    for _ in range(nellipsoids):
        c = [rng.uniform(0.1, 0.9) for _ in range(3)]
        ax = [rng.uniform(0.03, 0.15) for _ in range(3)]
        # random rotation from a uniform unit quaternion (Shoemake)
        u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
        q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
             math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
        qx, qy, qz, qw = q
        rot = np.array([
            [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
            [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
            [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
        value = rng.integer(60, 220) / 255.0
        dx, dy, dz = xx - c[0], yy - c[1], zz - c[2]
        inside = np.ones((nz, ny, nx), dtype=bool)
        acc = np.zeros((nz, ny, nx), dtype=np.float64)
        for k in range(3):
            pk = rot[0, k] * dx + rot[1, k] * dy + rot[2, k] * dz
            acc = acc + (pk / ax[k]) ** 2
        inside &= acc <= 1.0
        f[inside] = value
    lo, hi = f.min(), f.max()
    top = float(np.iinfo(dtype).max) if np.dtype(dtype).kind in "ui" else 1.0
    g = (f - lo) / (hi - lo) * top if hi > lo else np.zeros_like(f)
    if np.dtype(dtype).kind in "ui":
        return np.rint(g).astype(dtype)
    return g.astype(dtype)


def turbulence_modes(seed, nmodes=64, kmin=1.0, kmax=64.0):
    """(k vectors, phases, amplitudes) of the turbulence-like field."""
    rng = SplitMix64(seed)
    ks, ph, amp = [], [], []
    for _ in range(nmodes):
        mag = rng.uniform(kmin, kmax)
        cz = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0.0, 2 * math.pi)
        r = math.sqrt(max(0.0, 1 - cz * cz))
        ks.append((mag * r * math.cos(phi), mag * r * math.sin(phi), mag * cz))
        ph.append(rng.uniform(0.0, 2 * math.pi))
        amp.append(mag ** (-5.0 / 3.0))
    return np.array(ks), np.array(ph), np.array(amp)


def turbulence(dims, seed=2, nmodes=64, period=None):
    """Sum of nmodes Fourier modes A cos(2 pi / period * k.p + phi), |k| ~ U[1,64],
    A = |k|^(-5/3), normalised to [-1, 1] by its min/max; float32 [z][y][x]
    on the host (numpy; small sizes)."""
    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    period = float(period or max(dims))
    ks, ph, amp = turbulence_modes(seed, nmodes)
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)
    w = 2 * math.pi / period
    f = np.zeros((nz, ny, nx), dtype=np.float64)
    for m in range(nmodes):
        f += amp[m] * np.cos(w * (ks[m, 0] * x[None, None, :] + ks[m, 1] * y[None, :, None]
                                  + ks[m, 2] * z[:, None, None]) + ph[m])
    lo, hi = f.min(), f.max()
    return (2.0 * (f - lo) / (hi - lo) - 1.0).astype(np.float32)


def turbulence_torch(dims, seed=2, nmodes=64, period=None, device="cuda", slab=32):
    """turbulence() evaluated on the GPU with torch, slab by slab (FP64
    accumulation, float32 result).  Returns a contiguous [z][y][x] tensor."""
    import torch

    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    period = float(period or max(dims))
    ks, ph, amp = turbulence_modes(seed, nmodes)
    w = 2 * math.pi / period
    x = torch.arange(nx, dtype=torch.float64, device=device)
    y = torch.arange(ny, dtype=torch.float64, device=device)
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=device)
    lo, hi = math.inf, -math.inf
    raw = []
    for z0 in range(0, nz, slab):
        z1 = min(nz, z0 + slab)
        z = torch.arange(z0, z1, dtype=torch.float64, device=device)
        acc = torch.zeros((z1 - z0, ny, nx), dtype=torch.float64, device=device)
        for m in range(nmodes):
            px = (w * ks[m, 0]) * x
            py = (w * ks[m, 1]) * y
            pz = (w * ks[m, 2]) * z + ph[m]
            acc += amp[m] * torch.cos(pz[:, None, None] + py[None, :, None] + px[None, None, :])
        lo = min(lo, float(acc.min()))
        hi = max(hi, float(acc.max()))
        out[z0:z1] = acc.to(torch.float32)   # provisional, rescaled below
        raw.append((z0, z1))
        del acc
    # rescale in float32 from the provisional values: a deterministic map
    scale = 2.0 / (hi - lo)
    for z0, z1 in raw:
        t = out[z0:z1].to(torch.float64)
        out[z0:z1] = ((t - lo) * scale - 1.0).to(torch.float32)
    return out


def _splitmix_uniform_torch(state, n, device):
    """The next n outputs of a SplitMix64 stream at `state`, as uniforms in
    [0,1) (float64 tensor), bit-identical to SplitMix64.bulk_uniform.  int64
    arithmetic wraps like uint64; shifts are made logical by masking."""
    import torch

    def s64(u):
        return u - (1 << 64) if u >= (1 << 63) else u

    golden, m1, m2 = s64(0x9E3779B97F4A7C15), s64(0xBF58476D1CE4E5B9), s64(0x94D049BB133111EB)
    k = torch.arange(1, n + 1, dtype=torch.int64, device=device)
    z = k * golden + s64(state)

    def lshr(v, s):
        return (v >> s) & ((1 << (64 - s)) - 1)

    z = (z ^ lshr(z, 30)) * m1
    z = (z ^ lshr(z, 27)) * m2
    z = z ^ lshr(z, 31)
    return lshr(z, 11).to(torch.float64) * (1.0 / 9007199254740992.0)


def blobs_torch(dims, seed=0, nblobs=3, noise=0.01, device="cuda"):
    """blobs() evaluated with torch (on the GPU for large batches): the same
    formula and noise stream; on the CPU device the bytes equal blobs()'s, on
    the GPU the float64 exp may differ from the host's in the last bit."""
    import torch

    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    rng, params = blob_params(seed, nblobs)
    ax = [torch.arange(n, dtype=torch.float64, device=device) / max(n - 1, 1) for n in (nx, ny, nz)]
    f = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for (c, s) in params:
        g = [torch.exp(-((ax[i] - c[i]) ** 2) / (2 * s * s)) for i in range(3)]
        f += g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    if noise:
        u = _splitmix_uniform_torch(rng.state, nx * ny * nz, device).reshape(nz, ny, nx)
        f += (u * 2.0 - 1.0) * noise
    return f.to(torch.float32)


def batch_workload(n=256, seed=3000, dmin=32, dmax=256, max_blobs=16, max_levels=4):
    """Jobs of the batched workload: per volume, dims uniform in [dmin, dmax]
    per axis, a blob field with 1..max_blobs blobs and 1..max_levels isolevels
    uniform in [0.2, 0.8].  Returns [(dims (nx,ny,nz), blob seed, nblobs, levels)]."""
    rng = SplitMix64(seed)
    jobs = []
    for i in range(n):
        dims = tuple(rng.integer(dmin, dmax) for _ in range(3))
        nb = rng.integer(1, max_blobs)
        levels = [rng.uniform(0.2, 0.8) for _ in range(rng.integer(1, max_levels))]
        jobs.append((dims, seed + 1 + i, nb, levels))
    return jobs


def ct_like_torch(dims, seed=1, nellipsoids=20, dtype=np.uint8, lattice=32, device="cuda"):
    """ct_like() evaluated with torch (the same parameter stream and formula;
    large volumes on the GPU).  Returns a contiguous [z][y][x] tensor."""
    import torch

    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    rng = SplitMix64(seed)
    lat = torch.from_numpy(rng.bulk_uniform((lattice + 1) ** 3).reshape(lattice + 1, lattice + 1, lattice + 1)).to(device)

    def axis(n):
        t = torch.arange(n, dtype=torch.float64, device=device) / max(n - 1, 1)
        u = t * lattice
        i0 = torch.clamp(torch.floor(u).to(torch.int64), max=lattice - 1)
        return t, i0, u - i0

    x, ix, wx = axis(nx)
    y, iy, wy = axis(ny)
    z, iz, wz = axis(nz)
    f = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for dz in (0, 1):
        az = (wz if dz else 1 - wz)[:, None, None]
        for dy in (0, 1):
            ay = (wy if dy else 1 - wy)[None, :, None]
            for dx in (0, 1):
                ax_ = (wx if dx else 1 - wx)[None, None, :]
                f += az * ay * ax_ * lat[(iz + dz)[:, None, None], (iy + dy)[None, :, None], (ix + dx)[None, None, :]]
    zz, yy, xx = z[:, None, None], y[None, :, None], x[None, None, :]
    for _ in range(nellipsoids):
        c = [rng.uniform(0.1, 0.9) for _ in range(3)]
        ax = [rng.uniform(0.03, 0.15) for _ in range(3)]
        u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
        q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
             math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
        qx, qy, qz, qw = q
        rot = [[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
               [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
               [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]]
        value = rng.integer(60, 220) / 255.0
        dx, dy, dz = xx - c[0], yy - c[1], zz - c[2]
        acc = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
        for k in range(3):
            pk = rot[0][k] * dx + rot[1][k] * dy + rot[2][k] * dz
            acc = acc + (pk / ax[k]) ** 2
        f = torch.where(acc <= 1.0, torch.full_like(f, value), f)
        del acc
    lo, hi = float(f.min()), float(f.max())
    top = float(np.iinfo(dtype).max) if np.dtype(dtype).kind in "ui" else 1.0
    g = (f - lo) / (hi - lo) * top if hi > lo else torch.zeros_like(f)
    tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.uint16, np.dtype(np.int16): torch.int16,
           np.dtype(np.float32): torch.float32}[np.dtype(dtype)]
    if np.dtype(dtype).kind in "ui":
        return torch.round(g).to(tdt).contiguous()
    return g.to(tdt).contiguous()


def serving_volumes(n=32, seed=4000, sizes=tuple(range(64, 513, 32))):
    """Volumes of the serving workload: cubic edge drawn from `sizes`, value
    type cycling float32 (turbulence field), uint16 and uint8 (CT-like).
    Returns [(edge, dtype name, generator seed)]."""
    rng = SplitMix64(seed)
    kinds = ("float32", "uint16", "uint8")
    return [(sizes[rng.integer(0, len(sizes) - 1)], kinds[i % 3], seed + 1 + i) for i in range(n)]


# kvs::UnstructuredVolumeObject::CellType
TETRAHEDRA, HEXAHEDRA, PYRAMID, PRISM = 1, 2, 5, 7


def unstructured_mesh(dims, cell_type, seed=0, jitter=0.2, nblobs=4, dtype=np.float32):
    """An unstructured volume cut from a (nx, ny, nz) node grid with seeded
    jitter: per grid cube 6 tetrahedra around its main diagonal, 1 hexahedron,
    2 prisms, or 6 pyramids with an extra apex node at the cube centre.  Node
    values are the Gaussian-blob field of blobs() at the node positions.
    Returns (values [nnodes], coords float32 [nnodes, 3], connections uint32
    [ncells * nodes per cell]) in kvs::UnstructuredVolumeObject's layout."""
    if isinstance(dims, int):
        dims = (dims, dims, dims)
    nx, ny, nz = dims
    rng = SplitMix64(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float64)
    pos += (rng.bulk_uniform(pos.size).reshape(pos.shape) - 0.5) * jitter

    def node(i, j, k):
        return (k * ny + j) * nx + i

    cz, cy, cx = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()
    # cube corners in marching-cubes order: 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0), 4..7 the same at z+1
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    corner = np.stack([node(cx + a, cy + b, cz + c) for a, b, c in off], axis=1)
    if cell_type == TETRAHEDRA:
        tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
        conn = np.concatenate([corner[:, list(t)] for t in tets], axis=1)
    elif cell_type == HEXAHEDRA:
        # connections 4..7 become the mapper's local nodes 0..3 (MarchingHexahedra.cpp:163-170)
        conn = corner[:, [4, 5, 6, 7, 0, 1, 2, 3]]
    elif cell_type == PRISM:
        conn = np.concatenate([corner[:, [0, 1, 2, 4, 5, 6]], corner[:, [0, 2, 3, 4, 6, 7]]], axis=1)
    elif cell_type == PYRAMID:
        centre = pos.shape[0] + np.arange(corner.shape[0])
        cpos = np.mean(pos[corner], axis=1)
        pos = np.concatenate([pos, cpos])
        faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
        conn = np.concatenate([np.concatenate([centre[:, None], corner[:, list(f)]], axis=1) for f in faces], axis=1)
    else:
        raise ValueError(f"cell type {cell_type}")
    _, params = blob_params(seed + 7, nblobs)
    p = pos / np.array([max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1)], dtype=np.float64)
    f = np.zeros(pos.shape[0])
    for c, s in params:
        f += np.exp(-np.sum((p - np.asarray(c)) ** 2, axis=1) / (2 * s * s))
    if np.dtype(dtype).kind in "ui":
        top = float(np.iinfo(dtype).max if np.dtype(dtype).itemsize < 4 else 1000)
        values = np.rint(f / f.max() * top).astype(dtype)
    else:
        values = f.astype(dtype)
    return values, pos.astype(np.float32), conn.reshape(-1).astype(np.uint32)
