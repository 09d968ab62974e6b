"""Seeded synthetic scenes for the benchmark configurations and the tests.

Particles are seeded like FluidSimulation::_addMarkerParticlesToCell
(fluidsimulation.cpp:1223-1245): in every fluid cell, the 8 points at the
cell centre +- dx/4 on each axis, each moved by a uniform jitter of
+-0.025*dx (0.25 * jitter factor 0.1, fluidsimulation.cpp:1219-1221 and
fluidsimulation.h:1111).  Velocities are either Gaussian or a divergence-free
curl-noise field v = curl(psi), psi_c = sum_m a_m sin(k_m . x + phi_m),
differentiated analytically and scaled to a given RMS speed.

Randomness comes from numpy's PCG64 seeded per scene (seed 1000 + config
number for the benchmark configurations), so CPU and GPU runs see identical
inputs.
"""
import math

import numpy as np

# corner order of fluidsimulation.cpp:1227-1236
_CORNERS = np.array([(-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1),
                     (-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)], np.float64)


class SceneSpec:
    """A particle set on a MAC grid: pos/vel are (n, 3) float32."""

    def __init__(self, name, dims, dx, pos, vel, dt, body_force=(0.0, -9.8, 0.0)):
        self.name = name
        self.dims = tuple(int(d) for d in dims)
        self.dx = float(dx)
        self.pos = pos
        self.vel = vel
        self.dt = float(dt)
        self.body_force = tuple(float(b) for b in body_force)

    @property
    def n(self):
        return int(self.pos.shape[0])

    def describe(self):
        ni, nj, nk = self.dims
        return dict(workload=self.name, grid=[ni, nj, nk], dx=self.dx,
                    particles=self.n, dt=self.dt)


def cell_centres(dims, dx):
    ni, nj, nk = dims
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    return ((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx)


def seed_cells(mask, dx, rng, ppc=8, chunk=1 << 22):
    """Jittered particles in every True cell of mask (shape (nk, nj, ni)),
    cells in flat (i-fastest) order, ppc of the 8 corner points per cell."""
    idx = np.flatnonzero(mask.reshape(-1))
    nk, nj, ni = mask.shape
    q = 0.25 * dx
    jit = 0.25 * 0.1 * dx
    out = np.empty((idx.size * ppc, 3), np.float32)
    corners = (_CORNERS[:ppc] * q).astype(np.float32)
    for s in range(0, idx.size, chunk):
        c = idx[s:s + chunk]
        i = c % ni
        j = (c // ni) % nj
        k = c // (ni * nj)
        centre = np.stack([(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx],
                          axis=1).astype(np.float32)
        p = out[s * ppc:(s + c.size) * ppc].reshape(c.size, ppc, 3)
        rng.random(out=p, dtype=np.float32)
        p *= np.float32(2 * jit)
        p -= np.float32(jit)
        p += corners[None, :, :]
        p += centre[:, None, :]
    return out


def curl_noise(pos, extent, amplitude, rng, modes=16, chunk=1 << 22):
    """Divergence-free velocity at pos (n, 3): curl of a sum of plane-wave
    streams, |k| uniform in [2pi, 8pi]/extent, scaled to RMS speed
    ``amplitude``."""
    n = pos.shape[0]
    ks, amps, phis = [], [], []
    for _c in range(3):
        d = rng.normal(size=(modes, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        mag = rng.uniform(2 * math.pi, 8 * math.pi, size=(modes, 1)) / extent
        ks.append(d * mag)
        amps.append(rng.normal(size=modes) / mag[:, 0])
        phis.append(rng.uniform(0, 2 * math.pi, size=modes))
    # all 3*modes plane waves at once: phase = pos @ K + phi, and
    # curl = cos(phase) @ C, where C maps each wave's cosine to the three
    # curl components: d psi_c / d x_d = sum_m a_cm k_cmd cos(...)
    K = np.concatenate(ks, axis=0).T.astype(np.float32)          # (3, 3m)
    phi = np.concatenate(phis).astype(np.float32)                # (3m,)
    C = np.zeros((3 * modes, 3))
    for c in range(3):
        g = amps[c][:, None] * ks[c]                              # (m, 3): d psi_c/d x_d
        rows = slice(c * modes, (c + 1) * modes)
        if c == 0:    # psi_x enters v_y (+d/dz) and v_z (-d/dy)
            C[rows, 1] += g[:, 2]
            C[rows, 2] -= g[:, 1]
        elif c == 1:  # psi_y enters v_z (+d/dx) and v_x (-d/dz)
            C[rows, 2] += g[:, 0]
            C[rows, 0] -= g[:, 2]
        else:         # psi_z enters v_x (+d/dy) and v_y (-d/dx)
            C[rows, 0] += g[:, 1]
            C[rows, 1] -= g[:, 0]
    C = C.astype(np.float32)
    vel = np.empty((n, 3), np.float32)
    for s in range(0, n, chunk):
        ph = pos[s:s + chunk].astype(np.float32) @ K
        ph += phi
        np.cos(ph, out=ph)
        vel[s:s + chunk] = ph @ C
    rms = math.sqrt(float(np.mean(np.sum(vel.astype(np.float64) ** 2, axis=1)))) if n else 1.0
    return (vel * np.float32(amplitude / rms if rms > 0 else 0.0)).astype(np.float32)


# ---------------------------------------------------------------- configs

def sphere64(seed=1001, dims=(64, 64, 64), dx=0.125, radius=2.5):
    """64^3, dx 0.125: a fluid sphere of radius 2.5 at the centre, 8 per
    cell, Gaussian velocities (sigma 1 m/s), dt = 1/120."""
    rng = np.random.default_rng(seed)
    cx, cy, cz = cell_centres(dims, dx)
    ext = np.array(dims) * dx / 2
    mask = (cx - ext[0]) ** 2 + (cy - ext[1]) ** 2 + (cz - ext[2]) ** 2 < radius ** 2
    pos = seed_cells(mask, dx, rng)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    return SceneSpec("sphere%d" % dims[0], dims, dx, pos, vel, 1.0 / 120.0)


def dambreak128(seed=1002, dims=(128, 128, 128), dx=0.0625):
    """128^3, dx 0.0625: a block filling x<0.4, y<0.6 (all z) of the domain,
    8 per cell, curl noise at 2 m/s, dt = 1/240."""
    rng = np.random.default_rng(seed)
    cx, cy, cz = cell_centres(dims, dx)
    W = np.array(dims) * dx
    mask = (cx < 0.4 * W[0]) & (cy < 0.6 * W[1]) & (cz < W[2])
    pos = seed_cells(mask, dx, rng)
    vel = curl_noise(pos, W[0], 2.0, rng)
    return SceneSpec("dambreak%d" % dims[0], dims, dx, pos, vel, 1.0 / 240.0)


def pool256(seed=1003, dims=(256, 256, 256), dx=0.03125, nspheres=12):
    """256^3, dx 1/32: a bottom pool 20% deep plus 12 seeded spheres (radius
    0.05-0.2 of the domain width), 8 per cell, curl noise at 3 m/s,
    dt = 1/240."""
    rng = np.random.default_rng(seed)
    W = np.array(dims) * dx
    mask = np.zeros(dims[::-1], bool)
    nk, nj, ni = mask.shape
    jy = np.arange(nj)
    mask[:, (jy + 0.5) * dx < 0.2 * W[1], :] = True
    centres_i = (np.arange(ni) + 0.5) * dx
    centres_j = (np.arange(nj) + 0.5) * dx
    centres_k = (np.arange(nk) + 0.5) * dx
    for _s in range(nspheres):
        r = rng.uniform(0.05, 0.2) * W[0]
        c = np.array([rng.uniform(r, W[a] - r) for a in range(3)])
        dxs = (centres_i - c[0]) ** 2
        dys = (centres_j - c[1]) ** 2
        dzs = (centres_k - c[2]) ** 2
        inside = dzs[:, None, None] + dys[None, :, None] + dxs[None, None, :] < r * r
        mask |= inside
    pos = seed_cells(mask, dx, rng)
    vel = curl_noise(pos, W[0], 3.0, rng)
    return SceneSpec("pool%d" % dims[0], dims, dx, pos, vel, 1.0 / 240.0)


def random_blobs(seed, dims, dx, ppc=8, amplitude=2.0, fill=(0.1, 0.6),
                 dt=1.0 / 240.0):
    """A scene of random boxes and spheres covering a fraction of the domain
    drawn from ``fill``, ppc particles per cell, curl noise velocities."""
    rng = np.random.default_rng(seed)
    W = np.array(dims, np.float64) * dx
    cx, cy, cz = cell_centres(dims, dx)
    mask = np.zeros(dims[::-1], bool)
    target = rng.uniform(*fill)
    for _ in range(64):
        if mask.mean() >= target:
            break
        if rng.uniform() < 0.5:
            lo = rng.uniform(0, 0.8, size=3) * W
            hi = lo + rng.uniform(0.1, 0.5, size=3) * W
            mask |= ((cx >= lo[0]) & (cx < hi[0]) & (cy >= lo[1]) & (cy < hi[1]) &
                     (cz >= lo[2]) & (cz < hi[2]))
        else:
            r = rng.uniform(0.1, 0.3) * W.min()
            c = rng.uniform(0, 1, size=3) * W
            mask |= (cx - c[0]) ** 2 + (cy - c[1]) ** 2 + (cz - c[2]) ** 2 < r * r
    pos = seed_cells(mask, dx, rng, ppc=ppc)
    vel = curl_noise(pos, W.max(), amplitude, rng)
    return SceneSpec("blobs%dx%dx%d" % tuple(dims), dims, dx, pos, vel, dt)


def batch64(seed=1004, nscenes=64, dx=1.0 / 64.0, dt=1.0 / 240.0):
    """The batched workload: nscenes independent scenes, each grid axis drawn
    log-uniform from 32 to 128 cells (independently), 10-60% of the cells
    fluid as random boxes and spheres, 4-8 particles per cell, curl noise
    at 1-3 m/s.  Returns a list of SceneSpec."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nscenes):
        dims = tuple(int(round(np.exp(rng.uniform(np.log(32.0), np.log(128.0))))) for _ in range(3))
        ppc = int(rng.integers(4, 9))
        amp = float(rng.uniform(1.0, 3.0))
        sc = random_blobs(int(rng.integers(1 << 30)), dims, dx, ppc=ppc, amplitude=amp,
                          fill=(0.1, 0.6), dt=dt)
        sc.name = "batch%d_scene%d" % (nscenes, s)
        out.append(sc)
    return out


def slab(seed=1005, dims=(96, 96, 96), dx=1.0 / 96.0, ppc=8, y0=20, layers=27,
         amplitude=1.0, dt=1.0 / 240.0):
    """The differentiable workload: a jittered slab of ``layers`` cell layers
    from cell row y0 (all of x and z), ppc per cell, curl noise at
    ``amplitude`` m/s.  At 96^3 this is 1,990,656 particles."""
    rng = np.random.default_rng(seed)
    mask = np.zeros(dims[::-1], bool)
    mask[:, y0:y0 + layers, :] = True
    pos = seed_cells(mask, dx, rng, ppc=ppc)
    vel = curl_noise(pos, dims[0] * dx, amplitude, rng)
    return SceneSpec("slab%d" % dims[0], dims, dx, pos, vel, dt)


def slab_subset32(seed=1006):
    """The gradient-check subset: a 32^3 grid (dx 1/32) holding a 16x20x8-cell
    jittered block, 8 per cell (20,480 particles), curl noise at 1 m/s."""
    rng = np.random.default_rng(seed)
    dims, dx = (32, 32, 32), 1.0 / 32.0
    mask = np.zeros(dims[::-1], bool)
    mask[12:20, 6:26, 8:24] = True
    pos = seed_cells(mask, dx, rng)
    vel = curl_noise(pos, dims[0] * dx, 1.0, rng)
    return SceneSpec("slab_subset32", dims, dx, pos, vel, 1.0 / 240.0)


def density_target(dims, dx, seed, level=6.5, y_range=None):
    """A seeded smooth target for the density (splat weight) field on the
    cell centres, shaped (nk, nj, ni) float32: ``level`` times a Gaussian
    profile across y (centred on y_range, the slab's cell rows) times
    1 + 0.3 sin(.) sin(.) waves along x and z."""
    rng = np.random.default_rng(seed)
    cx, cy, cz = cell_centres(dims, dx)
    ext = np.array(dims, np.float64) * dx
    lo, hi = y_range if y_range is not None else (0, dims[1])
    yc = 0.5 * (lo + hi) * dx
    wy = 0.5 * (hi - lo) * dx
    kx, kz = rng.integers(1, 3, size=2)
    p1, p2 = rng.uniform(0, 2 * math.pi, size=2)
    prof = np.exp(-((cy - yc) / wy) ** 2)
    wave = 1.0 + 0.3 * np.sin(2 * math.pi * kx * cx / ext[0] + p1) * \
        np.sin(2 * math.pi * kz * cz / ext[2] + p2)
    return (level * prof * wave).astype(np.float32)


CONFIGS = {
    "sphere64": sphere64,
    "dambreak128": dambreak128,
    "pool256": pool256,
}
