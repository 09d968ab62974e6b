"""The seeded scene generator: counts, seeding pattern, divergence-free
velocities, determinism."""
import numpy as np

from gridfluidsim3d_amd import scenes


def test_sphere64_counts_and_seeding():
    sc = scenes.sphere64()
    assert sc.dims == (64, 64, 64) and sc.dx == 0.125
    # cells whose centre lies inside the radius-2.5 sphere, 8 particles each
    cx, cy, cz = scenes.cell_centres(sc.dims, sc.dx)
    ncell = int(((cx - 4) ** 2 + (cy - 4) ** 2 + (cz - 4) ** 2 < 2.5 ** 2).sum())
    assert sc.n == 8 * ncell == 268416
    # each particle sits within jitter of one of its cell's 8 sub-points
    # (centre +- dx/4), fluidsimulation.cpp:1223-1245
    rel = sc.pos / sc.dx - np.floor(sc.pos / sc.dx)
    off = np.abs(rel - 0.5)
    assert np.all(np.abs(off - 0.25) <= 0.025 + 1e-5)
    assert abs(float(np.std(sc.vel)) - 1.0) < 0.01


def test_curl_noise_is_divergence_free_and_scaled():
    rng = np.random.default_rng(0)
    p = rng.uniform(0, 4, size=(500, 3)).astype(np.float32)
    h = np.float32(2e-3)
    pts = [p]
    for a in range(3):
        e = np.zeros(3, np.float32)
        e[a] = h
        pts += [p + e, p - e]
    v = scenes.curl_noise(np.concatenate(pts), 4.0, 2.0, np.random.default_rng(1))
    v = v.astype(np.float64).reshape(7, 500, 3)
    div = sum((v[1 + 2 * a][:, a] - v[2 + 2 * a][:, a]) / (2 * h) for a in range(3))
    grad = np.abs((v[1][:, 0] - v[2][:, 0]) / (2 * h)).mean()
    assert np.abs(div).max() < 5e-3 * grad + 5e-3
    assert abs(np.sqrt((v ** 2).sum(-1).mean()) - 2.0) < 0.05


def test_generators_are_deterministic():
    a = scenes.random_blobs(3, (16, 12, 10), 0.1)
    b = scenes.random_blobs(3, (16, 12, 10), 0.1)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)
    c = scenes.random_blobs(4, (16, 12, 10), 0.1)
    assert not np.array_equal(a.pos[: min(a.n, c.n)], c.pos[: min(a.n, c.n)])
