"""gfsx_remove_particles against the oracle's removal (pinned by brute force
in test_removal.py): the same particles kept, copied bit for bit in input
order, ids carried, for caps from 1 to above every cell's count."""
import numpy as np
import pytest

from test_removal import crowded_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cap,seed,with_solid", [(20, 0, True), (8, 99, False), (1, 5, True),
                                                 (1000, 0, False)])
def test_removal_matches_oracle(ctx, orc, cap, seed, with_solid):
    import gridfluidsim3d_amd as gf
    dims, dx, pos, solid = crowded_scene()
    sol = solid if with_solid else None
    rng = np.random.default_rng(seed)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    st = gf.FlipState(pos, vel, dims, dx, track_ids=True, solid=sol)
    kept = ctx.remove_particles(st, cap, seed)
    keep = orc.remove_particles(pos, dims, dx, sol, cap, seed).astype(bool)
    assert kept == keep.sum() == st.n
    assert np.array_equal(st.positions().cpu().numpy(), pos[keep])
    assert np.array_equal(st.velocities().cpu().numpy(), vel[keep])
    assert np.array_equal(st.particle_ids().cpu().numpy(), np.flatnonzero(keep))
    if cap >= 1000 and sol is None:
        assert kept == len(pos)


def test_removal_then_step(ctx, orc):
    """A step on the shrunk state runs and matches the oracle."""
    import gridfluidsim3d_amd as gf
    from conftest import NORM_REL_TOL, norm_rel
    dims, dx, pos, solid = crowded_scene(5)
    vel = np.random.default_rng(5).normal(size=pos.shape).astype(np.float32)
    st = gf.FlipState(pos, vel, dims, dx)
    kept = ctx.remove_particles(st, 10, 1)
    assert kept < len(pos)
    p0 = st.positions().cpu().numpy()
    v0 = st.velocities().cpu().numpy()
    ctx.flip_step(st, 1.0 / 120.0, order=gf.ORDER_KEEP)
    exp = orc.flip_step(p0, v0, dims, dx, 1.0 / 120.0)
    assert norm_rel(st.velocities().cpu().numpy(), exp["vel"]) <= NORM_REL_TOL


def test_removal_rejects_bad_cap(ctx):
    import gridfluidsim3d_amd as gf
    dims, dx, pos, _ = crowded_scene()
    st = gf.FlipState(pos, pos, dims, dx)
    with pytest.raises(RuntimeError, match="max_per_cell"):
        ctx.remove_particles(st, 0)


def _splitmix_np(seed, idx):
    """SplitMix64 of (seed, index) over a uint64 array (wrapping products)."""
    z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


@pytest.mark.parametrize("cap", [100, 150_000])
def test_removal_one_clumped_cell(ctx, cap):
    """200,000 particles in one cell among ordinary ones: the cell keeps
    exactly the cap particles of smallest key (numpy over all its keys), in
    time linear in its count."""
    import gridfluidsim3d_amd as gf
    from test_removal import splitmix
    idx = np.array([0, 1, 7, 123456], np.uint64)
    assert [int(k) for k in _splitmix_np(42, idx)] == [splitmix(42, int(i)) for i in idx]
    rng = np.random.default_rng(9)
    dims, dx = (16, 16, 16), 0.1
    bg = rng.uniform(0, 1.6, size=(30_000, 3))
    clump = (np.array([5, 6, 7]) + rng.uniform(0.01, 0.99, size=(200_000, 3))) * dx
    pos = np.concatenate([bg, clump]).astype(np.float32)
    rng.shuffle(pos)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    c = np.floor(pos.astype(np.float64) / dx).astype(np.int64)
    flat = c[:, 0] + 16 * (c[:, 1] + 16 * c[:, 2])
    hot = 5 + 16 * (6 + 16 * 7)
    keys = _splitmix_np(3, np.arange(len(pos)))
    keep = np.ones(len(pos), bool)
    counts = np.bincount(flat, minlength=16 ** 3)
    for cell in np.flatnonzero(counts > cap):
        members = np.flatnonzero(flat == cell)
        keep[members] = False
        keep[members[np.argsort(keys[members])[:cap]]] = True
    assert counts[hot] >= 200_000
    st = gf.FlipState(pos, vel, dims, dx, track_ids=True)
    kept = ctx.remove_particles(st, cap, 3)
    assert kept == keep.sum() == st.n
    assert np.array_equal(st.particle_ids().cpu().numpy(), np.flatnonzero(keep))
    assert np.array_equal(st.positions().cpu().numpy(), pos[keep])
