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
