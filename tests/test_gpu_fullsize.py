"""The fused step at the benchmark's full size (the 256^3 pool scene,
43,194,016 particles), checked through properties that do not need a
reference run at that size:

* two runs of the same step give identical bits (grids, snapshot, weights,
  positions, velocities, particle labels);
* the particle labels are a permutation and, with GFSX_ORDER_SORTED, the
  output order is a stable sort of the input by P2G brick (4^3 cells, the
  8 bricks of every 8^3-cell G2P brick consecutive, the pad cells in the
  edge bricks, the particles beyond them last): keys non-decreasing, and
  input indices increasing within each brick;
* P2G is exactly linear in the velocities: with the body force off,
  doubling every velocity doubles the pre-force snapshot bit for bit and
  leaves the weights unchanged (the scatter's fixed-point terms are the same
  integers at the doubled scale).
The sizes the oracle can check directly are covered by test_gpu_parity.py.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    from gridfluidsim3d_amd import scenes
    return scenes.pool256()


def _step(ctx, sc, vel, body_force):
    import gridfluidsim3d_amd as gf
    st = gf.FlipState(sc.pos, vel, sc.dims, sc.dx, track_ids=True, keep_old=True,
                      keep_weights=True)
    ctx.flip_step(st, sc.dt, body_force, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
    torch.cuda.synchronize()
    return st


def test_fullsize_step_deterministic_and_sorted(ctx, pool):
    a = _step(ctx, pool, pool.vel, pool.body_force)
    b = _step(ctx, pool, pool.vel, pool.body_force)
    for ta, tb in zip(a.grid + a.grid_old + a.weights, b.grid + b.grid_old + b.weights):
        assert torch.equal(ta, tb)
    for c in range(3):
        assert torch.equal(a.cur_x[c], b.cur_x[c])
        assert torch.equal(a.cur_v[c], b.cur_v[c])
    ids = a.particle_ids().long()
    assert torch.equal(a.particle_ids(), b.particle_ids())
    n = pool.pos.shape[0]
    assert torch.equal(torch.sort(ids).values, torch.arange(n, device=ids.device))
    # brick of every input position, in output order
    p = torch.as_tensor(pool.pos, device=ids.device)[ids].double()
    dims = torch.tensor(pool.dims, device=ids.device)
    cell = torch.floor(p * (1.0 / pool.dx)).long()
    inside = ((cell >= -1) & (cell <= dims)).all(1)
    b = torch.minimum(torch.clamp(cell, min=0), dims - 1) >> 2
    gb = ((dims + 3) // 4 + 1) // 2
    g = (b[:, 0] >> 1) + gb[0] * ((b[:, 1] >> 1) + gb[1] * (b[:, 2] >> 1))
    key = (g << 3) | (b[:, 0] & 1) | ((b[:, 1] & 1) << 1) | ((b[:, 2] & 1) << 2)
    key = torch.where(inside, key, torch.full_like(key, int(8 * gb.prod())))
    assert bool((key[1:] >= key[:-1]).all())
    same = key[1:] == key[:-1]
    assert bool((ids[1:][same] > ids[:-1][same]).all())
    # every particle advanced: no NaN in the outputs of this scene
    for c in range(3):
        assert bool(torch.isfinite(a.cur_x[c]).all()) and bool(torch.isfinite(a.cur_v[c]).all())


def test_fullsize_p2g_linear_in_velocity(ctx, pool):
    zero = (0.0, 0.0, 0.0)
    a = _step(ctx, pool, pool.vel, zero)
    b = _step(ctx, pool, (2.0 * pool.vel).astype(np.float32), zero)
    for c in range(3):
        assert torch.equal(b.grid_old[c], 2.0 * a.grid_old[c])
        assert torch.equal(b.grid[c], 2.0 * a.grid[c])
        assert torch.equal(b.weights[c], a.weights[c])
