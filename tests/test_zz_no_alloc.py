"""gfsx_flip_step_noalloc: the fused step with every workspace buffer in the
library's static device scratch gives the bits of gfsx_flip_step.

Two calls in a row on the same context, the second on a smaller scene (the
scratch keeps the first call's contents), each against gfsx_flip_step on a
copy of the same state; both stencils and output orders, with the snapshot
in the scratch (no u_old..) and in the caller's buffers.  A step whose
workspace exceeds the scratch is refused with a message before any work is
enqueued, and the next call still works.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BF = (0.0, -9.8, 0.0)


def _result(st, keep_old):
    import torch
    torch.cuda.synchronize()
    out = dict(pos=st.positions().cpu().numpy(), vel=st.velocities().cpu().numpy(),
               id=st.particle_ids().cpu().numpy(),
               u=st.grid[0].cpu().numpy(), v=st.grid[1].cpu().numpy(),
               w=st.grid[2].cpu().numpy())
    if keep_old:
        for nm, t in zip(("u_old", "v_old", "w_old"), st.grid_old):
            out[nm] = t.cpu().numpy()
        for nm, t in zip(("wu", "wv", "ww"), st.weights):
            out[nm] = t.cpu().numpy()
    return out


def _pair(ctx, sc, stencil, order, keep_old):
    import gridfluidsim3d_amd as gf
    mk = lambda: gf.FlipState(sc.pos, sc.vel, sc.dims, sc.dx, track_ids=True,
                              keep_old=keep_old, keep_weights=keep_old)
    a, b = mk(), mk()
    print("[no_alloc] %d particles, dims %s, stencil %d, order %d, keep_old %d"
          % (sc.n, sc.dims, stencil, order, keep_old), flush=True)
    ctx.flip_step_noalloc(a, sc.dt, BF, 0.05, stencil, order)
    ra = _result(a, keep_old)
    ctx.flip_step(b, sc.dt, BF, 0.05, stencil, order)
    rb = _result(b, keep_old)
    for key in rb:
        assert np.array_equal(ra[key], rb[key], equal_nan=True), key
    return a, b


@pytest.mark.parametrize("stencil", [0, 1])
@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_matches_step_two_calls(ctx, stencil, order):
    from gridfluidsim3d_amd import scenes
    big = scenes.random_blobs(501, (48, 40, 36), 1.0 / 40, amplitude=2.0)
    small = scenes.random_blobs(502, (20, 16, 12), 0.1, amplitude=2.0)
    assert small.n < big.n
    _pair(ctx, big, stencil, order, keep_old=False)
    _pair(ctx, small, stencil, order, keep_old=True)


def test_noalloc_second_step_from_first(ctx):
    """Two steps in a row on the same states (the scratch holds the previous
    step's sort and tiles when the second starts)."""
    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(503, (32, 24, 28), 1.0 / 32, amplitude=3.0)
    a, b = _pair(ctx, sc, 0, 0, keep_old=True)
    ctx.flip_step_noalloc(a, sc.dt, BF, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
    ctx.flip_step(b, sc.dt, BF, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
    ra, rb = _result(a, True), _result(b, True)
    for key in rb:
        assert np.array_equal(ra[key], rb[key], equal_nan=True), key


def test_noalloc_refuses_oversize_then_works(ctx):
    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    dims = (160, 160, 160)
    n = 2_000_000
    rng = np.random.default_rng(7)
    pos = (rng.uniform(size=(n, 3)) * np.array(dims) * 0.01).astype(np.float32)
    vel = np.zeros((n, 3), np.float32)
    st = gf.FlipState(pos, vel, dims, 0.01)
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        ctx.flip_step_noalloc(st, 1 / 240, BF, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
    del st
    small = scenes.random_blobs(504, (16, 16, 16), 1.0 / 16, amplitude=1.0)
    _pair(ctx, small, 0, 1, keep_old=True)
