"""Parity of the fused step on every benchmark workload, at full size where
the CPU path finishes in seconds and on a slab of the workload where it does
not.

* sphere64 (64^3, 268,416 particles): the plain step, both stencils, both
  output orders, against the reference's own compiled CPU code (the C
  restatement where that was not built).
* dambreak128 (128^3, 4.0M particles): trilinear at full size against the
  CPU path; tricubic on the x-slab [0, 24) of the same scene.
* pool256 (256^3, 43.2M particles): the x-slab [0, 24) of the scene
  (24 x 256 x 256 cells, 2.5M particles), both stencils.
* batch64 (64 scenes, 43.7M particles): the batched step gives every scene
  the bits of a step on that scene alone, and the eight smallest scenes
  match the CPU path.

Bar: max|gpu - cpu| / max|cpu| <= 1e-5 per field (grids after the body
force, snapshots, weights, particle velocities and positions), and the
displacement x' - x within 1e-4 of its own size.
"""
import numpy as np
import pytest

from conftest import norm_rel
from test_gpu_parity import TRICUBIC, TRILINEAR, _check_step, _run_step

pytestmark = pytest.mark.gpu

BF = (0.0, -9.8, 0.0)


def _cpu(orc, ref):
    return ref if ref is not None else orc


def _slab(sc, cells):
    """The particles of scene sc with x < cells*dx, on a cells x nj x nk grid."""
    keep = sc.pos[:, 0] < np.float32(cells * sc.dx)
    return (np.ascontiguousarray(sc.pos[keep]), np.ascontiguousarray(sc.vel[keep]),
            (cells, sc.dims[1], sc.dims[2]))


def _compare(ctx, cpu, pos, vel, dims, dx, dt, stencil, keep):
    print("[configs] %s particles, dims %s, stencil %d, keep %d" % (len(pos), dims, stencil,
                                                                     keep), flush=True)
    exp = cpu.flip_step(pos, vel, dims, dx, dt, BF, 0.05, stencil)
    res = _run_step(ctx, pos, vel, dims, dx, dt, BF, 0.05, stencil, int(keep))
    got_pos, _ = _check_step(res, exp, keep)
    assert norm_rel(got_pos - pos, exp["pos"] - pos) <= 1e-4


@pytest.fixture(scope="module")
def sphere():
    from gridfluidsim3d_amd import scenes
    return scenes.sphere64()


@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
@pytest.mark.parametrize("keep", [False, True])
def test_sphere64_plain_step(ctx, orc, ref, sphere, stencil, keep):
    sc = sphere
    _compare(ctx, _cpu(orc, ref), sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, stencil, keep)


@pytest.fixture(scope="module")
def dambreak():
    from gridfluidsim3d_amd import scenes
    return scenes.dambreak128()


def test_dambreak128_full_trilinear(ctx, orc, ref, dambreak):
    sc = dambreak
    assert sc.n > 4_000_000
    _compare(ctx, _cpu(orc, ref), sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, TRILINEAR, False)


def test_dambreak128_slab_tricubic(ctx, orc, ref, dambreak):
    pos, vel, dims = _slab(dambreak, 24)
    _compare(ctx, _cpu(orc, ref), pos, vel, dims, dambreak.dx, dambreak.dt, TRICUBIC, False)


@pytest.fixture(scope="module")
def pool_slab():
    from gridfluidsim3d_amd import scenes
    sc = scenes.pool256()
    pos, vel, dims = _slab(sc, 24)
    return pos, vel, dims, sc.dx, sc.dt


@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
def test_pool256_slab(ctx, orc, ref, pool_slab, stencil):
    pos, vel, dims, dx, dt = pool_slab
    assert len(pos) > 2_000_000
    _compare(ctx, _cpu(orc, ref), pos, vel, dims, dx, dt, stencil, False)


def test_batch64_batched_matches_alone_and_cpu(ctx, orc, ref):
    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    from test_gpu_parity import _state_result
    scs = scenes.batch64()
    assert len(scs) == 64
    mk = lambda s: gf.FlipState(s.pos, s.vel, s.dims, s.dx, track_ids=True, keep_old=True,
                                keep_weights=True)
    batch = [mk(s) for s in scs]
    print("[configs] batch64: %d particles" % sum(s.n for s in scs), flush=True)
    ctx.flip_step_batched(batch, [s.dt for s in scs], BF, 0.05, TRILINEAR, gf.ORDER_SORTED)
    torch.cuda.synchronize()
    cpu = _cpu(orc, ref)
    smallest = set(np.argsort([s.n for s in scs])[:8].tolist())
    for i, s in enumerate(scs):
        alone = mk(s)
        ctx.flip_step(alone, s.dt, BF, 0.05, TRILINEAR, gf.ORDER_SORTED)
        torch.cuda.synchronize()
        ra, rb = _state_result(batch[i]), _state_result(alone)
        for key in ra:
            assert np.array_equal(ra[key], rb[key], equal_nan=True), (i, key)
        if i in smallest:
            exp = cpu.flip_step(s.pos, s.vel, s.dims, s.dx, s.dt, BF, 0.05, TRILINEAR)
            _check_step(ra, exp, False)
        del alone
