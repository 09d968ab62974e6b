"""The brick sort and the scatter P2G of the default fused step.

* The particles come out of a GFSX_ORDER_SORTED step in the order of a
  stable counting sort by P2G brick (4^3 cells; the pad cells just outside
  the grid belong to the edge bricks; anything further out, and NaN, to one
  overflow bin after the bricks), bit-exact against numpy's stable argsort
  of the same keys: for inputs close to brick order (few runs per wave),
  shuffled inputs (a run per particle), a brick fed by thousands of waves
  and a crowded overflow bin (the work-group path for bins of many runs),
  and a batch of scenes.
* P2G sums without a bias: on faces reached only by slow particles in a
  brick that also holds a fast one, each face's value keeps FP32 precision
  against the FP64 evaluation of the same step.
* A velocity that is not finite marks exactly the faces the reference's
  float sums mark (the nodes inside its radius), against the oracle.
"""
import numpy as np
import pytest

from conftest import NORM_REL_TOL, norm_rel
from test_gpu_parity import TRILINEAR, _check_step, _run_step

pytestmark = pytest.mark.gpu

BF = (0.0, -9.8, 0.0)


def brick_keys(pos, dims, dx):
    """Local bin of every particle: (G2P brick << 3) | octant of its P2G brick
    in that G2P brick, or the overflow bin 8 * gbx * gby * gbz."""
    dims = np.array(dims)
    nb = (dims + 3) // 4
    gb = (nb + 1) // 2
    c = np.floor(pos.astype(np.float64) * (1.0 / dx))
    with np.errstate(invalid="ignore"):
        ok = np.all((c >= -1) & (c <= dims), axis=1)
    c = np.where(ok[:, None], c, 0).astype(np.int64)
    b = np.clip(c, 0, dims - 1) >> 2
    g = (b[:, 0] >> 1) + gb[0] * ((b[:, 1] >> 1) + gb[1] * (b[:, 2] >> 1))
    key = (g << 3) | (b[:, 0] & 1) | ((b[:, 1] & 1) << 1) | ((b[:, 2] & 1) << 2)
    return np.where(ok, key, 8 * gb[0] * gb[1] * gb[2])


def _sorted_ids(ctx, pos, vel, dims, dx, steps=1):
    import torch

    import gridfluidsim3d_amd as gf
    st = gf.FlipState(pos, vel, dims, dx, track_ids=True)
    orders = []
    for _ in range(steps):
        before = st.positions().cpu().numpy()
        ids_before = st.particle_ids().cpu().numpy().astype(np.int64)
        ctx.flip_step(st, 1 / 240, BF, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
        torch.cuda.synchronize()
        ids = st.particle_ids().cpu().numpy().astype(np.int64)
        # slot order of this step's input: a stable sort of its keys
        expect = ids_before[np.argsort(brick_keys(before, dims, dx), kind="stable")]
        orders.append((ids, expect))
    return orders


def _scene(kind, rng):
    from gridfluidsim3d_amd import scenes
    if kind == "blobs":
        sc = scenes.random_blobs(3, (37, 29, 21), 0.1, amplitude=3.0)
        return sc.pos, sc.vel, sc.dims, sc.dx
    if kind == "shuffled":
        sc = scenes.random_blobs(4, (24, 40, 16), 0.05, amplitude=2.0)
        p = rng.permutation(sc.n)
        return sc.pos[p], sc.vel[p], sc.dims, sc.dx
    if kind == "crowded":
        # one brick holding 60,000 particles spread over the input, and a
        # crowded overflow bin: far particles interleaved everywhere
        dims, dx = (16, 12, 20), 0.1
        base = rng.uniform(-0.05, np.array(dims) * dx + 0.05, size=(40000, 3))
        brick = (np.array([1, 1, 2]) * 4 + rng.uniform(0, 4, size=(60000, 3))) * dx
        far = rng.uniform(-30, 30, size=(30000, 3))
        pos = np.concatenate([base, brick, far]).astype(np.float32)
        pos[rng.integers(0, len(pos), 50)] = np.nan
        p = rng.permutation(len(pos))
        return pos[p], rng.normal(size=pos.shape).astype(np.float32), dims, dx
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["blobs", "shuffled", "crowded"])
def test_brick_sort_is_a_stable_sort(ctx, kind):
    rng = np.random.default_rng(17)
    pos, vel, dims, dx = _scene(kind, rng)
    print("[brick] %s: %d particles" % (kind, len(pos)), flush=True)
    for step, (ids, expect) in enumerate(_sorted_ids(ctx, pos, vel, dims, dx, steps=3)):
        assert np.array_equal(ids, expect), (kind, step)


def test_brick_sort_batched(ctx):
    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    rng = np.random.default_rng(5)
    specs = []
    for k, dims in enumerate([(9, 13, 7), (33, 11, 16), (5, 6, 7), (17, 17, 40)]):
        sc = scenes.random_blobs(100 + k, dims, 0.1, amplitude=2.0)
        p = rng.permutation(sc.n) if k % 2 else np.arange(sc.n)
        specs.append((sc.pos[p], sc.vel[p], dims, 0.1))
    specs.insert(1, (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), (4, 4, 4), 0.1))
    states = [gf.FlipState(p, v, d, h, track_ids=True) for p, v, d, h in specs]
    ctx.flip_step_batched(states, 1 / 240, BF, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
    torch.cuda.synchronize()
    for (p, v, d, h), st in zip(specs, states):
        ids = st.particle_ids().cpu().numpy().astype(np.int64)
        assert np.array_equal(ids, np.argsort(brick_keys(p, d, h), kind="stable"))


def test_scatter_keeps_precision_beside_a_fast_particle(ctx, orc64):
    """Slow particles (|v| ~ 1e-3) and one fast one (|v| = 3) in one brick:
    every face the slow particles reach keeps its FP32 precision (a bias of
    the brick's largest |v| would leave an absolute error of ~2^-24 * 3)."""
    rng = np.random.default_rng(9)
    dims, dx = (8, 8, 8), 0.1
    pos = rng.uniform(0.05, 0.75, size=(3000, 3)).astype(np.float32)
    vel = (1e-3 * rng.normal(size=pos.shape)).astype(np.float32)
    pos[0] = (0.21, 0.22, 0.23)
    vel[0] = (3.0, -3.0, 3.0)
    res = _run_step(ctx, pos, vel, dims, dx, 1 / 240, (0.0, 0.0, 0.0), 0.05, TRILINEAR, 1)
    e64 = orc64.flip_step(pos, vel, dims, dx, 1 / 240, (0.0, 0.0, 0.0), 0.05, TRILINEAR)
    for key, wkey in (("u", "wu"), ("v", "wv"), ("w", "ww")):
        got, exp = res[key].astype(np.float64), e64[key]
        m = (e64[wkey] > 1e-3) & (np.abs(exp) > 1e-5)
        rel = np.abs(got[m] - exp[m]) / np.abs(exp[m])
        print("[brick] %s: %d faces, worst per-face relative error %.2e" % (key, m.sum(), rel.max()))
        assert rel.max() <= 2e-5, key


@pytest.mark.parametrize("variant", [4, 0])
def test_nonfinite_velocities_mark_the_oracles_faces(ctx, orc, variant):
    rng = np.random.default_rng(41)
    dims, dx = (9, 11, 7), 0.1
    W = np.array(dims) * dx
    pos = rng.uniform(-0.15, W + 0.15, size=(6000, 3)).astype(np.float32)
    vel = (3.0 * rng.normal(size=pos.shape)).astype(np.float32)
    vel[10, 0] = np.inf
    vel[20, 1] = -np.inf
    vel[30, 2] = np.nan
    vel[40] = (np.inf, -np.inf, np.nan)
    exp = orc.flip_step(pos, vel, dims, dx, 0.01, BF, 0.05, TRILINEAR)
    old = ctx.get_option("p2g_variant")
    ctx.set_option("p2g_variant", variant)
    try:
        res = _run_step(ctx, pos, vel, dims, dx, 0.01, BF, 0.05, TRILINEAR, 1)
    finally:
        ctx.set_option("p2g_variant", old)
    for key in ("u", "v", "w", "u_old", "v_old", "w_old", "wu", "wv", "ww"):
        a, b = res[key], exp[key]
        assert np.array_equal(np.isnan(a), np.isnan(b)), key
        assert np.array_equal(np.isposinf(a), np.isposinf(b)), key
        assert np.array_equal(np.isneginf(a), np.isneginf(b)), key
        f = np.isfinite(b)
        assert norm_rel(a[f], b[f]) <= NORM_REL_TOL, key
    assert not np.isfinite(res["u"]).all()


def test_step_matches_oracle_after_sorted_steps(ctx, orc):
    """Three sorted-order steps in a row (each input in the previous step's
    brick order) against the oracle stepping the same particles."""
    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(13, (26, 19, 23), 0.1, amplitude=3.0)
    st = gf.FlipState(sc.pos, sc.vel, sc.dims, sc.dx, track_ids=True, keep_old=True,
                      keep_weights=True)
    pos, vel = sc.pos.copy(), sc.vel.copy()
    for _ in range(3):
        ctx.flip_step(st, sc.dt, BF, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
        torch.cuda.synchronize()
        ids = st.particle_ids().cpu().numpy().astype(np.int64)
        exp = orc.flip_step(pos, vel, sc.dims, sc.dx, sc.dt, BF, 0.05, TRILINEAR)
        res = dict(pos=st.positions().cpu().numpy(), vel=st.velocities().cpu().numpy(),
                   u=st.grid[0].cpu().numpy(), v=st.grid[1].cpu().numpy(),
                   w=st.grid[2].cpu().numpy(), u_old=st.grid_old[0].cpu().numpy(),
                   v_old=st.grid_old[1].cpu().numpy(), w_old=st.grid_old[2].cpu().numpy(),
                   wu=st.weights[0].cpu().numpy(), wv=st.weights[1].cpu().numpy(),
                   ww=st.weights[2].cpu().numpy(), id=ids)
        got_pos, got_vel = _check_step(res, exp, False)
        # next step from the GPU's own state, on both sides
        pos, vel = got_pos, got_vel
