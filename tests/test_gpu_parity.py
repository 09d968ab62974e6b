"""GPU parity of every entry point against the CPU oracle and the golden
vectors written by the reference's own sources.

Bars (written per check below):
  * integer work (the cell sort) is bit-exact;
  * FP32 fields: max|gpu - ref| / max|ref| <= 1e-5 (NORM_REL_TOL), and the
    GPU error against the FP64 oracle is no worse than 10x the FP32 oracle's;
  * the P2G sum order is fixed, so two runs are bit-identical.
"""
import numpy as np
import pytest

from conftest import NORM_REL_TOL, load_golden, norm_rel

pytestmark = pytest.mark.gpu

TRILINEAR, TRICUBIC = 0, 1


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _np(t):
    return t.detach().cpu().numpy()


# --------------------------------------------------------------- cell sort

@pytest.mark.parametrize("seed,n", [(1, 1), (2, 1000), (3, 20000)])
def test_cell_sort_bit_exact(ctx, orc, seed, n):
    rng = np.random.default_rng(seed)
    dims = (13, 7, 9)
    dx = 0.1
    pos = rng.uniform(-0.2, np.array(dims) * dx + 0.2, size=(n, 3)).astype(np.float32)
    # some exact cell boundaries, and clustered particles (large bins)
    pos[: n // 10] = np.round(pos[: n // 10] / dx) * dx
    pos[n // 10: n // 5] = pos[0]
    perm_ref, start_ref = orc.cell_sort(pos, dx, dims)
    x, y, z = (_dev(pos[:, a].copy()) for a in range(3))
    perm, start = ctx.cell_sort_device(x, y, z, dx, dims)
    perm = _np(perm)
    start = _np(start)
    ncell = dims[0] * dims[1] * dims[2]
    assert np.array_equal(start, start_ref.astype(np.int64))
    m = start_ref[ncell]
    assert np.array_equal(perm[:m], perm_ref.astype(np.int64))
    # particles outside the grid follow, in index order
    outside = np.setdiff1d(np.arange(n), perm_ref)
    assert np.array_equal(perm[m:], outside)


def _clustered(rng, dims, dx, sizes, background):
    """Positions with one cluster per entry of sizes (all in one cell, at
    cells spread over the grid) plus uniform background particles."""
    parts = [rng.uniform(-0.15, np.array(dims) * dx + 0.15, size=(background, 3))]
    for c, m in enumerate(sizes):
        cell = np.array([(3 * c + 1) % dims[0], (5 * c + 2) % dims[1], (7 * c + 3) % dims[2]])
        parts.append((cell + rng.uniform(0.01, 0.99, size=(m, 3))) * dx)
    pos = np.concatenate(parts).astype(np.float32)
    return pos[rng.permutation(len(pos))]


def test_cell_sort_bin_size_classes(ctx, orc):
    """Bins ordered by a thread (up to 16), a wave (17..256) and a wave over
    LDS tiles (beyond, including more than one 2048-key tile)."""
    rng = np.random.default_rng(41)
    dims = (11, 9, 10)
    dx = 0.1
    pos = _clustered(rng, dims, dx, [17, 40, 64, 65, 130, 256, 257, 1500, 4500], 3000)
    perm_ref, start_ref = orc.cell_sort(pos, dx, dims)
    x, y, z = (_dev(pos[:, a].copy()) for a in range(3))
    perm, start = ctx.cell_sort_device(x, y, z, dx, dims)
    perm, start = _np(perm), _np(start)
    assert np.array_equal(start, start_ref.astype(np.int64))
    m = start_ref[dims[0] * dims[1] * dims[2]]
    assert np.array_equal(perm[:m], perm_ref.astype(np.int64))


def test_step_bin_size_classes(ctx, orc):
    """The fused step on cells of every ordering class and bricks that need
    several LDS chunks, against the oracle."""
    rng = np.random.default_rng(43)
    dims = (14, 12, 13)
    dx = 0.1
    pos = _clustered(rng, dims, dx, [20, 70, 300, 2600, 5000], 6000)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    exp = orc.flip_step(pos, vel, dims, dx, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
    for keep in (0, 1):
        res = _run_step(ctx, pos, vel, dims, dx, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR,
                        keep)
        _check_step(res, exp, bool(keep))


def test_cell_sort_nan_and_empty(ctx):
    torch = _torch()
    dims = (4, 4, 4)
    x = torch.tensor([0.05, float("nan"), 0.35, -1.0], device="cuda")
    y = torch.tensor([0.05, 0.1, 0.35, 0.1], device="cuda")
    z = torch.tensor([0.05, 0.1, 0.35, 0.1], device="cuda")
    perm, start = ctx.cell_sort_device(x, y, z, 0.1, dims)
    assert _np(perm).tolist() == [0, 2, 1, 3]
    assert int(start[-1]) == 2
    e = torch.empty(0, device="cuda")
    perm, start = ctx.cell_sort_device(e, e, e, 0.1, dims)
    assert perm.numel() == 0 and int(start.max()) == 0


# --------------------------------------------------------------------- P2G

@pytest.mark.parametrize("case", ["p2g_r1", "p2g_r17"])
def test_p2g_golden(ctx, orc64, case):
    g = load_golden(case)
    nd = tuple(int(v) for v in g["node_dims"])
    S, W = ctx.p2g(g["pos"], g["val"], float(g["radius"]), g["offset"], float(g["dx"]), nd)
    assert norm_rel(W, g["W"]) <= NORM_REL_TOL
    assert norm_rel(S, g["S"]) <= NORM_REL_TOL
    S64, W64 = orc64.p2g(g["pos"], g["val"], float(g["radius"]), g["offset"],
                         float(g["dx"]), nd)
    # FP64 oracle: the GPU is no further from it than the reference is (x10)
    assert norm_rel(S, S64) <= max(10 * norm_rel(g["S"], S64), 1e-6)
    assert norm_rel(W, W64) <= max(10 * norm_rel(g["W"], W64), 1e-6)
    # normalised field (ScalarField::applyWeightField)
    Sn = np.where(W > 0, S / np.where(W > 0, W, 1), S)
    assert norm_rel(Sn, g["S_norm"]) <= NORM_REL_TOL


def test_addpoints_golden(ctx):
    g = load_golden("addpoints")
    nd = tuple(int(v) for v in g["node_dims"])
    S, W = ctx.p2g(g["pos"], None, float(g["radius"]), g["offset"], float(g["dx"]), nd,
                   weight=False)
    assert W is None
    assert norm_rel(S, g["S"]) <= NORM_REL_TOL


def test_p2g_accumulates_and_is_deterministic(ctx, orc):
    g = load_golden("p2g_r1")
    nd = tuple(int(v) for v in g["node_dims"])
    args = (g["pos"], g["val"], float(g["radius"]), g["offset"], float(g["dx"]), nd)
    S1, W1 = ctx.p2g(*args)
    S2, W2 = ctx.p2g(*args)
    assert np.array_equal(S1, S2) and np.array_equal(W1, W2)
    base = np.random.default_rng(0).normal(size=S1.shape).astype(np.float32)
    S3, W3 = ctx.p2g(*args, S=base.copy(), W=base.copy())
    assert np.array_equal(S3, base + S1)
    assert np.array_equal(W3, base + W1)


def test_p2g_single_particle_closed_form(ctx, orc):
    # one particle: S = v K, W = K, so S/W = v on every node it reaches
    dx = 0.1
    p = np.array([[0.43, 0.27, 0.51]], np.float32)
    S, W = ctx.p2g(p, np.array([1.75], np.float32), dx, (0, 0, 0), dx, (10, 10, 10))
    hit = W > 0
    assert hit.sum() >= 4
    np.testing.assert_allclose(S[hit] / W[hit], 1.75, rtol=1e-6)
    k, j, i = np.nonzero(hit)
    pf = p[0].astype(np.float64)
    node = lambda a: (a * dx).astype(np.float32).astype(np.float64)
    d2 = (node(i) - pf[0]) ** 2 + (node(j) - pf[1]) ** 2 + (node(k) - pf[2]) ** 2
    Kexp = np.array([orc.kernel_weight(dx, float(t)) for t in d2])
    np.testing.assert_allclose(W[hit], Kexp, rtol=2e-6, atol=1e-7)


def test_p2g_device_matches_host(ctx):
    torch = _torch()
    g = load_golden("p2g_r17")
    nd = tuple(int(v) for v in g["node_dims"])
    S, W = ctx.p2g(g["pos"], g["val"], float(g["radius"]), g["offset"], float(g["dx"]), nd)
    Sd = torch.zeros(nd[::-1], device="cuda")
    Wd = torch.zeros(nd[::-1], device="cuda")
    ctx.p2g_device(_dev(g["pos"]), _dev(g["val"]), float(g["radius"]), g["offset"],
                   float(g["dx"]), nd, Sd, Wd)
    torch.cuda.synchronize()
    assert np.array_equal(_np(Sd), S) and np.array_equal(_np(Wd), W)
    ctx.apply_weight_device(Sd, Wd)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(Sd), np.where(W > 0, S / np.where(W > 0, W, 1), S))


def test_p2g_empty(ctx):
    S, W = ctx.p2g(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), 0.1,
                   (0, 0, 0), 0.1, (4, 5, 6))
    assert not S.any() and not W.any()


# --------------------------------------------------------------------- G2P

@pytest.mark.parametrize("stencil,case", [(TRILINEAR, "g2p_trilinear"),
                                          (TRICUBIC, "g2p_tricubic")])
def test_g2p_and_rk_golden(ctx, orc64, stencil, case):
    g = load_golden(case)
    dims = tuple(int(v) for v in g["dims"])
    dx = float(g["dx"])
    out = ctx.g2p(g["pos"], g["u"], g["v"], g["w"], dims, dx, stencil)
    assert norm_rel(out, g["out"]) <= NORM_REL_TOL
    # particles outside the grid read exactly zero, like the reference
    outside = np.all(g["out"] == 0, axis=1)
    assert np.all(out[outside] == 0)
    o64 = orc64.g2p(g["pos"], g["u"], g["v"], g["w"], dims, dx, stencil)
    assert norm_rel(out, o64) <= max(10 * norm_rel(g["out"], o64), 1e-6)
    for order in (1, 2, 3, 4):
        adv = ctx.advect(g["pos"], g["u"], g["v"], g["w"], dims, dx, float(g["dt"]),
                         order, stencil)
        exp = g["rk%d" % order]
        assert norm_rel(adv - g["pos"], exp - g["pos"]) <= NORM_REL_TOL, order
        assert norm_rel(adv, exp) <= 1e-6, order


def test_g2p_nan_field_is_zeroed(ctx):
    dims = (4, 4, 4)
    shapes = [(4, 4, 5), (4, 5, 4), (5, 4, 4)]
    u, v, w = [np.ones(s, np.float32) for s in shapes]
    u[2, 2, 2] = np.nan
    # particle 0 reads the NaN face; particle 1 sits 0.2 cells from the
    # low corner, where one of two samples per axis is out of range (0):
    # each component is 0.7 * 0.7 = 0.49 (macvelocityfield.cpp:470-478)
    pos = np.array([[0.2, 0.25, 0.25], [0.02, 0.02, 0.02]], np.float32)
    out = ctx.g2p(pos, u, v, w, dims, 0.1, TRILINEAR)
    assert np.array_equal(out[0], [0, 0, 0])
    np.testing.assert_allclose(out[1], [0.49, 0.49, 0.49], rtol=1e-5)


def test_g2p_device_in_place(ctx):
    torch = _torch()
    g = load_golden("g2p_trilinear")
    dims = tuple(int(v) for v in g["dims"])
    exp = ctx.g2p(g["pos"], g["u"], g["v"], g["w"], dims, float(g["dx"]), TRILINEAR)
    p = _dev(g["pos"])
    ctx.g2p_device(p, _dev(g["u"]), _dev(g["v"]), _dev(g["w"]), dims, float(g["dx"]),
                   TRILINEAR, out=p)
    torch.cuda.synchronize()
    assert np.array_equal(_np(p), exp)


# ---------------------------------------------------------------- the step

def _run_step(ctx, pos, vel, dims, dx, dt, bf, ratio, stencil, order, ids=True):
    import gridfluidsim3d_amd as gf
    st = gf.FlipState(pos, vel, dims, dx, track_ids=ids, keep_old=True, keep_weights=True)
    ctx.flip_step(st, dt, bf, ratio, stencil, order)
    _torch().cuda.synchronize()
    res = dict(pos=_np(st.positions()), vel=_np(st.velocities()),
               u=_np(st.grid[0]), v=_np(st.grid[1]), w=_np(st.grid[2]),
               u_old=_np(st.grid_old[0]), v_old=_np(st.grid_old[1]),
               w_old=_np(st.grid_old[2]), wu=_np(st.weights[0]),
               wv=_np(st.weights[1]), ww=_np(st.weights[2]))
    if ids:
        res["id"] = _np(st.particle_ids())
    return res


def _check_step(res, exp, order_keep):
    if not order_keep:
        idx = res["id"].astype(np.int64)
        assert np.array_equal(np.sort(idx), np.arange(idx.size))
        pos = np.empty_like(res["pos"])
        vel = np.empty_like(res["vel"])
        pos[idx] = res["pos"]
        vel[idx] = res["vel"]
    else:
        pos, vel = res["pos"], res["vel"]
    for key in ("u", "v", "w", "u_old", "v_old", "w_old", "wu", "wv", "ww"):
        assert norm_rel(res[key], exp[key]) <= NORM_REL_TOL, key
    assert norm_rel(vel, exp["vel"]) <= NORM_REL_TOL
    assert norm_rel(pos, exp["pos"]) <= NORM_REL_TOL
    return pos, vel


@pytest.mark.parametrize("stencil,case", [(TRILINEAR, "step_trilinear"),
                                          (TRICUBIC, "step_tricubic")])
@pytest.mark.parametrize("keep", [False, True])
def test_step_golden(ctx, stencil, case, keep):
    g = load_golden(case)
    dims = tuple(int(v) for v in g["dims"])
    res = _run_step(ctx, g["pos"], g["vel"], dims, float(g["dx"]), float(g["dt"]),
                    tuple(g["body_force"]), float(g["pic_ratio"]), stencil, int(keep))
    exp = {k[4:]: v for k, v in g.items() if k.startswith("out_")}
    pos, _ = _check_step(res, exp, keep)
    # displacement parity, relative to the displacement itself
    assert norm_rel(pos - g["pos"], exp["pos"] - g["pos"]) <= 1e-4


@pytest.mark.parametrize("seed,dims", [(11, (24, 16, 20)), (12, (9, 31, 7))])
@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
def test_step_random_scenes(ctx, orc, orc64, seed, dims, stencil):
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(seed, dims, 0.1, ppc=8 - seed % 3, amplitude=3.0)
    # a few particles just outside the domain and one far away
    extra = np.array([[-0.05, 0.5, 0.5], [0.5, dims[1] * 0.1 + 0.03, 0.5],
                      [50.0, 50.0, 50.0]], np.float32)
    pos = np.concatenate([sc.pos, extra])
    vel = np.concatenate([sc.vel, np.ones((3, 3), np.float32)])
    exp = orc.flip_step(pos, vel, dims, 0.1, sc.dt, (0.0, -9.8, 0.0), 0.05, stencil)
    res = _run_step(ctx, pos, vel, dims, 0.1, sc.dt, (0.0, -9.8, 0.0), 0.05, stencil, 0)
    pos_g, vel_g = _check_step(res, exp, False)
    # the far particle: zero velocity sample -> stays put, keeps 0.95 v
    assert np.array_equal(pos_g[-1], pos[-1])
    np.testing.assert_allclose(vel_g[-1], 0.95, rtol=1e-6)
    # error against FP64 is of the same size as the reference path's
    e64 = orc64.flip_step(pos, vel, dims, 0.1, sc.dt, (0.0, -9.8, 0.0), 0.05, stencil)
    for key, got in (("u", res["u"]), ("vel", vel_g)):
        assert norm_rel(got, e64[key]) <= max(10 * norm_rel(exp[key], e64[key]), 2e-6), key


def test_step_constant_velocity_closed_form(ctx):
    # uniform velocity c, no body force: u = c wherever W > 0, and every
    # particle keeps c (PIC and FLIP both return c)
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(5, (20, 20, 20), 0.1, amplitude=1.0)
    c = np.array([0.3, -0.7, 1.1], np.float32)
    vel = np.tile(c, (sc.n, 1))
    res = _run_step(ctx, sc.pos, vel, (20, 20, 20), 0.1, 1 / 240, (0.0, 0.0, 0.0), 0.05,
                    TRILINEAR, 1, ids=False)
    for comp, key, wk in ((0, "u", "wu"), (1, "v", "wv"), (2, "w", "ww")):
        hit = res[wk] > 0
        np.testing.assert_allclose(res[key][hit], c[comp], rtol=2e-6)
        assert np.all(res[key][~hit] == 0)
    # particles whose 8 trilinear stencil faces all carry weight (W > 0) for
    # every component sample u = c exactly, so PIC and FLIP both return c
    inner = np.ones(sc.n, bool)
    g = sc.pos.astype(np.float64) / 0.1
    for comp, wk in ((0, "wu"), (1, "wv"), (2, "ww")):
        base = [np.floor(g[:, a] - (0.0 if a == comp else 0.5)).astype(int) for a in range(3)]
        Wc = res[wk]
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    i, j, k = base[0] + di, base[1] + dj, base[2] + dk
                    ok = ((i >= 0) & (i < Wc.shape[2]) & (j >= 0) & (j < Wc.shape[1]) &
                          (k >= 0) & (k < Wc.shape[0]))
                    val = np.zeros(sc.n)
                    val[ok] = Wc[k[ok], j[ok], i[ok]]
                    inner &= val > 0
    assert inner.sum() > sc.n // 2
    np.testing.assert_allclose(res["vel"][inner], np.tile(c, (inner.sum(), 1)), rtol=1e-5,
                               atol=1e-6)


def test_step_empty(ctx):
    res = _run_step(ctx, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32),
                    (5, 6, 7), 0.1, 0.01, (0.0, -9.8, 0.0), 0.05, TRILINEAR, 0)
    assert res["pos"].shape == (0, 3)
    assert not res["u"].any() and not res["v"].any() and not res["w"].any()


def test_bad_arguments_raise(ctx):
    with pytest.raises(RuntimeError, match="stencil"):
        ctx.g2p(np.zeros((1, 3), np.float32), *[np.zeros(s, np.float32) for s in
                                                [(2, 2, 3), (2, 3, 2), (3, 2, 2)]],
                (2, 2, 2), 0.1, 7)
    with pytest.raises(RuntimeError, match="order"):
        ctx.advect(np.zeros((1, 3), np.float32), *[np.zeros(s, np.float32) for s in
                                                   [(2, 2, 3), (2, 3, 2), (3, 2, 2)]],
                   (2, 2, 2), 0.1, 0.01, order=5)
    with pytest.raises(RuntimeError, match="radius"):
        ctx.p2g(np.zeros((1, 3), np.float32), None, -1.0, (0, 0, 0), 0.1, (2, 2, 2))


@pytest.mark.parametrize("variant", [0, 1, 4])
def test_step_p2g_variants(ctx, orc, variant):
    """Brick-staged (0) and direct (1) P2G gathers and the tile scatter (4)
    all match the oracle."""
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(21, (27, 18, 13), 0.1, amplitude=2.0)
    exp = orc.flip_step(sc.pos, sc.vel, sc.dims, 0.1, sc.dt, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
    old = ctx.get_option("p2g_variant")
    ctx.set_option("p2g_variant", variant)
    try:
        res = _run_step(ctx, sc.pos, sc.vel, sc.dims, 0.1, sc.dt, (0.0, -9.8, 0.0), 0.05,
                        TRILINEAR, 0)
    finally:
        ctx.set_option("p2g_variant", old)
    _check_step(res, exp, False)


def _scatter_edge_scene(kind):
    rng = np.random.default_rng(41)
    if kind == "clump":
        # a clump of 600 particles in one cell inside ordinary fluid, some
        # particles nearly at rest beside fast ones
        dims, dx = (12, 10, 9), 0.1
        pos = (rng.uniform(0.2, 0.9, size=(4000, 3)) * np.array(dims) * dx).astype(np.float32)
        clump = np.array([0.55, 0.45, 0.35]) + rng.uniform(0.0, 0.1, size=(600, 3))
        pos = np.concatenate([pos, clump]).astype(np.float32)
        vel = rng.normal(size=pos.shape).astype(np.float32)
        vel[:100] *= 1e-3
        return pos, vel, dims, dx
    # particles in the pad layers and beyond them, non-finite velocities
    dims, dx = (9, 11, 7), 0.1
    W = np.array(dims) * dx
    pos = rng.uniform(-0.15, W + 0.15, size=(6000, 3)).astype(np.float32)
    vel = (3.0 * rng.normal(size=pos.shape)).astype(np.float32)
    vel[10, 0] = np.inf
    vel[20, 1] = -np.inf
    vel[30, 2] = np.nan
    vel[40] = (np.inf, -np.inf, np.nan)
    return pos, vel, dims, dx


@pytest.mark.parametrize("variant", [0, 4])
def test_step_clumped_cells_against_oracle(ctx, orc, variant):
    """Cells of 65..256 particles (sorted in registers by one wave, four
    keys per lane) among ordinary ones: both P2G forms match the oracle.
    The gather needs each particle's rank inside its half-cell from that
    sort, including runs that cross from lane 0 into lane 1."""
    pos, vel, dims, dx = _scatter_edge_scene("clump")
    exp = orc.flip_step(pos, vel, dims, dx, 0.01, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
    old = ctx.get_option("p2g_variant")
    ctx.set_option("p2g_variant", variant)
    try:
        res = _run_step(ctx, pos, vel, dims, dx, 0.01, (0.0, -9.8, 0.0), 0.05, TRILINEAR, 0)
    finally:
        ctx.set_option("p2g_variant", old)
    _check_step(res, exp, False)


@pytest.mark.parametrize("kind", ["clump", "wild"])
def test_step_p2g_scatter_matches_gather(ctx, kind):
    """The tile scatter (4) against the brick gather (0) on a clumped cell,
    the pad layers and non-finite velocities: the same faces are NaN and
    +-inf (what the float sums give: K v with v infinite is +-inf, or NaN
    where K is 0), and the finite values agree within the tolerance."""
    pos, vel, dims, dx = _scatter_edge_scene(kind)
    old = ctx.get_option("p2g_variant")
    out = []
    try:
        for variant in (0, 4):
            ctx.set_option("p2g_variant", variant)
            out.append(_run_step(ctx, pos, vel, dims, dx, 0.01, (0.0, -9.8, 0.0), 0.05,
                                 TRILINEAR, 1))
    finally:
        ctx.set_option("p2g_variant", old)
    a, b = out
    nonfinite = 0
    for key in ("u", "v", "w", "u_old", "v_old", "w_old", "wu", "wv", "ww", "vel", "pos"):
        fa, fb = np.isfinite(a[key]), np.isfinite(b[key])
        assert np.array_equal(fa, fb), key
        assert np.array_equal(np.isnan(a[key]), np.isnan(b[key])), key
        assert np.array_equal(a[key][np.isinf(a[key])], b[key][np.isinf(b[key])]), key
        assert norm_rel(b[key][fb], a[key][fa]) <= NORM_REL_TOL, key
        nonfinite += int((~fa).sum())
    if kind == "wild":
        assert nonfinite > 0


def _g2p_variant_scene(kind):
    from gridfluidsim3d_amd import scenes
    rng = np.random.default_rng(31)
    if kind == "blobs":
        sc = scenes.random_blobs(22, (27, 18, 13), 0.1, amplitude=2.0)
        return sc.pos, sc.vel, sc.dims, 0.1, sc.dt
    if kind == "interior":
        # 8^3-cell bricks two cells or more inside the grid on every side,
        # dx a power of two (FP32 cell coordinates)
        sc = scenes.random_blobs(23, (40, 36, 44), 1.0 / 32.0, amplitude=2.0)
        return sc.pos, sc.vel, sc.dims, 1.0 / 32.0, sc.dt
    if kind == "tiny":
        # one brick per axis, holding both pad layers
        dims = (5, 6, 7)
        pos = rng.uniform(-0.12, 0.8, size=(3000, 3)).astype(np.float32)
        return pos, rng.normal(size=pos.shape).astype(np.float32), dims, 0.1, 0.01
    # fast particles whose RK stages leave the staged boxes, particles in the
    # pad layers and beyond them, and a NaN position
    dims = (19, 17, 21)
    W = np.array(dims) * 0.1
    pos = rng.uniform(-0.4, W + 0.4, size=(20000, 3)).astype(np.float32)
    pos[7] = np.nan
    vel = (15.0 * rng.normal(size=pos.shape)).astype(np.float32)
    return pos, vel, dims, 0.1, 0.01


@pytest.mark.parametrize("kind", ["blobs", "tiny", "wild", "interior"])
@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
@pytest.mark.parametrize("keep", [False, True])
def test_step_g2p_variants_bit_identical(ctx, kind, stencil, keep):
    """The LDS-staged G2P (0) over bricks of 8^3 and of 4^3 cells and the
    direct one (1) give identical bits, including samples that fall back to
    global memory, particles outside the bin grid and bricks far enough
    inside the grid that their samples skip the grid tests."""
    pos, vel, dims, dx, dt = _g2p_variant_scene(kind)
    old = ctx.get_option("g2p_variant"), ctx.get_option("g2p_brick")
    out = []
    try:
        for variant, brick in ((0, 8), (0, 4), (1, 8)):
            ctx.set_option("g2p_variant", variant)
            ctx.set_option("g2p_brick", brick)
            out.append(_run_step(ctx, pos, vel, dims, dx, dt, (0.0, -9.8, 0.0), 0.05,
                                 stencil, int(keep)))
    finally:
        ctx.set_option("g2p_variant", old[0])
        ctx.set_option("g2p_brick", old[1])
    for other in out[1:]:
        for key in out[0]:
            assert np.array_equal(out[0][key], other[key], equal_nan=True), key


def _gather_runs(ctx, fn):
    old = ctx.get_option("particle_gather")
    out = []
    try:
        for gather in (1, 2, 0):
            ctx.set_option("particle_gather", gather)
            out.append(fn())
    finally:
        ctx.set_option("particle_gather", old)
    return out


@pytest.mark.parametrize("kind", ["blobs", "tiny", "wild"])
@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
@pytest.mark.parametrize("g2p_brick", [8, 4])
def test_step_particle_gather_bit_identical(ctx, kind, stencil, g2p_brick):
    """P2G and G2P reading each sort slot's particle through the sort's
    permutation (particle_gather 1), P2G doing so and writing the sorted
    copies for G2P (2), and the sort writing them (0) give the same bits,
    three sorted steps in a row with particle ids: the first from an
    unsorted input, the next from the previous step's sorted output."""
    import gridfluidsim3d_amd as gf
    pos, vel, dims, dx, dt = _g2p_variant_scene(kind)

    def run():
        st = gf.FlipState(pos, vel, dims, dx, track_ids=True, keep_old=True,
                          keep_weights=True)
        res = []
        for _ in range(3):
            ctx.flip_step(st, dt, (0.0, -9.8, 0.0), 0.05, stencil, gf.ORDER_SORTED)
            _torch().cuda.synchronize()
            res.append(dict(pos=_np(st.positions()), vel=_np(st.velocities()),
                            id=_np(st.particle_ids()), u=_np(st.grid[0]),
                            v=_np(st.grid[1]), w=_np(st.grid[2]), wu=_np(st.weights[0])))
        return res

    old = ctx.get_option("g2p_brick")
    try:
        ctx.set_option("g2p_brick", g2p_brick)
        runs = _gather_runs(ctx, run)
    finally:
        ctx.set_option("g2p_brick", old)
    for other in runs[1:]:
        for step, (ra, rb) in enumerate(zip(runs[0], other)):
            for key in ra:
                assert np.array_equal(ra[key], rb[key], equal_nan=True), (step, key)


def _pow2_scene(kind):
    """Scenes with 1/dx a power of two, where the brick G2P takes cell
    coordinates in FP32: pad-layer particles (x < 0), positions on cell
    faces, tiny and subnormal coordinates of either sign, NaN and far-out
    particles."""
    from gridfluidsim3d_amd import scenes
    rng = np.random.default_rng(37)
    if kind == "blobs":
        sc = scenes.random_blobs(23, (27, 18, 13), 0.0625, amplitude=2.0)
        return sc.pos, sc.vel, sc.dims, 0.0625, sc.dt
    dx = 0.125 if kind == "edges" else 2.0
    dims = (19, 17, 21)
    W = np.array(dims) * dx
    pos = rng.uniform(-1.5 * dx, W + 1.5 * dx, size=(20000, 3)).astype(np.float32)
    k = np.arange(4000)
    pos[:4000, 0] = (rng.integers(-1, dims[0] + 1, 4000) * dx).astype(np.float32)
    pos[4000:5000, 1] = rng.choice(np.array([0.0, -0.0, 1e-30, -1e-30, 1e-42, -1e-42,
                                             np.float32(1e-7) * dx, -np.float32(1e-7) * dx],
                                            np.float32), 1000)
    pos[5000:5100] = rng.choice(np.array([1e30, -1e30, 3e38], np.float32), (100, 3))
    pos[7] = np.nan
    pos[k[::97], 2] = np.nextafter(np.float32(W[2]), np.float32(0))
    vel = (4.0 / 0.1 * dx * rng.normal(size=pos.shape)).astype(np.float32)
    return pos, vel, dims, dx, 0.01


@pytest.mark.parametrize("kind", ["blobs", "edges", "coarse"])
@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
@pytest.mark.parametrize("order", [0, 1])
def test_step_fp32_cells_bit_identical(ctx, kind, stencil, order):
    """With 1/dx a power of two the brick G2P takes cell coordinates in FP32
    (fp32_cells 1): the same bits as the FP64 coordinates (0), three steps
    in a row, G2P bricks of 8^3 and 4^3 cells."""
    import gridfluidsim3d_amd as gf
    pos, vel, dims, dx, dt = _pow2_scene(kind)

    def run():
        st = gf.FlipState(pos, vel, dims, dx, track_ids=True, keep_old=True,
                          keep_weights=True)
        res = []
        for _ in range(3):
            ctx.flip_step(st, dt, (0.0, -9.8, 0.0), 0.05, stencil, order)
            _torch().cuda.synchronize()
            res.append(dict(pos=_np(st.positions()), vel=_np(st.velocities()),
                            id=_np(st.particle_ids()), u=_np(st.grid[0]),
                            v=_np(st.grid[1]), w=_np(st.grid[2]), wu=_np(st.weights[0]),
                            u_old=_np(st.grid_old[0])))
        return res

    old = ctx.get_option("fp32_cells"), ctx.get_option("g2p_brick")
    runs = []
    try:
        for cells, brick in ((1, 8), (0, 8), (1, 4), (0, 4)):
            ctx.set_option("fp32_cells", cells)
            ctx.set_option("g2p_brick", brick)
            runs.append(run())
    finally:
        ctx.set_option("fp32_cells", old[0])
        ctx.set_option("g2p_brick", old[1])
    for other in runs[1:]:
        for step, (ra, rb) in enumerate(zip(runs[0], other)):
            for key in ra:
                assert np.array_equal(ra[key], rb[key], equal_nan=True), (step, key)


def _brick_keys(pos, dims, dx):
    """The brick sort's key of every position (numpy restatement): P2G brick
    of 4^3 cells, the 8 bricks of an 8^3-cell G2P brick consecutive, pad
    cells in the edge bricks, the rest in the overflow bin."""
    d = np.array(dims)
    cell = np.floor(pos.astype(np.float64) * (1.0 / dx))
    inside = ((cell >= -1) & (cell <= d)).all(1)
    c = np.where(np.isfinite(cell), cell, 0).astype(np.int64)
    b = np.minimum(np.maximum(c, 0), d - 1) >> 2
    gb = ((d + 3) // 4 + 1) // 2
    g = (b[:, 0] >> 1) + gb[0] * ((b[:, 1] >> 1) + gb[1] * (b[:, 2] >> 1))
    key = (g << 3) | (b[:, 0] & 1) | ((b[:, 1] & 1) << 1) | ((b[:, 2] & 1) << 2)
    return np.where(inside, key, 8 * int(np.prod(gb)))


@pytest.mark.parametrize("gather", [2, 0])
def test_step_sort_of_unsorted_input(ctx, orc, gather):
    """The brick sort on input in no useful order: a cell-ordered scene
    followed by uniformly random particles over many bricks (work-groups of
    the sort whose particles fall in more bins than their LDS table holds
    take every run as a chunk of its own), pad-layer and far-out particles.
    The output order is a stable sort by brick, and the step matches the
    oracle."""
    from gridfluidsim3d_amd import scenes
    dims, dx = (64, 48, 40), 1.0 / 64.0
    sc = scenes.random_blobs(41, dims, dx, ppc=4, amplitude=1.0, fill=(0.1, 0.2))
    rng = np.random.default_rng(43)
    W = np.array(dims) * dx
    extra = rng.uniform(-1.5 * dx, W + 1.5 * dx, size=(120000, 3)).astype(np.float32)
    extra[:200] *= np.float32(3.0)
    pos = np.concatenate([sc.pos, extra])
    vel = np.concatenate([sc.vel, rng.normal(size=extra.shape).astype(np.float32)])
    old = ctx.get_option("particle_gather")
    try:
        ctx.set_option("particle_gather", gather)
        res = _run_step(ctx, pos, vel, dims, dx, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR, 0)
    finally:
        ctx.set_option("particle_gather", old)
    ids = res["id"].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(len(pos)))
    key = _brick_keys(pos[ids], dims, dx)
    assert (key[1:] >= key[:-1]).all()
    same = key[1:] == key[:-1]
    assert (ids[1:][same] > ids[:-1][same]).all()
    exp = orc.flip_step(pos, vel, dims, dx, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
    _check_step(res, exp, False)


def test_step_dense_cells_use_fallback(ctx, orc):
    """Bricks holding more particles than fit in LDS take the global-memory
    gather; results still match the oracle."""
    rng = np.random.default_rng(5)
    dims = (12, 12, 12)
    dx = 0.1
    dense = rng.uniform(0.3, 0.7, size=(6000, 3)).astype(np.float32)   # ~94 per cell
    sparse = rng.uniform(0.0, 1.2, size=(2000, 3)).astype(np.float32)
    pos = np.concatenate([dense, sparse])
    vel = rng.normal(size=pos.shape).astype(np.float32)
    exp = orc.flip_step(pos, vel, dims, dx, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
    res = _run_step(ctx, pos, vel, dims, dx, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR, 0)
    _check_step(res, exp, False)


def test_step_is_deterministic(ctx):
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(8, (30, 22, 17), 0.05, amplitude=3.0)
    r1 = _run_step(ctx, sc.pos, sc.vel, sc.dims, 0.05, sc.dt, (0.0, -9.8, 0.0), 0.05,
                   TRILINEAR, 0)
    r2 = _run_step(ctx, sc.pos, sc.vel, sc.dims, 0.05, sc.dt, (0.0, -9.8, 0.0), 0.05,
                   TRILINEAR, 0)
    for key in r1:
        assert np.array_equal(r1[key], r2[key]), key


def test_options(ctx):
    assert ctx.get_option("p2g_variant") == 4
    assert ctx.get_option("g2p_variant") == 0
    assert ctx.get_option("g2p_brick") == 8
    assert ctx.get_option("particle_gather") == 2
    assert ctx.get_option("fp32_cells") == 1
    with pytest.raises(RuntimeError, match="fp32_cells"):
        ctx.set_option("fp32_cells", 2)
    with pytest.raises(RuntimeError, match="particle_gather"):
        ctx.set_option("particle_gather", 3)
    with pytest.raises(RuntimeError, match="g2p_variant"):
        ctx.set_option("g2p_variant", 2)
    with pytest.raises(RuntimeError, match="g2p_brick"):
        ctx.set_option("g2p_brick", 6)
    with pytest.raises(RuntimeError, match="unknown option"):
        ctx.set_option("no_such_option", 1)
    with pytest.raises(RuntimeError, match="p2g_variant"):
        ctx.set_option("p2g_variant", 7)


# ------------------------------------------------------------ batched step

def _batch_scenes(seed):
    """Uneven scenes: different grids and counts, one without particles, one
    with particles beyond its padded grid."""
    from gridfluidsim3d_amd import scenes
    rng = np.random.default_rng(seed)
    out = []
    for k, dims in enumerate([(9, 13, 7), (20, 11, 16), (5, 6, 7), (17, 17, 9)]):
        sc = scenes.random_blobs(seed + k, dims, 0.1, amplitude=2.0)
        pos, vel = sc.pos, sc.vel
        if k == 3:
            W = np.array(dims) * 0.1
            extra = rng.uniform(-0.5, W + 0.5, size=(500, 3)).astype(np.float32)
            pos = np.concatenate([pos, extra])
            vel = np.concatenate([vel, rng.normal(size=extra.shape).astype(np.float32)])
        out.append((pos, vel, dims, 0.1))
    out.insert(2, (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), (6, 5, 4), 0.1))
    return out


def _state_result(st):
    res = dict(pos=_np(st.positions()), vel=_np(st.velocities()),
               u=_np(st.grid[0]), v=_np(st.grid[1]), w=_np(st.grid[2]),
               u_old=_np(st.grid_old[0]), v_old=_np(st.grid_old[1]), w_old=_np(st.grid_old[2]),
               wu=_np(st.weights[0]), wv=_np(st.weights[1]), ww=_np(st.weights[2]))
    res["id"] = _np(st.particle_ids())
    return res


@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
@pytest.mark.parametrize("keep", [False, True])
def test_step_batched_matches_single(ctx, stencil, keep):
    """Every scene of a batched step gives the bits of a step on that scene
    alone; two steps in a row, with per-scene time steps."""
    import gridfluidsim3d_amd as gf
    sc = _batch_scenes(61)
    dts = [1 / 240, 1 / 120, 1 / 200, 1 / 300, 1 / 240]
    mk = lambda p, v, d, h: gf.FlipState(p, v, d, h, track_ids=True, keep_old=True,
                                         keep_weights=True)
    batch = [mk(*s) for s in sc]
    alone = [mk(*s) for s in sc]
    for _ in range(2):
        ctx.flip_step_batched(batch, dts, (0.0, -9.8, 0.0), 0.05, stencil, int(keep))
        for st, dt in zip(alone, dts):
            ctx.flip_step(st, dt, (0.0, -9.8, 0.0), 0.05, stencil, int(keep))
    _torch().cuda.synchronize()
    for i, (a, b) in enumerate(zip(batch, alone)):
        ra, rb = _state_result(a), _state_result(b)
        for key in ra:
            assert np.array_equal(ra[key], rb[key], equal_nan=True), (i, key)


def test_step_batched_against_oracle(ctx, orc):
    import gridfluidsim3d_amd as gf
    sc = _batch_scenes(71)
    states = [gf.FlipState(p, v, d, h, track_ids=True, keep_old=True, keep_weights=True)
              for p, v, d, h in sc]
    ctx.flip_step_batched(states, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR, 0)
    _torch().cuda.synchronize()
    for (p, v, d, h), st in zip(sc, states):
        if len(p) == 0:
            continue
        exp = orc.flip_step(p, v, d, h, 1 / 240, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
        _check_step(_state_result(st), exp, False)


@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
def test_step_batched_g2p_bricks_agree(ctx, stencil):
    """The batched step over G2P bricks of 4^3 cells gives the bits of the
    one over 8^3-cell bricks, two steps in a row."""
    import gridfluidsim3d_amd as gf
    sc = _batch_scenes(71)
    dts = [1 / 240, 1 / 120, 1 / 200, 1 / 300, 1 / 240]
    mk = lambda p, v, d, h: gf.FlipState(p, v, d, h, track_ids=True, keep_old=True,
                                         keep_weights=True)
    runs = []
    old = ctx.get_option("g2p_brick")
    try:
        for brick in (8, 4):
            ctx.set_option("g2p_brick", brick)
            states = [mk(*s) for s in sc]
            for _ in range(2):
                ctx.flip_step_batched(states, dts, (0.0, -9.8, 0.0), 0.05, stencil,
                                      gf.ORDER_SORTED)
            _torch().cuda.synchronize()
            runs.append([_state_result(st) for st in states])
    finally:
        ctx.set_option("g2p_brick", old)
    for i, (a, b) in enumerate(zip(*runs)):
        for key in a:
            assert np.array_equal(a[key], b[key], equal_nan=True), (i, key)


@pytest.mark.parametrize("stencil", [TRILINEAR, TRICUBIC])
def test_step_batched_particle_gather_agrees(ctx, stencil):
    """The batched step reading particles through the sort's permutation
    gives the bits of the one over sorted copies, three steps in a row."""
    import gridfluidsim3d_amd as gf
    sc = _batch_scenes(73)
    dts = [1 / 240, 1 / 120, 1 / 200, 1 / 300, 1 / 240]

    def run():
        states = [gf.FlipState(p, v, d, h, track_ids=True, keep_old=True, keep_weights=True)
                  for p, v, d, h in sc]
        for _ in range(3):
            ctx.flip_step_batched(states, dts, (0.0, -9.8, 0.0), 0.05, stencil,
                                  gf.ORDER_SORTED)
        _torch().cuda.synchronize()
        return [_state_result(st) for st in states]

    runs = _gather_runs(ctx, run)
    for other in runs[1:]:
        for i, (ra, rb) in enumerate(zip(runs[0], other)):
            for key in ra:
                assert np.array_equal(ra[key], rb[key], equal_nan=True), (i, key)


def test_step_batched_rejects_direct_variants(ctx):
    import gridfluidsim3d_amd as gf
    sc = _batch_scenes(81)[:2]
    states = [gf.FlipState(p, v, d, h) for p, v, d, h in sc]
    old = ctx.get_option("g2p_variant")
    ctx.set_option("g2p_variant", 1)
    try:
        with pytest.raises(RuntimeError, match="one scene"):
            ctx.flip_step_batched(states, 1 / 240)
    finally:
        ctx.set_option("g2p_variant", old)


def test_step_batched_graph_replays_match_eager(ctx):
    """The batched step captured into HIP graphs gives, replay after replay
    (with other work and synchronisations between replays), the bits of the
    eager batched step."""
    import gridfluidsim3d_amd as gf
    torch = _torch()
    sc = _batch_scenes(91)
    mk = lambda p, v, d, h: gf.FlipState(p, v, d, h, track_ids=True, keep_old=True,
                                         keep_weights=True)
    graph_states = [mk(*s) for s in sc]
    eager_states = [mk(*s) for s in sc]
    step = ctx.capture_batched(graph_states, 1 / 240)
    # capture_batched ran two eager steps first: bring the others level
    for _ in range(2):
        ctx.flip_step_batched(eager_states, 1 / 240)
    for rep in range(4):
        step()
        ctx.flip_step_batched(eager_states, 1 / 240)
        junk = torch.ones(1 << 20, device="cuda").cumsum(0)   # unrelated eager work
        torch.cuda.synchronize()
        del junk
        for i, (a, b) in enumerate(zip(graph_states, eager_states)):
            ra, rb = _state_result(a), _state_result(b)
            for key in ra:
                assert np.array_equal(ra[key], rb[key], equal_nan=True), (rep, i, key)


def test_autotune_picks_a_variant(ctx):
    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import autotune
    sc = _batch_scenes(95)
    states = [gf.FlipState(p, v, d, h) for p, v, d, h in sc]
    before = [st.positions().clone() for st in states]
    old = ctx.get_option("p2g_variant")
    try:
        res = autotune.tune(ctx, states, 1 / 240, steps=1)
        best, times = res["p2g_variant"]
        assert best in autotune.CANDIDATES["p2g_variant"] and best == ctx.get_option("p2g_variant")
        assert set(times) == set(autotune.CANDIDATES["p2g_variant"])
    finally:
        ctx.set_option("p2g_variant", old)
    # the caller's states did not move
    for st, b in zip(states, before):
        assert _torch().equal(st.positions(), b)


def test_step_on_loaded_save_state(ctx, orc, tmp_path):
    """A scene read from a save state (the reference's binary layout,
    gridfluidsim3d_amd/savestate.py) through the fused step, against the
    oracle."""
    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import savestate, scenes
    sc = scenes.random_blobs(61, (24, 20, 18), 0.05, ppc=8, amplitude=1.5)
    f = tmp_path / "scene.state"
    savestate.write_state(str(f), sc.dims, sc.dx, sc.pos, sc.vel, frame=3)
    r = savestate.SaveStateReader()
    assert r.load_state(str(f))
    loaded = r.to_scene(1.0 / 120.0)
    st = gf.FlipState(loaded.pos, loaded.vel, loaded.dims, loaded.dx, track_ids=True)
    ctx.flip_step(st, loaded.dt, loaded.body_force, 0.05, gf.TRILINEAR, gf.ORDER_SORTED)
    ref = orc.flip_step(loaded.pos, loaded.vel, loaded.dims, loaded.dx, loaded.dt,
                        loaded.body_force, 0.05, gf.TRILINEAR)
    ids = st.particle_ids().cpu().numpy()
    assert norm_rel(st.positions().cpu().numpy()[np.argsort(ids)] - loaded.pos,
                    ref["pos"] - loaded.pos) <= NORM_REL_TOL
    assert norm_rel(st.velocities().cpu().numpy()[np.argsort(ids)], ref["vel"]) <= NORM_REL_TOL
