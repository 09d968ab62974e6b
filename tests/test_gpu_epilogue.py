"""The fused step with FluidSimulation's epilogue (GFSX_EPILOGUE_FLUIDSIM:
face fill of fluidsimulation.cpp:2597-2730, extrapolation of
macvelocityfield.cpp:577-798 on the snapshot and on the field) against the
oracle's restatement (gfo_flip_step_fs, pinned bit for bit to the
reference's own MACVelocityField extrapolation in test_oracle_epilogue.py).
Norm-wise relative error bar 1e-5 on every field and particle array."""
import numpy as np
import pytest

from conftest import NORM_REL_TOL, norm_rel

pytestmark = pytest.mark.gpu

TRILINEAR, TRICUBIC = 0, 1


def border_solid(dims):
    ni, nj, nk = dims
    s = np.zeros((nk, nj, ni), np.uint8)
    s[0, :, :] = s[-1, :, :] = 1
    s[:, 0, :] = s[:, -1, :] = 1
    s[:, :, 0] = s[:, :, -1] = 1
    return s


def outside_solid(pos, dx, solid):
    nk, nj, ni = solid.shape
    c = np.floor(pos.astype(np.float64) / dx).astype(np.int64)
    ok = ((c >= 0) & (c < np.array([ni, nj, nk]))).all(1)
    keep = np.ones(len(pos), bool)
    keep[ok] = solid[c[ok, 2], c[ok, 1], c[ok, 0]] == 0
    return keep


def run_gpu(ctx, pos, vel, dims, dx, dt, stencil, order, solid, layers):
    import gridfluidsim3d_amd as gf
    st = gf.FlipState(pos, vel, dims, dx, track_ids=True, keep_old=True, solid=solid)
    ctx.flip_step(st, dt, (0.0, -9.8, 0.0), 0.05, stencil, order, gf.EPILOGUE_FLUIDSIM, layers)
    ids = st.particle_ids().cpu().numpy()
    inv = np.argsort(ids)
    return dict(pos=st.positions().cpu().numpy()[inv], vel=st.velocities().cpu().numpy()[inv],
                u=st.grid[0].cpu().numpy(), v=st.grid[1].cpu().numpy(),
                w=st.grid[2].cpu().numpy(), u_old=st.grid_old[0].cpu().numpy(),
                v_old=st.grid_old[1].cpu().numpy(), w_old=st.grid_old[2].cpu().numpy())


@pytest.mark.parametrize("solid", [False, True])
@pytest.mark.parametrize("stencil,order", [(TRILINEAR, 0), (TRILINEAR, 1), (TRICUBIC, 0)])
def test_step_fs_matches_oracle(ctx, orc, solid, stencil, order):
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(81, (40, 34, 28), 1.0 / 32.0, ppc=8, amplitude=2.0,
                             fill=(0.15, 0.35))
    sol = border_solid(sc.dims) if solid else None
    pos, vel = sc.pos, sc.vel
    if solid:
        keep = outside_solid(pos, sc.dx, sol)
        pos, vel = pos[keep], vel[keep]
    dt = 1.0 / 120.0
    got = run_gpu(ctx, pos, vel, sc.dims, sc.dx, dt, stencil, order, sol, 7)
    exp = orc.flip_step_fs(pos, vel, sc.dims, sc.dx, dt, (0.0, -9.8, 0.0), 0.05, stencil,
                           sol, 7)
    for key in ("u", "v", "w", "u_old", "v_old", "w_old"):
        assert norm_rel(got[key], exp[key]) <= NORM_REL_TOL, key
    assert norm_rel(got["vel"], exp["vel"]) <= NORM_REL_TOL
    assert norm_rel(got["pos"] - pos, exp["pos"] - pos) <= NORM_REL_TOL
    # the epilogue did something: faces off the fluid were filled or extrapolated
    plain = orc.flip_step(pos, vel, sc.dims, sc.dx, dt, (0.0, -9.8, 0.0), 0.05, stencil)
    assert np.count_nonzero(exp["u"]) > np.count_nonzero(plain["u"])


@pytest.mark.parametrize("layers", [1, 3])
def test_step_fs_layers_sphere64(ctx, orc, layers):
    """Config 1 (64^3 sphere, Gaussian velocities) with few layers."""
    from gridfluidsim3d_amd import scenes
    sc = scenes.sphere64()
    got = run_gpu(ctx, sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, TRILINEAR, 0, None, layers)
    exp = orc.flip_step_fs(sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, (0.0, -9.8, 0.0), 0.05,
                           TRILINEAR, None, layers)
    for key in ("u", "v", "w", "u_old", "v_old", "w_old", "vel"):
        assert norm_rel(got[key], exp[key]) <= NORM_REL_TOL, key


def test_step_fs_is_deterministic(ctx):
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(82, (32, 30, 26), 1.0 / 32.0, ppc=8)
    a = run_gpu(ctx, sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, TRILINEAR, 0, None, 7)
    b = run_gpu(ctx, sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, TRILINEAR, 0, None, 7)
    for key in a:
        assert np.array_equal(a[key], b[key]), key


def test_step_fs_rejections(ctx):
    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(83, (12, 12, 12), 0.1, ppc=4)
    st = gf.FlipState(sc.pos, sc.vel, sc.dims, sc.dx, solid=border_solid(sc.dims))
    with pytest.raises(RuntimeError, match="solid"):
        ctx.flip_step(st, sc.dt)                     # solids without the epilogue
    st2 = gf.FlipState(sc.pos, sc.vel, sc.dims, sc.dx)
    f = gf.gfsx.pb.lib_func("gfsx_flip_step_batched",
                            [gf.gfsx.c_void_p, gf.gfsx.POINTER(gf.gfsx.Scene), gf.gfsx.c_int,
                             gf.gfsx.POINTER(gf.gfsx.StepParams)])
    scs = (gf.gfsx.Scene * 2)(ctx._scene_of(st2), ctx._scene_of(st2))
    prs = (gf.gfsx.StepParams * 2)(*[ctx._params(0.01, (0, -9.8, 0), 0.05, 0, 0,
                                                  gf.EPILOGUE_FLUIDSIM) for _ in range(2)])
    with pytest.raises(RuntimeError, match="one scene"):
        gf.gfsx.pb.execute_lib_func(f, [ctx(), scs, 2, prs])
    with pytest.raises(RuntimeError, match="epilogue"):
        ctx.flip_step(st2, sc.dt, epilogue=5)


@pytest.mark.parametrize("order", [0, 1])
def test_step_fs_collisions_match_oracle(ctx, orc, order):
    """Particles driven into the solid border: the collision fix
    (_resolveParticleSolidCellCollision, fluidsimulation.cpp:3145-3207)
    against the oracle, which matches the reference's own Collision
    functions bit for bit."""
    from test_oracle_epilogue import wall_scene
    dims, dx, pos, vel = wall_scene()
    sol = border_solid(dims)
    dt = 1.0 / 20.0
    got = run_gpu(ctx, pos, vel, dims, dx, dt, TRILINEAR, order, sol, 7)
    exp = orc.flip_step_fs(pos, vel, dims, dx, dt, (0.0, -9.8, 0.0), 0.05, TRILINEAR, sol, 7)
    assert norm_rel(got["pos"] - pos, exp["pos"] - pos) <= NORM_REL_TOL
    assert norm_rel(got["vel"], exp["vel"]) <= NORM_REL_TOL
    cells = np.floor(got["pos"].astype(np.float64) / dx).astype(int)
    assert (cells[:, 0] >= 1).all()
    # every particle that the oracle pulled back was pulled back here too
    free = orc.flip_step_fs(pos, vel, dims, dx, dt, (0.0, -9.8, 0.0), 0.05, TRILINEAR, None, 7)
    assert (free["pos"][:, 0] < dx).sum() > len(pos) // 4


def test_step_fs_collisions_particle_gather_bit_identical(ctx):
    """The collision fix reads each particle's start position through the
    sort's permutation when the step keeps no sorted copies (particle_gather
    1): the same bits as with the copies (2: written by P2G, 0: by the
    sort)."""
    from test_oracle_epilogue import wall_scene
    dims, dx, pos, vel = wall_scene()
    sol = border_solid(dims)
    old = ctx.get_option("particle_gather")
    out = []
    try:
        for gather in (1, 2, 0):
            ctx.set_option("particle_gather", gather)
            out.append(run_gpu(ctx, pos, vel, dims, dx, 1.0 / 20.0, TRILINEAR, 0, sol, 7))
    finally:
        ctx.set_option("particle_gather", old)
    for other in out[1:]:
        for key in out[0]:
            assert np.array_equal(out[0][key], other[key]), key
