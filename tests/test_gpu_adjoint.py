"""The differentiable form on the GPU: gfsx_flip_step_backward,
gfsx_p2g_backward_device and the torch.autograd.Function layer, checked
against

* the dense FP64 PyTorch reference (tests/dense_flip.py, pinned against the
  C oracle in test_adjoint_cpu.py): full gradients at small size;
* FP64 central finite differences of the C oracle's rollout on the 32^3 /
  20,480-particle subset: directional derivatives of the 16-step density
  loss w.r.t. the initial velocities.

Tolerances, as norm-wise relative errors max|g - g_ref| / max|g_ref| (or
|d - d_ref| / |d_ref| for a directional derivative): the adjoint runs in
FP32 on FP32 states; a single step's gradient is held to 2e-5 (measured
~1.5e-6), a 4-step rollout's to 1e-4 (measured ~3e-6), and the 16-step
rollout's directional derivatives, which also carry the FP32 trajectory's
drift from the FP64 one and the finite-difference error, to 2e-4 (measured
~1e-5).
"""
import numpy as np
import pytest
import torch

import dense_flip as dfl
from test_adjoint_cpu import oracle_rollout_loss, small_scene

pytestmark = pytest.mark.gpu

F64 = torch.float64
STEP_TOL = 2e-5
ROLLOUT_TOL = 2e-4


def nrel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def dev_soa(a):
    """(n, 3) array -> (3, n) float32 CUDA tensor."""
    return torch.tensor(np.ascontiguousarray(np.asarray(a, np.float32).T), device="cuda")


@pytest.mark.parametrize("seed", [41, 42, 43])
def test_step_backward_matches_dense(ctx, seed):
    dims, dx, dt = (10, 9, 8), 0.1, 1.0 / 60.0
    pos, vel = small_scene(seed, dims, dx)
    rng = np.random.default_rng(seed + 7)
    a = rng.normal(size=pos.shape)
    b = rng.normal(size=pos.shape)
    gx, gv = ctx.flip_step_backward(dev_soa(pos), dev_soa(vel), dev_soa(a), dev_soa(b),
                                    dims, dx, dt)
    torch.cuda.synchronize()
    x = torch.tensor(pos, dtype=F64, requires_grad=True)
    v = torch.tensor(vel, dtype=F64, requires_grad=True)
    x1, v1 = dfl.step(x, v, dims, dx, dt)
    L = (torch.tensor(a) * x1).sum() + (torch.tensor(b) * v1).sum()
    rx, rv = torch.autograd.grad(L, (x, v))
    ex = nrel(gx.cpu().numpy().T, rx.numpy())
    ev = nrel(gv.cpu().numpy().T, rv.numpy())
    print("seed %d: dL/dx err %.2e, dL/dv err %.2e" % (seed, ex, ev))
    assert ex < STEP_TOL and ev < STEP_TOL


def test_step_backward_fast_and_outside_particles(ctx):
    """Stages that move several cells (the brick kernel's global fallback)
    and particles outside the grid, beyond the bin grid's pad layer (its
    overflow bin), against the dense reference."""
    dims, dx, dt = (10, 9, 8), 0.1, 1.0 / 60.0
    pos, vel = small_scene(47, dims, dx, amp=25.0)
    rng = np.random.default_rng(47)
    far = (rng.uniform(-0.4, 1.4, size=(40, 3)) * np.array(dims) * dx).astype(np.float32)
    pos = np.concatenate([pos, far])
    vel = np.concatenate([vel, rng.normal(size=(40, 3)).astype(np.float32) * 5])
    a = rng.normal(size=pos.shape)
    b = rng.normal(size=pos.shape)
    for variant in (0, 1):
        ctx.set_option("adjoint_variant", variant)
        try:
            gx, gv = ctx.flip_step_backward(dev_soa(pos), dev_soa(vel), dev_soa(a), dev_soa(b),
                                            dims, dx, dt)
        finally:
            ctx.set_option("adjoint_variant", 0)
        x = torch.tensor(pos, dtype=F64, requires_grad=True)
        v = torch.tensor(vel, dtype=F64, requires_grad=True)
        x1, v1 = dfl.step(x, v, dims, dx, dt)
        L = (torch.tensor(a) * x1).sum() + (torch.tensor(b) * v1).sum()
        rx, rv = torch.autograd.grad(L, (x, v))
        ex = nrel(gx.cpu().numpy().T, rx.numpy())
        ev = nrel(gv.cpu().numpy().T, rv.numpy())
        print("variant %d: dL/dx err %.2e, dL/dv err %.2e" % (variant, ex, ev))
        assert ex < STEP_TOL and ev < STEP_TOL


def test_adjoint_variants_agree(ctx):
    """The LDS-brick adjoint and the global-atomic adjoint on a 40^3 scene
    of random blobs: equal up to the order of FP32 additions."""
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(48, (40, 36, 44), 1.0 / 40.0, ppc=6, amplitude=2.0)
    rng = np.random.default_rng(48)
    a = dev_soa(rng.normal(size=sc.pos.shape))
    b = dev_soa(rng.normal(size=sc.pos.shape))
    x, v = dev_soa(sc.pos), dev_soa(sc.vel)
    out = []
    for variant in (0, 1):
        ctx.set_option("adjoint_variant", variant)
        try:
            out.append([t.cpu().numpy() for t in ctx.flip_step_backward(x, v, a, b, sc.dims,
                                                                        sc.dx, sc.dt)])
        finally:
            ctx.set_option("adjoint_variant", 0)
    assert nrel(out[0][0], out[1][0]) < 1e-5
    assert nrel(out[0][1], out[1][1]) < 1e-5


def test_step_backward_is_deterministic(ctx):
    """The default adjoint sums its grid adjoints in fixed point: two runs
    give bit-identical gradients."""
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(49, (40, 36, 44), 1.0 / 40.0, ppc=6, amplitude=2.0)
    rng = np.random.default_rng(49)
    a = dev_soa(rng.normal(size=sc.pos.shape))
    b = dev_soa(rng.normal(size=sc.pos.shape))
    x, v = dev_soa(sc.pos), dev_soa(sc.vel)
    r1 = [t.cpu().numpy() for t in ctx.flip_step_backward(x, v, a, b, sc.dims, sc.dx, sc.dt)]
    r2 = [t.cpu().numpy() for t in ctx.flip_step_backward(x, v, a, b, sc.dims, sc.dx, sc.dt)]
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])


def _tape_scene(seed, n_far):
    """random_blobs scene plus n_far particles scattered around and outside
    the grid (the sort's overflow bin)."""
    from gridfluidsim3d_amd import scenes
    sc = scenes.random_blobs(seed, (40, 36, 44), 1.0 / 40.0, ppc=6, amplitude=2.0)
    rng = np.random.default_rng(seed)
    ext = np.array(sc.dims) * sc.dx
    far = (rng.uniform(-0.4, 1.4, size=(n_far, 3)) * ext).astype(np.float32)
    pos = np.concatenate([sc.pos, far])
    vel = np.concatenate([sc.vel, rng.normal(size=(n_far, 3)).astype(np.float32)])
    return sc, pos, vel, rng


@pytest.mark.parametrize("seed,n_far", [(51, 0), (52, 300)])
def test_step_record_and_tape_backward_bit_identical(ctx, seed, n_far):
    """gfsx_flip_step_record gives gfsx_flip_step's outputs and grids bit for
    bit, and gfsx_flip_step_backward_tape gives gfsx_flip_step_backward's
    gradients bit for bit (the tape replaces the recompute of the same sort
    and grids)."""
    sc, pos, vel, rng = _tape_scene(seed, n_far)
    dims, dx, dt = sc.dims, sc.dx, sc.dt
    x, v = dev_soa(pos), dev_soa(vel)
    xo, vo, grids = ctx.flip_step_soa(x, v, dims, dx, dt)
    xr, vr, tape = ctx.flip_step_record(x, v, dims, dx, dt)
    for got, want in zip([xr, vr] + tape.grids[0], [xo, vo] + grids):
        assert torch.equal(got, want)
    a = dev_soa(rng.normal(size=pos.shape))
    b = dev_soa(rng.normal(size=pos.shape))
    r = [t.cpu().numpy() for t in ctx.flip_step_backward(x, v, a, b, dims, dx, dt)]
    t = [t.cpu().numpy() for t in ctx.flip_step_backward(x, v, a, b, dims, dx, dt, tape=tape)]
    assert np.array_equal(r[0], t[0]) and np.array_equal(r[1], t[1])
    # the tape outlives later steps on the same context
    ctx.flip_step_record(xr, vr, dims, dx, dt)
    t2 = [t.cpu().numpy() for t in ctx.flip_step_backward(x, v, a, b, dims, dx, dt, tape=tape)]
    assert np.array_equal(r[0], t2[0]) and np.array_equal(r[1], t2[1])


def test_rollout_tape_matches_recompute(ctx):
    """A 6-step autograd rollout with taped steps (the default) and with the
    recomputing backward: the same loss and bit-identical gradients."""
    from gridfluidsim3d_amd import autograd as ga
    sc, pos, vel, rng = _tape_scene(53, 100)
    dims, dx, dt = sc.dims, sc.dx, sc.dt
    target = torch.tensor(rng.uniform(0.0, 4.0, size=(dims[2], dims[1], dims[0])),
                          dtype=torch.float32, device="cuda")
    grads = []
    for tape in (True, False):
        x0 = dev_soa(pos).requires_grad_(True)
        v0 = dev_soa(vel).requires_grad_(True)
        x, _v = ga.rollout(ctx, x0, v0, dims, dx, 6, dt, tape=tape)
        L = ga.density_loss(ctx, x, target, dims, dx)
        L.backward()
        grads.append((L.item(), x0.grad.cpu().numpy(), v0.grad.cpu().numpy()))
    assert grads[0][0] == grads[1][0]
    assert np.array_equal(grads[0][1], grads[1][1]) and np.array_equal(grads[0][2], grads[1][2])


def test_step_tape_empty_step(ctx):
    """A recorded step without particles: empty outputs, zero grids, empty
    gradients, no error."""
    dims, dx, dt = (6, 5, 4), 0.1, 0.01
    x = torch.empty((3, 0), dtype=torch.float32, device="cuda")
    v = torch.empty((3, 0), dtype=torch.float32, device="cuda")
    xo, vo, tape = ctx.flip_step_record(x, v, dims, dx, dt)
    assert xo.shape == (3, 0) and vo.shape == (3, 0)
    for g in tape.grids[2]:
        assert not bool(g.any())
    gx, gv = ctx.flip_step_backward(x, v, x, v, dims, dx, dt, tape=tape)
    assert gx.shape == (3, 0) and gv.shape == (3, 0)


def test_step_tape_rejects_mismatch(ctx):
    import ctypes
    from gridfluidsim3d_amd import gfsx
    from gridfluidsim3d_amd import pybindings as pb
    dims, dx, dt = (6, 6, 6), 0.1, 0.01
    pos, vel = small_scene(46, dims, dx, box=((1, 4), (1, 4), (1, 4)))
    x, v = dev_soa(pos), dev_soa(vel)
    _xo, _vo, tape = ctx.flip_step_record(x, v, dims, dx, dt)
    # a tape of another grid or particle count
    with pytest.raises(ValueError, match="another step"):
        ctx.flip_step_backward(x, v, x, v, (6, 6, 7), dx, dt, tape=tape)
    f = pb.lib_func("gfsx_flip_step_record",
                    [ctypes.c_void_p, ctypes.POINTER(gfsx.Scene), ctypes.POINTER(gfsx.StepParams),
                     ctypes.POINTER(gfsx.StepTape)])
    pr = ctx._params(dt, (0.0, -9.8, 0.0), 0.05, gfsx.TRILINEAR, gfsx.ORDER_KEEP)
    xo, vo = torch.empty_like(x), torch.empty_like(v)
    sc = ctx._soa_scene(x, v, dims, dx, xo, vo)
    t = tape.c_tape()
    with pytest.raises(RuntimeError, match="every grid"):
        pb.execute_lib_func(f, [ctx(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(t)])
    tape.set_grids(sc)
    t.starts -= 1
    with pytest.raises(RuntimeError, match="gfsx_step_tape_starts"):
        pb.execute_lib_func(f, [ctx(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(t)])
    t.starts += 1
    t.n -= 1
    with pytest.raises(RuntimeError, match="particle count"):
        pb.execute_lib_func(f, [ctx(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(t)])


def test_step_backward_position_only_and_velocity_only(ctx):
    """Adjoints of x' alone and of v' alone, each against the dense
    reference (covers the RK4 path and the PIC/FLIP path separately)."""
    dims, dx, dt = (10, 9, 8), 0.1, 1.0 / 60.0
    pos, vel = small_scene(44, dims, dx)
    rng = np.random.default_rng(9)
    a = rng.normal(size=pos.shape)
    for which in ("x", "v"):
        za = a if which == "x" else np.zeros_like(a)
        zb = a if which == "v" else np.zeros_like(a)
        gx, gv = ctx.flip_step_backward(dev_soa(pos), dev_soa(vel), dev_soa(za), dev_soa(zb),
                                        dims, dx, dt)
        x = torch.tensor(pos, dtype=F64, requires_grad=True)
        v = torch.tensor(vel, dtype=F64, requires_grad=True)
        x1, v1 = dfl.step(x, v, dims, dx, dt)
        L = (torch.tensor(za) * x1).sum() + (torch.tensor(zb) * v1).sum()
        rx, rv = torch.autograd.grad(L, (x, v))
        assert nrel(gx.cpu().numpy().T, rx.numpy()) < STEP_TOL, which
        assert nrel(gv.cpu().numpy().T, rv.numpy()) < STEP_TOL, which


@pytest.mark.parametrize("radius_cells,with_val", [(1.0, True), (1.5, True), (1.0, False)])
def test_p2g_backward_matches_dense(ctx, radius_cells, with_val):
    rng = np.random.default_rng(51)
    dims, dx = (9, 8, 7), 0.125
    n = 300
    pos = (rng.uniform(-0.1, 1.1, size=(n, 3)) * np.array(dims) * dx).astype(np.float32)
    val = rng.normal(size=n).astype(np.float32) if with_val else None
    off = (0.5 * dx, 0.0, 0.25 * dx)
    R = radius_cells * dx
    gS = rng.normal(size=(dims[2], dims[1], dims[0])).astype(np.float32)
    gW = rng.normal(size=(dims[2], dims[1], dims[0])).astype(np.float32)
    tpos = torch.tensor(pos, device="cuda")
    tval = None if val is None else torch.tensor(val, device="cuda")
    gpos, gval = ctx.p2g_backward_device(tpos, tval, R, off, dx, dims,
                                         gS=torch.tensor(gS, device="cuda"),
                                         gW=torch.tensor(gW, device="cuda"))
    x = torch.tensor(pos, dtype=F64, requires_grad=True)
    vv = None if val is None else torch.tensor(val, dtype=F64, requires_grad=True)
    S, W = dfl.splat(x, vv, dims, off, dx, R)
    L = (torch.tensor(gS, dtype=F64) * S).sum() + (torch.tensor(gW, dtype=F64) * W).sum()
    if vv is None:
        (rx,) = torch.autograd.grad(L, (x,))
    else:
        rx, rval = torch.autograd.grad(L, (x, vv))
        assert nrel(gval.cpu().numpy(), rval.numpy()) < 1e-5
    assert nrel(gpos.cpu().numpy(), rx.numpy()) < 1e-5


def test_rollout_autograd_matches_dense(ctx):
    """4 steps through FlipStepFunction plus the density loss: gradients
    w.r.t. the initial positions and velocities against the dense FP64
    rollout."""
    from gridfluidsim3d_amd import autograd as ga
    dims, dx, dt = (10, 9, 8), 0.1, 1.0 / 60.0
    pos, vel = small_scene(45, dims, dx)
    rng = np.random.default_rng(145)
    target = rng.uniform(0.0, 2.0, size=(dims[2], dims[1], dims[0]))
    x0 = dev_soa(pos).requires_grad_(True)
    v0 = dev_soa(vel).requires_grad_(True)
    x, _v = ga.rollout(ctx, x0, v0, dims, dx, 4, dt)
    L = ga.density_loss(ctx, x, torch.tensor(target, dtype=torch.float32, device="cuda"),
                        dims, dx)
    L.backward()
    xr = torch.tensor(pos, dtype=F64, requires_grad=True)
    vr = torch.tensor(vel, dtype=F64, requires_grad=True)
    xx, vv = xr, vr
    for _ in range(4):
        xx, vv = dfl.step(xx, vv, dims, dx, dt)
    Lr = ((dfl.density(xx, dims, dx) - torch.tensor(target)) ** 2).sum()
    rx, rv = torch.autograd.grad(Lr, (xr, vr))
    assert abs(L.item() - Lr.item()) <= 1e-4 * abs(Lr.item())
    ex = nrel(x0.grad.cpu().numpy().T, rx.numpy())
    ev = nrel(v0.grad.cpu().numpy().T, rv.numpy())
    print("rollout: dL/dx0 err %.2e, dL/dv0 err %.2e" % (ex, ev))
    assert ex < 1e-4 and ev < 1e-4


def test_rollout_gradient_matches_fd_subset32(ctx, orc64):
    """The configuration's gradient check: 16 steps on the 32^3 / 20,480
    particle subset, density loss against a seeded smooth target, gradient
    w.r.t. the initial velocities vs FP64 central finite differences of the
    C oracle along random directions."""
    from gridfluidsim3d_amd import autograd as ga
    from gridfluidsim3d_amd import scenes
    sc = scenes.slab_subset32()
    dims, dx, dt, steps = sc.dims, sc.dx, sc.dt, 16
    target = scenes.density_target(dims, dx, 5, y_range=(6, 26))
    x0 = dev_soa(sc.pos)
    v0 = dev_soa(sc.vel).requires_grad_(True)
    x, _v = ga.rollout(ctx, x0, v0, dims, dx, steps, dt)
    L = ga.density_loss(ctx, x, torch.tensor(target, device="cuda"), dims, dx)
    L.backward()
    g = v0.grad.cpu().numpy().T.astype(np.float64)
    rng = np.random.default_rng(77)
    # small enough that no particle crosses a cell or support boundary
    # (where the loss jumps) in the perturbed rollouts, large enough for
    # FP64 rounding (the quotient agrees to ~1e-6 for eps in 1e-7..1e-6)
    eps = 1e-7
    worst = 0.0
    for _ in range(3):
        d = rng.normal(size=sc.vel.shape)
        lp = oracle_rollout_loss(orc64, sc.pos, sc.vel + eps * d, dims, dx, dt, steps, target)
        lm = oracle_rollout_loss(orc64, sc.pos, sc.vel - eps * d, dims, dx, dt, steps, target)
        fd = (lp - lm) / (2 * eps)
        an = float((g * d).sum())
        err = abs(fd - an) / abs(fd)
        worst = max(worst, err)
        print("FD %.8e  adjoint %.8e  rel %.2e" % (fd, an, err))
    assert worst < ROLLOUT_TOL


def test_step_backward_rejects_unsupported(ctx):
    from gridfluidsim3d_amd import gfsx
    dims, dx = (6, 6, 6), 0.1
    pos, vel = small_scene(46, dims, dx, box=((1, 4), (1, 4), (1, 4)))
    x, v = dev_soa(pos), dev_soa(vel)
    sc = ctx._soa_scene(x, v, dims, dx)
    g = gfsx.SceneGrad()
    import ctypes
    from gridfluidsim3d_amd import pybindings as pb
    f = pb.lib_func("gfsx_flip_step_backward",
                    [ctypes.c_void_p, ctypes.POINTER(gfsx.Scene), ctypes.POINTER(gfsx.StepParams),
                     ctypes.POINTER(gfsx.SceneGrad)])
    for stencil, order, msg in ((gfsx.TRICUBIC, gfsx.ORDER_KEEP, "trilinear"),
                                (gfsx.TRILINEAR, gfsx.ORDER_SORTED, "input order")):
        pr = ctx._params(0.01, (0.0, -9.8, 0.0), 0.05, stencil, order)
        with pytest.raises(RuntimeError, match=msg):
            pb.execute_lib_func(f, [ctx(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(g)])
    pr = ctx._params(0.01, (0.0, -9.8, 0.0), 0.05, gfsx.TRILINEAR, gfsx.ORDER_KEEP)
    with pytest.raises(RuntimeError, match="null"):
        pb.execute_lib_func(f, [ctx(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(g)])
