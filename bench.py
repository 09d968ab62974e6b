#!/usr/bin/env python3
"""Throughput of the plain FLIP transfer + advection step on one GPU.

One step = brick sort -> P2G (u, v, w with weights, normalisation, snapshot,
body force on fluid-bordering faces) -> G2P with PIC/FLIP blend and RK4
advection, over every particle of the scene, inputs resident in HBM.
The default workload is the 256^3 pool-plus-spheres scene (~43M particles,
dx = 1/32, dt = 1/240, curl noise at 3 m/s, seed 1003), trilinear stencil.

Prints ONE JSON line (rank 0) with the throughput in particle-steps/s, the
roofline of the dominant kernel (HIP events on the library's stream), the
whole-step roofline, and a CPU baseline: the reference's own CPU code
(oracle/_ref, single thread) run on a bounded x-slab of the same scene.
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FLIP transfer+advect particle-steps/s at 256³; batched & fwd+bwd vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def face_count(dims):
    ni, nj, nk = dims
    return (ni + 1) * nj * nk + ni * (nj + 1) * nk + ni * nj * (nk + 1)


def algorithmic_bytes(P, F, C):
    """Bytes each stage must move at minimum (FP32 / uint32):
    sort: read positions and velocities (24P), write them in brick order
          (24P), one pass over the bins' counts and starts (16C, C bins:
          64-bit counters);
    p2g : read x, v once (24P) + write the field and its pre-force
          snapshot, one value per face each (8F);
    g2p : read x, v (24P) + write x', v' (24P) + read the field and the
          snapshot once (8F);
    step: the whole step: read x, v and write x', v' (48P), write the field
          and read it back (8F)."""
    return dict(sort=48 * P + 16 * C, p2g=24 * P + 8 * F, g2p=48 * P + 8 * F,
                step=48 * P + 8 * F)


def brick_bins(dims):
    """Bins of the brick sort: 8 per G2P brick (8^3 cells) plus the overflow
    bin."""
    g = [((d + 3) // 4 + 1) // 2 for d in dims]
    return 8 * g[0] * g[1] * g[2] + 1


# MI355X's published FP32 vector rate; it counts packed FP32 instructions
# (two FMAs per lane), which these kernels are built without, so their
# arithmetic ceiling is half of it.
FP32_PEAK_TFLOPS = 157.3


def flops_per_particle(stencil):
    """Useful arithmetic of one plain step per particle, counted from the
    kernels' formulas with an FMA as two FLOPs (multiply + add); compares,
    floors, clamps and index arithmetic are not counted, so it is a lower
    bound on the work done.
    P2G, per (node, particle) pair: offset (3), squared distance (5),
    t = d2/R^2 (1), Wyvill a^2 (5/9 + 4a/9) with a = 1 - t (5), S and W sums
    (3) = 17.  With R = dx a particle reaches the 2x2x2 nodes of each
    component around it: 24 pairs = 408.  (The gather kernel evaluates the
    2x3x3 cells of candidates per face, 54 pairs per particle; only the 24
    are counted.)
    G2P: a trilinear component sample is 7 lerps of 3 = 21; a tricubic one
    21 Catmull-Rom evaluations of 18 = 378.  PIC/FLIP samples the field and
    the snapshot (6 component samples) and RK4 three more positions (9);
    the blend is 15 and the RK4 combinations 39."""
    sample = 21 if stencil == "trilinear" else 378
    return dict(p2g=24 * 17, g2p=15 * sample + 15 + 39)


# The kernels each stage of gfsx_flip_step launches (HIP events bracket the
# stage on the library's stream).
STAGE_KERNELS = {
    "sort": "brick sort: k_bsort_bin, k_tile_scan64, k_bsort_list, k_bsort_order_*, "
            "k_bsort_scatter",
    "p2g": "k_p2g_bscatter + k_p2g_bcombine",
    "g2p": "k_flip_g2p_bb_tl (trilinear) / k_flip_g2p_bb_tc (tricubic)",
}


def kernel_source_hash():
    """sha1 over the HIP sources the library is built from."""
    h = hashlib.sha1()
    for f in sorted(glob.glob(os.path.join(ROOT, "gridfluidsim3d_amd", "csrc", "*"))):
        with open(f, "rb") as fh:
            h.update(os.path.basename(f).encode())
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic.json, which
    holds rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch scaled by the same
    counters' reading for a plain copy of known size.  Only used when it was
    measured on the current kernel sources."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("src_sha1") != kernel_source_hash():
        return None
    ent = rec.get("workloads", {}).get(workload, {}).get(kernel)
    return None if ent is None else ent.get("hbm_bytes")


def copy_calibration(nbytes, reps=5):
    """A plain device-to-device copy of nbytes, for scaling the FETCH_SIZE /
    WRITE_SIZE counters (each copy reads and writes nbytes once)."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "reps": reps}), flush=True)


def cpu_baseline(scene, stencil, target_s=15.0, steps=None, gpu_check=False):
    """Time the reference's CPU path (its own compiled sources, one thread).

    steps=None: one step on an x-slab of the scene sized for ~target_s
    seconds.  steps=k: k steps of the whole scene.  With gpu_check, the GPU
    runs the same first step on the same particles and the record carries the
    per-field norm-wise error of the GPU against the reference's output
    (max|gpu - ref| / max|ref|), so the benchmark scene itself is checked."""
    import numpy as np
    from oracle import oracle
    lib = oracle.load_reference()
    kind = "reference"
    if lib is None:
        lib = oracle.load_oracle()
        kind = "port"
    ni, nj, nk = scene.dims

    def run(cells):
        keep = scene.pos[:, 0] < np.float32(cells * scene.dx)
        pos = np.ascontiguousarray(scene.pos[keep])
        vel = np.ascontiguousarray(scene.vel[keep])
        t0 = time.perf_counter()
        out = lib.flip_step(pos, vel, (cells, nj, nk), scene.dx, scene.dt, scene.body_force,
                            0.05, stencil)
        return pos, vel, out, time.perf_counter() - t0

    if steps is None:
        cells = max(1, ni // 64)
        pos, vel, out, t = run(cells)
        if t > 0 and cells < ni:
            cells = int(min(ni, max(cells, round(cells * target_s / t))))
            pos, vel, out, t = run(cells)
        p = pos.shape[0]
        rec = dict(value=p / t, unit="particle-steps/s", cores=1, kind=kind,
                   sample="one plain step on the x-slab [0,%d) x %d x %d of the same scene "
                          "(%d particles), %.1f s" % (cells, nj, nk, p, t),
                   nproc=os.cpu_count(), cpu=_cpu_model())
        dims = (cells, nj, nk)
    else:
        pos = np.ascontiguousarray(scene.pos)
        vel = np.ascontiguousarray(scene.vel)
        first = None
        t = 0.0
        x, v = pos, vel
        for k in range(steps):
            t0 = time.perf_counter()
            o = lib.flip_step(x, v, scene.dims, scene.dx, scene.dt, scene.body_force, 0.05,
                              stencil)
            t += time.perf_counter() - t0
            first = o if first is None else first
            x, v = o["pos"], o["vel"]
        out = first
        p = pos.shape[0]
        rec = dict(value=p * steps / t, unit="particle-steps/s", cores=1, kind=kind,
                   sample="%d plain steps of the whole scene (%d particles), %.1f s"
                          % (steps, p, t),
                   nproc=os.cpu_count(), cpu=_cpu_model())
        dims = scene.dims
    if gpu_check:
        rec["gpu_vs_cpu"] = gpu_step_errors(pos, vel, dims, scene.dx, scene.dt,
                                            scene.body_force, stencil, out)
    return rec


def gpu_step_errors(pos, vel, dims, dx, dt, body_force, stencil, ref):
    """Norm-wise error max|gpu - ref| / max|ref| per field of one GPU step
    (GFSX_ORDER_KEEP) against a CPU step's output dict."""
    import numpy as np
    import torch

    import gridfluidsim3d_amd as gf
    ctx = gf.Context(0)
    st = gf.FlipState(pos, vel, dims, dx, keep_old=True, keep_weights=True)
    ctx.flip_step(st, dt, body_force, 0.05, stencil, gf.ORDER_KEEP)
    torch.cuda.synchronize()
    got = dict(u=st.grid[0], v=st.grid[1], w=st.grid[2], u_old=st.grid_old[0],
               wu=st.weights[0], vel=st.velocities(), pos=st.positions())
    out = {}
    for k, t in got.items():
        a = t.cpu().numpy().astype(np.float64)
        b = np.asarray(ref[k], np.float64)
        den = np.abs(b).max()
        out[k] = float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())
    x0 = pos.astype(np.float64)
    a = got["pos"].cpu().numpy().astype(np.float64) - x0
    b = np.asarray(ref["pos"], np.float64) - x0
    out["displacement"] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    del st
    ctx.close()
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def parity_report(log_fn):
    """Parity of the fused step on the CPU-runnable configuration (64^3
    sphere, 268,416 particles, dt = 1/120) for both stencils: per field
    (u, v, w after the body force; the u-face weights W; particle velocity;
    particle displacement) the norm-wise relative error max|gpu-cpu|/max|cpu|
    against the reference's own CPU code (oracle/_ref; the C restatement
    where that was not built), the largest per-element relative error over
    elements with |cpu| >= 1e-3 max|cpu|, and the norm-wise errors of the GPU
    and of the CPU path against the FP64 evaluation of the same step."""
    import numpy as np
    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes
    from oracle import oracle

    sc = scenes.sphere64()
    cpu = oracle.load_reference()
    kind = "reference"
    if cpu is None:
        cpu = oracle.load_oracle()
        kind = "port"
    o64 = oracle.Oracle64()
    ctx = gf.Context(0)

    def nrel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        den = np.abs(b).max()
        return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())

    def erel(a, b):
        a = np.asarray(a, np.float64).ravel()
        b = np.asarray(b, np.float64).ravel()
        m = np.abs(b) >= 1e-3 * np.abs(b).max()
        return float((np.abs(a[m] - b[m]) / np.abs(b[m])).max()) if m.any() else 0.0

    out = dict(workload="sphere64", particles=sc.n, cpu_kind=kind, bar=1e-5)
    for stencil, name in ((gf.TRILINEAR, "trilinear"), (gf.TRICUBIC, "tricubic")):
        st = gf.FlipState(sc.pos, sc.vel, sc.dims, sc.dx, keep_weights=True)
        ctx.flip_step(st, sc.dt, sc.body_force, 0.05, stencil, gf.ORDER_KEEP)
        torch.cuda.synchronize()
        g = dict(u=st.grid[0].cpu().numpy(), v=st.grid[1].cpu().numpy(),
                 w=st.grid[2].cpu().numpy(), W=st.weights[0].cpu().numpy(),
                 vel=st.velocities().cpu().numpy(), pos=st.positions().cpu().numpy())
        c = cpu.flip_step(sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, sc.body_force, 0.05, stencil)
        c = dict(u=c["u"], v=c["v"], w=c["w"], W=c["wu"], vel=c["vel"], pos=c["pos"])
        d = o64.flip_step(sc.pos, sc.vel, sc.dims, sc.dx, sc.dt, sc.body_force, 0.05, stencil)
        d = dict(u=d["u"], v=d["v"], w=d["w"], W=d["wu"], vel=d["vel"], pos=d["pos"])
        fields = {}
        for k in ("u", "v", "w", "W", "vel", "pos"):
            fields[k] = dict(norm_rel=nrel(g[k], c[k]), elem_rel=erel(g[k], c[k]),
                             gpu_vs_fp64=nrel(g[k], d[k]), cpu_vs_fp64=nrel(c[k], d[k]))
        # the displacement x' - x, for information: its error is set by the
        # FP32 rounding of the positions themselves (1 ulp of |x| ~ 4 m is
        # 4.8e-7 m against displacements of a few cm), which the CPU path
        # shows against FP64 as well
        x0 = sc.pos.astype(np.float64)
        fields["displacement_info"] = dict(
            norm_rel=nrel(g["pos"] - x0, c["pos"] - x0),
            gpu_vs_fp64=nrel(g["pos"] - x0, d["pos"] - x0),
            cpu_vs_fp64=nrel(c["pos"] - x0, d["pos"] - x0))
        out[name] = fields
        worst = max(f["norm_rel"] for k, f in fields.items() if k != "displacement_info")
        out[name + "_worst_norm_rel"] = worst
        log_fn("parity %s: worst norm-wise %.2e (bar 1e-5)" % (name, worst))
    ctx.close()
    return out


def batched_measure(steps, stencil, log_fn):
    """The batched form (64 independent scenes of uneven size): particle-steps/s
    of 64 sequential plain steps per step, of the batched step (after the
    autotuner picked its variants), and of the batched step replayed from HIP
    graphs, each timed over `steps` steps from the scenes' initial state."""
    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import autotune, scenes

    t0 = time.time()
    specs = scenes.batch64()
    P = sum(sc.n for sc in specs)
    F = sum(face_count(sc.dims) for sc in specs)
    log_fn("batch64: %d scenes, %d particles, %d faces, built in %.1f s"
           % (len(specs), P, F, time.time() - t0))
    ctx = gf.Context(0)
    dts = [sc.dt for sc in specs]

    def fresh():
        return [gf.FlipState(sc.pos, sc.vel, sc.dims, sc.dx) for sc in specs]

    def timed(run, states):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(steps):
            run(states)
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps

    def sequential(states):
        for st, dt in zip(states, dts):
            ctx.flip_step(st, dt, (0.0, -9.8, 0.0), 0.05, stencil, gf.ORDER_SORTED)

    def batched(states):
        ctx.flip_step_batched(states, dts, (0.0, -9.8, 0.0), 0.05, stencil, gf.ORDER_SORTED)

    # warm both paths (workspace, tables) on scratch states
    warm = fresh()
    sequential(warm)
    batched(warm)
    del warm
    ms_seq = timed(sequential, fresh())
    tuned = autotune.tune(ctx, fresh(), dts, (0.0, -9.8, 0.0), 0.05, stencil, gf.ORDER_SORTED)
    log_fn("autotune: %s" % tuned)
    ms_bat = timed(batched, fresh())
    states = fresh()
    replay = ctx.capture_batched(states, dts, (0.0, -9.8, 0.0), 0.05, stencil, gf.ORDER_SORTED)
    # the capture advanced the states two steps; time the replays from a
    # fresh start by copying the initial state into the captured buffers
    init = fresh()
    for st, st0 in zip(states, init):
        for a in range(3):
            st.cur_x[a].copy_(st0.cur_x[a])
            st.cur_v[a].copy_(st0.cur_v[a])
    del init
    ms_graph = timed(lambda _: replay(), states)
    ctx.close()
    alg = sum(48 * sc.n + 8 * face_count(sc.dims) for sc in specs)
    out = dict(workload="batch64", scenes=len(specs), particles=P, faces=F, steps=steps,
               ms_per_step_sequential=round(ms_seq, 4), ms_per_step_batched=round(ms_bat, 4),
               ms_per_step_graph=round(ms_graph, 4),
               value_sequential=P / (ms_seq / 1e3), value_batched=P / (ms_bat / 1e3),
               value_graph=P / (ms_graph / 1e3), unit="particle-steps/s",
               speedup_graph_vs_sequential=round(ms_seq / ms_graph, 3),
               autotune={k: v[0] for k, v in tuned.items()},
               step_roofline_graph=dict(achieved=round(alg / (ms_graph / 1e3) / 1e9, 1),
                                        peak=HBM_PEAK_GBS, unit="GB/s",
                                        frac=round(alg / (ms_graph / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                                        algorithmic_bytes=alg))
    log_fn("batched: %s" % out)
    return out


def fwdbwd_measure(reps, log_fn, steps=16, opt_iters=3):
    """The differentiable form: a 16-step rollout of the 96^3 jittered slab
    (1,990,656 particles, dx = 1/96) through the autograd step, the density
    loss against a seeded smooth target, and the backward pass through the
    adjoint kernels to the initial velocities.  Reports particle-steps/s of
    forward+backward (P * steps / time; each step keeps its sort and grids
    for the backward pass) next to the same rollout forward only and to the
    backward that recomputes them, and a few gradient-descent iterations on
    the initial velocities."""
    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import autograd as ga
    from gridfluidsim3d_amd import scenes

    t0 = time.time()
    sc = scenes.slab()
    dims, dx, dt = sc.dims, sc.dx, sc.dt
    P, F = sc.n, face_count(dims)
    target = torch.tensor(scenes.density_target(dims, dx, 7, y_range=(20, 47)), device="cuda")
    log_fn("slab96: %d particles, %d faces, built in %.1f s" % (P, F, time.time() - t0))
    ctx = gf.Context(0)
    x0 = torch.tensor(sc.pos.T.copy(), device="cuda")
    v0 = torch.tensor(sc.vel.T.copy(), device="cuda")

    def fwd(v_init=v0):
        with torch.no_grad():
            x, _v = ga.rollout(ctx, x0, v_init, dims, dx, steps, dt)
            return ga.density_loss(ctx, x, target, dims, dx)

    def fwd_bwd(v_init, tape=True):
        v = v_init.detach().clone().requires_grad_(True)
        x, _v = ga.rollout(ctx, x0, v, dims, dx, steps, dt, tape=tape)
        L = ga.density_loss(ctx, x, target, dims, dx)
        L.backward()
        return L.detach(), v.grad

    def timed(fn):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    fwd()
    fwd_bwd(v0)
    fwd_bwd(v0, tape=False)
    ms_f = timed(fwd)
    ms_fb = timed(lambda: fwd_bwd(v0))
    ms_fb_re = timed(lambda: fwd_bwd(v0, tape=False))
    # gradient descent on the initial velocities, step length fixed in
    # velocity units (0.05 m/s on the largest gradient component)
    losses = []
    v = v0
    for _ in range(opt_iters):
        L, g = fwd_bwd(v)
        losses.append(float(L))
        v = v - 0.05 * g / g.abs().max()
    losses.append(float(fwd(v)))
    ctx.close()
    alg = (120 * P + 24 * F) * steps
    gbs = alg / (ms_fb / 1e3) / 1e9
    out = dict(workload="slab96", grid=list(dims), dx=dx, particles=P, steps=steps, reps=reps,
               ms_forward=round(ms_f, 3), ms_forward_backward=round(ms_fb, 3),
               ms_forward_backward_recompute=round(ms_fb_re, 3),
               value_forward=P * steps / (ms_f / 1e3),
               value_forward_backward=P * steps / (ms_fb / 1e3), unit="particle-steps/s",
               roofline_forward_backward=dict(achieved=round(gbs, 1), peak=HBM_PEAK_GBS,
                                              unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                                              algorithmic_bytes=alg),
               loss_descent=[round(l, 3) for l in losses])
    log_fn("fwd+bwd: %s" % out)
    return out


def plain_measure(ctx, scene, stencil, steps, warmup, reps, epilogue=0):
    """Plain steps of one scene: warm-up steps on a scratch copy, then `reps`
    timed runs of `steps` steps, each from the scene's initial state (HIP
    events on the library's stream around the step loop; per-stage events
    inside the library).  Reports the median run."""
    import torch

    import gridfluidsim3d_amd as gf
    stream = torch.cuda.current_stream()

    def step(st):
        ctx.flip_step(st, scene.dt, scene.body_force, 0.05, stencil, gf.ORDER_SORTED, epilogue)

    st = gf.FlipState(scene.pos, scene.vel, scene.dims, scene.dx)
    for _ in range(warmup):
        step(st)
    torch.cuda.synchronize()
    del st
    runs = []
    for _ in range(reps):
        st = gf.FlipState(scene.pos, scene.vel, scene.dims, scene.dx)
        torch.cuda.synchronize()
        ctx.timing_begin()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            step(st)
        end.record(stream)
        torch.cuda.synchronize()
        nsteps, stage_ms = ctx.timing_end()
        assert nsteps == steps
        runs.append((start.elapsed_time(end) / steps, {k: v / nsteps for k, v in stage_ms.items()}))
        del st
    runs_sorted = sorted(runs, key=lambda r: r[0])
    ms, stages = runs_sorted[len(runs) // 2]
    P, F = scene.n, face_count(scene.dims)
    alg = algorithmic_bytes(P, F, brick_bins(scene.dims))
    gbs = alg["step"] / (ms / 1e3) / 1e9
    return dict(workload=scene.name, particles=P, faces=F, steps=steps,
                ms_per_step=ms, value=P / (ms / 1e3), unit="particle-steps/s",
                ms_per_step_runs=[round(r[0], 4) for r in runs],
                stages_ms={k: round(v, 4) for k, v in stages.items()},
                step_roofline=dict(achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                                   frac=round(gbs / HBM_PEAK_GBS, 4),
                                   algorithmic_bytes=alg["step"]))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reps", type=int, default=3,
                    help="timed runs of --steps steps, each from the scene's initial state; "
                         "the median is reported")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 64^3 and 128^3 workloads (GPU and CPU)")
    ap.add_argument("--workload", default="pool256",
                    choices=["pool256", "dambreak128", "sphere64"])
    ap.add_argument("--stencil", default="trilinear", choices=["trilinear", "tricubic"])
    ap.add_argument("--epilogue", default="none", choices=["none", "fluidsim"],
                    help="fluidsim: the step with FluidSimulation's face fill and "
                         "velocity extrapolation (7 layers)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the parity report")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tricubic", action="store_true",
                    help="skip the tricubic-stencil measurement (same scene, 5 steps)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-form measurement (64 scenes)")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--no-fwdbwd", action="store_true",
                    help="skip the differentiable-form measurement (96^3 slab)")
    ap.add_argument("--fwdbwd-reps", type=int, default=3)
    ap.add_argument("--only", choices=["batched", "fwdbwd"], default=None,
                    help="run only that form's measurement and print its record "
                         "(for profiling it alone)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain copy of this many bytes (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(requested --gpus %d, WORLD_SIZE %d)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return

    import torch

    import gridfluidsim3d_amd as gf
    from gridfluidsim3d_amd import scenes

    stencil = gf.TRILINEAR if args.stencil == "trilinear" else gf.TRICUBIC
    if args.only == "batched":
        print(json.dumps(batched_measure(args.batched_steps, stencil, log)), flush=True)
        return
    if args.only == "fwdbwd":
        print(json.dumps(fwdbwd_measure(args.fwdbwd_reps, log)), flush=True)
        return
    t0 = time.time()
    scene = scenes.CONFIGS[args.workload]()
    P = scene.n
    F = face_count(scene.dims)
    C = brick_bins(scene.dims)
    log("scene %s: %d particles, %d faces, built in %.1f s" % (scene.name, P, F, time.time() - t0))

    ctx = gf.Context(0)
    log(ctx.device_info)
    epilogue = gf.EPILOGUE_FLUIDSIM if args.epilogue == "fluidsim" else gf.EPILOGUE_NONE

    head = plain_measure(ctx, scene, stencil, args.steps, args.warmup, args.reps, epilogue)
    value = head["value"]
    ms_per_step = head["ms_per_step"]
    stage_avg = head["stages_ms"]
    log("ms/step %.3f (median of %d runs: %s)  stages %s" % (
        ms_per_step, args.reps, head["ms_per_step_runs"], stage_avg))

    alg = algorithmic_bytes(P, F, C)
    dom = max(("sort", "p2g", "g2p"), key=lambda k: stage_avg[k])
    achieved = alg[dom] / (stage_avg[dom] / 1e3) / 1e9
    roofline = dict(bound="hbm", kernel=STAGE_KERNELS[dom],
                    achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=measured_traffic(args.workload, dom),
                    algorithmic_bytes=alg[dom], avg_ms=round(stage_avg[dom], 4))
    step_roofline = head["step_roofline"]

    fpp = flops_per_particle(args.stencil)
    compute = dict(flops_per_particle=fpp["p2g"] + fpp["g2p"],
                   p2g_tflops=round(fpp["p2g"] * P / (stage_avg["p2g"] / 1e3) / 1e12, 2),
                   g2p_tflops=round(fpp["g2p"] * P / (stage_avg["g2p"] / 1e3) / 1e12, 2),
                   step_tflops=round((fpp["p2g"] + fpp["g2p"]) * P / (ms_per_step / 1e3) / 1e12, 2),
                   peak_fp32_tflops=FP32_PEAK_TFLOPS,
                   peak_fp32_unpacked_tflops=FP32_PEAK_TFLOPS / 2)
    compute["step_frac_fp32"] = round(compute["step_tflops"] / FP32_PEAK_TFLOPS, 4)
    st = None

    tricubic = None
    if not args.no_tricubic and args.stencil == "trilinear":
        # the ParticleAdvector stencil on the same scene: 1 warm-up step on a
        # copy, then 5 steps from the initial state
        torch.cuda.empty_cache()
        tc = plain_measure(ctx, scene, gf.TRICUBIC, 5, 1, 1)
        tf = flops_per_particle("tricubic")
        g2p_tf = tf["g2p"] * P / (tc["stages_ms"]["g2p"] / 1e3) / 1e12
        tricubic = dict(steps=5, ms_per_step=tc["ms_per_step"], value=tc["value"],
                        unit="particle-steps/s", stages_ms=tc["stages_ms"],
                        g2p_flops_per_particle=tf["g2p"], g2p_tflops=round(g2p_tf, 2),
                        g2p_frac_fp32=round(g2p_tf / FP32_PEAK_TFLOPS, 4))
        log("tricubic: %s" % tricubic)

    configs_out = None
    if not args.no_configs and args.workload == "pool256":
        # the other plain workloads: 64^3 sphere and 128^3 dam break, both
        # stencils, 10 steps from the initial state after 2 warm-up steps
        configs_out = {}
        for name in ("sphere64", "dambreak128"):
            sc = scenes.CONFIGS[name]()
            for stn, sname in ((gf.TRILINEAR, "trilinear"), (gf.TRICUBIC, "tricubic")):
                torch.cuda.empty_cache()
                r = plain_measure(ctx, sc, stn, 10, 2, 1)
                r.pop("ms_per_step_runs", None)
                configs_out["%s_%s" % (name, sname)] = r
                log("%s %s: %.3f ms/step, %.3e particle-steps/s, step roofline %.4f" % (
                    name, sname, r["ms_per_step"], r["value"], r["step_roofline"]["frac"]))
            del sc

    batched_out = None
    if not args.no_batched:
        torch.cuda.empty_cache()
        batched_out = batched_measure(args.batched_steps, stencil, log)
        batched_out["vs_plain"] = round(batched_out["value_graph"] / value, 4)

    fwdbwd_out = None
    if not args.no_fwdbwd:
        torch.cuda.empty_cache()
        fwdbwd_out = fwdbwd_measure(args.fwdbwd_reps, log)
        fwdbwd_out["vs_plain"] = round(fwdbwd_out["value_forward_backward"] / value, 4)
        fwdbwd_out["vs_forward"] = round(fwdbwd_out["value_forward_backward"] /
                                         fwdbwd_out["value_forward"], 4)

    cpu = None
    parity = None
    cpu_configs = None
    if not args.no_cpu:
        parity = parity_report(log)
        log("CPU baseline ...")
        cpu = cpu_baseline(scene, stencil, args.cpu_seconds, gpu_check=True)
        log("CPU baseline: %s" % cpu)
        if args.workload == "pool256" and not args.no_configs:
            cpu_configs = {}
            for name, k in (("sphere64", 10), ("dambreak128", 1)):
                sc = scenes.CONFIGS[name]()
                cpu_configs[name] = cpu_baseline(sc, stencil, steps=k, gpu_check=True)
                log("CPU baseline %s: %s" % (name, cpu_configs[name]))
                del sc

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(scene.describe(), stencil=args.stencil,
                       form="plain" if args.epilogue == "none" else "plain+fluidsim-epilogue",
                       pic_ratio=0.05, body_force=list(scene.body_force)),
        "roofline": roofline,
        "step_roofline": step_roofline,
        "compute": compute,
        "tricubic": tricubic,
        "stages_ms": stage_avg,
        "ms_per_step_runs": head["ms_per_step_runs"],
        "cpu_baseline": cpu,
        "cpu_baseline_configs": cpu_configs,
        "parity": parity,
        "configs": configs_out,
        "batched": batched_out,
        "fwd_bwd": fwdbwd_out,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
