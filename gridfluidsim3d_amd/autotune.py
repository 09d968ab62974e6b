"""Autotuner for the fused step: times the kernel variants of
``gfsx_flip_step`` / ``gfsx_flip_step_batched`` on the caller's workload and
keeps the fastest.

The variants are selected with ``Context.set_option``; every variant gives
the same result for G2P and results within FP32 rounding of each other for
P2G (the gathers sum the candidates of a face in the same order with the
arithmetic contracted differently; the scatter sums in fixed point), so the
choice is about speed only.  Timing runs on copies of the states, so the
caller's states do not advance.

Run-to-run bit identity holds for a fixed choice of variants.  A timed
choice can differ between runs (timing noise), and with it the rounding of
the P2G sums; a caller that needs repeatable bits keeps the choice ``tune``
returns and applies it again with ``apply`` instead of timing.
"""
import copy

# option -> candidate values (direct variants take one scene only)
CANDIDATES = {
    "p2g_variant": (4, 0, 2, 3),
    "g2p_brick": (8, 4),
}


def _clone(st):
    out = copy.copy(st)
    out.x = [[t.clone() for t in b] for b in st.x]
    out.v = [[t.clone() for t in b] for b in st.v]
    out.grid = [t.clone() for t in st.grid]
    out.grid_old = None if st.grid_old is None else [t.clone() for t in st.grid_old]
    out.weights = None if st.weights is None else [t.clone() for t in st.weights]
    out.ids = None if st.ids is None else [t.clone() for t in st.ids]
    return out


def tune(ctx, states, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05, stencil=0,
         order=0, steps=3, candidates=None):
    """Pick, one option at a time, the value with the lowest time per step
    over ``steps`` steps (after one untimed step) of the batched step on
    copies of ``states``; leave the winners set on ``ctx``.  Returns
    {option: (best value, {value: ms per step})}."""
    import torch
    candidates = CANDIDATES if candidates is None else candidates
    result = {}
    for opt, values in candidates.items():
        times = {}
        for v in values:
            ctx.set_option(opt, v)
            work = [_clone(st) for st in states]
            ctx.flip_step_batched(work, dt, body_force, pic_ratio, stencil, order)
            torch.cuda.synchronize()
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(steps):
                ctx.flip_step_batched(work, dt, body_force, pic_ratio, stencil, order)
            t1.record()
            torch.cuda.synchronize()
            times[v] = t0.elapsed_time(t1) / steps
            del work
        best = min(times, key=times.get)
        ctx.set_option(opt, best)
        result[opt] = (best, times)
    return result


def choice(result):
    """{option: value} of a ``tune`` result (the form ``apply`` takes)."""
    return {opt: best for opt, (best, _times) in result.items()}


def apply(ctx, chosen):
    """Set a recorded choice ({option: value}) on ``ctx`` without timing."""
    for opt, value in chosen.items():
        if opt not in CANDIDATES:
            raise ValueError("unknown tunable %r" % opt)
        ctx.set_option(opt, int(value))
