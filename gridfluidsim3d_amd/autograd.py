"""Differentiable form of the FLIP transfer + advection step.

``FlipStepFunction`` wraps one plain step (trilinear stencil, particles kept
in input order) as a ``torch.autograd.Function``: forward is
``gfsx_flip_step``, backward is ``gfsx_flip_step_backward``, both called
through the library's ctypes layer with device pointers (the pattern of
pyfluid's ``init_lib_func`` / ``execute_lib_func``, pybindings.py:10-36).
The tensors autograd saves for each step are that step's inputs (x_t, v_t)
and, by default, its tape (``gfsx_flip_step_record``: the brick sort of x_t
and the step's grids), so that the backward pass recomputes only the RK4
stages; with ``tape=False`` it recomputes the sort and the grids as well
(24 bytes per particle per step saved instead of about 40 plus 36 per face).

``SplatWeightFunction`` is the weight field of the point splat
(``gfsx_p2g_device`` with no values, CLScalarField::addPoints) with its
adjoint ``gfsx_p2g_backward_device``; on cell centres with radius dx it is
the particle density used by the rollout loss.

Particles are (3, n) float32 CUDA tensors (rows x, y, z): the structure of
arrays the kernels read.
"""
import torch

from .gfsx import mac_shapes


def _zeros_if_none(g, like):
    return torch.zeros_like(like) if g is None else g.contiguous()


class FlipStepFunction(torch.autograd.Function):
    """(x, v) -> (x', v') of one plain step; see the module docstring."""

    @staticmethod
    def forward(fctx, x, v, gctx, dims, dx, dt, body_force, pic_ratio, tape=True):
        x = x.contiguous()
        v = v.contiguous()
        fctx.tape = None
        if tape and any(fctx.needs_input_grad[:2]):
            x_out, v_out, fctx.tape = gctx.flip_step_record(x, v, dims, dx, dt, body_force,
                                                            pic_ratio)
        else:
            x_out, v_out, _ = gctx.flip_step_soa(x, v, dims, dx, dt, body_force, pic_ratio)
        fctx.save_for_backward(x, v)
        fctx.gctx = gctx
        fctx.args = (dims, dx, dt, body_force, pic_ratio)
        return x_out, v_out

    @staticmethod
    def backward(fctx, gx_out, gv_out):
        x, v = fctx.saved_tensors
        gx_out = _zeros_if_none(gx_out, x)
        gv_out = _zeros_if_none(gv_out, v)
        gx, gv = fctx.gctx.flip_step_backward(x, v, gx_out, gv_out, *fctx.args,
                                              tape=fctx.tape)
        fctx.tape = None
        return gx, gv, None, None, None, None, None, None, None


class SplatWeightFunction(torch.autograd.Function):
    """x -> W, W[node] = sum_p K(|node - x_p|^2) on the node grid
    (ni, nj, nk) at offset + (i, j, k) * dx, returned shaped (nk, nj, ni)."""

    @staticmethod
    def forward(fctx, x, gctx, node_dims, dx, radius, offset):
        pos = x.t().contiguous()
        ni, nj, nk = node_dims
        S = torch.zeros((nk, nj, ni), dtype=torch.float32, device=x.device)
        W = torch.zeros_like(S)
        gctx.p2g_device(pos, None, radius, offset, dx, node_dims, S, W)
        fctx.save_for_backward(pos)
        fctx.gctx = gctx
        fctx.args = (radius, offset, dx, node_dims)
        return W

    @staticmethod
    def backward(fctx, gW):
        (pos,) = fctx.saved_tensors
        radius, offset, dx, node_dims = fctx.args
        gpos, _ = fctx.gctx.p2g_backward_device(pos, None, radius, offset, dx, node_dims,
                                                gS=None, gW=gW.contiguous())
        return gpos.t().contiguous(), None, None, None, None, None


def flip_step(gctx, x, v, dims, dx, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
              tape=True):
    """One differentiable plain step (``tape``: keep the step's sort and
    grids for the backward pass instead of recomputing them)."""
    return FlipStepFunction.apply(x, v, gctx, tuple(dims), float(dx), float(dt),
                                  tuple(float(b) for b in body_force), float(pic_ratio),
                                  bool(tape))


def rollout(gctx, x0, v0, dims, dx, steps, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
            tape=True):
    """``steps`` differentiable plain steps from (x0, v0); returns (x, v).
    Autograd keeps one (x_t, v_t) checkpoint per step (24 bytes per
    particle), plus with ``tape`` the step's sort and grids (16 bytes per
    particle, 36 per face)."""
    x, v = x0, v0
    for _ in range(steps):
        x, v = flip_step(gctx, x, v, dims, dx, dt, body_force, pic_ratio, tape)
    return x, v


def density(gctx, x, dims, dx, radius=None):
    """Particle density on the cell centres of an (ni, nj, nk) grid: the
    splat weight field with radius dx (or ``radius``), shaped (nk, nj, ni)."""
    h = 0.5 * dx
    return SplatWeightFunction.apply(x, gctx, tuple(dims), float(dx),
                                     float(dx if radius is None else radius), (h, h, h))


def density_loss(gctx, x, target, dims, dx):
    """L = sum over cells of (density(x) - target)^2."""
    d = density(gctx, x, dims, dx)
    return ((d - target) ** 2).sum()


__all__ = ["FlipStepFunction", "SplatWeightFunction", "flip_step", "rollout", "density",
           "density_loss", "mac_shapes"]
