"""Dense FP64 PyTorch restatement of the plain FLIP step, for gradients.

Test infrastructure only: the step is written as plain differentiable torch
operations (every particle against every node for P2G; explicit corner
gathers for the trilinear samples), so ``torch.autograd`` gives the exact
FP64 gradient that the library's hand-written adjoints must reproduce.

It follows the same reference behaviour as the C oracle's FP64 step
(oracle/gfsx_oracle.c, ``gfo64_flip_step``) and is pinned against it:
  * Wyvill kernel in the reference's expanded form
    1 - 4/9 t^3 + 17/9 t^2 - 22/9 t, t = r^2/R^2 < 1 (scalarfield.cpp:40-46,
    569-571);
  * P2G with R = dx onto the u, v, w faces at their offsets, normalised where
    W > 0 (fluidsimulation.cpp:2526-2570, scalarfield.cpp:90-106);
  * body force on faces bordering a cell that holds a particle
    (fluidsimulation.cpp:2010-2017, 2766-2805, fluidmaterialgrid.cpp:119-143);
  * MAC trilinear samples, zero outside [0, n*dx) (macvelocityfield.cpp:
    452-575), PIC/FLIP (fluidsimulation.cpp:3115-3127) and RK4
    (particleadvector.cpp:1045-1054).
Floors (cell indices, the fluid mask) are constants for autograd, which is
the almost-everywhere derivative.
"""
import torch

F64 = torch.float64


def wyvill(t):
    return torch.where(t < 1.0, 1.0 - (4.0 / 9.0) * t ** 3 + (17.0 / 9.0) * t ** 2
                       - (22.0 / 9.0) * t, torch.zeros_like(t))


def splat(x, val, node_dims, offset, dx, radius):
    """S = sum_p K val_p and W = sum_p K on the nodes offset + (i,j,k) dx,
    shaped (nk, nj, ni); every particle against every node."""
    ni, nj, nk = node_dims
    k, j, i = torch.meshgrid(torch.arange(nk, dtype=F64), torch.arange(nj, dtype=F64),
                             torch.arange(ni, dtype=F64), indexing="ij")
    nodes = torch.stack([i, j, k], -1).reshape(-1, 3) * dx
    p = x - torch.as_tensor(offset, dtype=F64)
    d = nodes[None, :, :] - p[:, None, :]
    K = wyvill((d * d).sum(-1) / (radius * radius))
    W = K.sum(0).reshape(nk, nj, ni)
    S = W if val is None else (K * val[:, None]).sum(0).reshape(nk, nj, ni)
    return S, W


def sample(f, q, dims, dx):
    """MAC trilinear velocity of the face arrays f = (u, v, w) at q (n, 3)."""
    ni, nj, nk = dims
    inside = ((q >= 0).all(1) & (q[:, 0] < ni * dx) & (q[:, 1] < nj * dx) &
              (q[:, 2] < nk * dx))
    out = []
    for c in range(3):
        arr = f[c]
        d_, h_, w_ = arr.shape
        shift = torch.full((3,), 0.5 * dx, dtype=F64)
        shift[c] = 0.0
        s = (q - shift) / dx
        base = torch.floor(s).detach()
        t = s - base
        b = base.long()
        val = torch.zeros(q.shape[0], dtype=F64)
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    ii, jj, kk = b[:, 0] + di, b[:, 1] + dj, b[:, 2] + dk
                    ok = ((ii >= 0) & (ii < w_) & (jj >= 0) & (jj < h_) &
                          (kk >= 0) & (kk < d_))
                    v = arr[kk.clamp(0, d_ - 1), jj.clamp(0, h_ - 1), ii.clamp(0, w_ - 1)]
                    v = torch.where(ok, v, torch.zeros_like(v))
                    wx = t[:, 0] if di else 1.0 - t[:, 0]
                    wy = t[:, 1] if dj else 1.0 - t[:, 1]
                    wz = t[:, 2] if dk else 1.0 - t[:, 2]
                    val = val + wx * wy * wz * v
        out.append(torch.where(inside, val, torch.zeros_like(val)))
    return torch.stack(out, 1)


def border_masks(x, dims, dx):
    """Per component, faces that border a cell holding a particle."""
    ni, nj, nk = dims
    c = torch.floor(x.detach() / dx)
    ok = ((c >= 0).all(1) & (c[:, 0] < ni) & (c[:, 1] < nj) & (c[:, 2] < nk))
    c = c[ok].long()
    fluid = torch.zeros((nk, nj, ni), dtype=torch.bool)
    fluid[c[:, 2], c[:, 1], c[:, 0]] = True
    masks = []
    for axis in (2, 1, 0):   # u pads along i (dim 2), v along j, w along k
        pad = [0, 0, 0, 0, 0, 0]
        pad[2 * (2 - axis)] = 1
        pad[2 * (2 - axis) + 1] = 1
        fp = torch.nn.functional.pad(fluid.to(torch.uint8), pad).bool()
        n = fp.shape[axis]
        lo = fp.narrow(axis, 0, n - 1)
        hi = fp.narrow(axis, 1, n - 1)
        masks.append(lo | hi)
    return masks


def step(x, v, dims, dx, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05):
    """(x, v) (n, 3) float64 -> (x', v'), differentiable."""
    ni, nj, nk = dims
    masks = border_masks(x, dims, dx)
    grids, olds = [], []
    for c in range(3):
        off = [0.5 * dx] * 3
        off[c] = 0.0
        nd = (ni + (c == 0), nj + (c == 1), nk + (c == 2))
        S, W = splat(x, v[:, c], nd, off, dx, dx)
        n = torch.where(W > 0, S / torch.where(W > 0, W, torch.ones_like(W)), torch.zeros_like(S))
        bf = float(torch.tensor(body_force[c], dtype=torch.float32))
        g = n + masks[c].to(F64) * (bf * dt) if bf != 0.0 else n
        grids.append(g)
        olds.append(n)
    vnew = sample(grids, x, dims, dx)
    vold = sample(olds, x, dims, dx)
    v1 = pic_ratio * vnew + (1.0 - pic_ratio) * (v + vnew - vold)
    h = 0.5 * dt
    k1 = vnew
    k2 = sample(grids, x + h * k1, dims, dx)
    k3 = sample(grids, x + h * k2, dims, dx)
    k4 = sample(grids, x + dt * k3, dims, dx)
    x1 = x + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    return x1, v1


def density(x, dims, dx):
    """Splat weight on the cell centres with radius dx, shaped (nk, nj, ni)."""
    h = 0.5 * dx
    return splat(x, None, dims, (h, h, h), dx, dx)[1]
