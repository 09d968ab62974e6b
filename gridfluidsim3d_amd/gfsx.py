"""Python host API over include/gfsx.h (ctypes).

``Context`` mirrors the two reference accelerators this library replaces:

* ``p2g`` / ``p2g_device``         -- CLScalarField::addPointValues / addPoints
                                      (clscalarfield.cpp:68-301)
* ``g2p`` / ``g2p_device``         -- ParticleAdvector::tricubicInterpolate
                                      (particleadvector.cpp:401-465)
* ``advect`` / ``advect_device``   -- ParticleAdvector::advectParticlesRK1..4
                                      (particleadvector.cpp:209-399)
* ``flip_step``                    -- the fused plain FLIP step on device
                                      tensors (P2G, normalisation, body force,
                                      PIC/FLIP, RK4)

Host entry points take numpy arrays; *_device entry points and ``flip_step``
take torch tensors on the context's GPU and run on torch's current stream.
"""
import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_void_p

import numpy as np

from . import pybindings as pb

TRILINEAR = 0
TRICUBIC = 1
ORDER_SORTED = 0
ORDER_KEEP = 1
EPILOGUE_NONE = 0
EPILOGUE_FLUIDSIM = 1
# bytes of the static device scratch of gfsx_flip_step_noalloc (GFSX_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20

_fp = POINTER(c_float)
_u32p = POINTER(ctypes.c_uint32)


class Scene(ctypes.Structure):
    """ctypes image of ``gfsx_scene`` (include/gfsx.h)."""
    _fields_ = [
        ("ni", c_int), ("nj", c_int), ("nk", c_int),
        ("dx", c_double),
        ("n", c_int64),
        ("x_in", c_void_p), ("y_in", c_void_p), ("z_in", c_void_p),
        ("vx_in", c_void_p), ("vy_in", c_void_p), ("vz_in", c_void_p),
        ("x_out", c_void_p), ("y_out", c_void_p), ("z_out", c_void_p),
        ("vx_out", c_void_p), ("vy_out", c_void_p), ("vz_out", c_void_p),
        ("id_in", c_void_p), ("id_out", c_void_p),
        ("u", c_void_p), ("v", c_void_p), ("w", c_void_p),
        ("u_old", c_void_p), ("v_old", c_void_p), ("w_old", c_void_p),
        ("wu", c_void_p), ("wv", c_void_p), ("ww", c_void_p),
        ("solid", c_void_p),
    ]


class StepParams(ctypes.Structure):
    """ctypes image of ``gfsx_step_params`` (include/gfsx.h)."""
    _fields_ = [
        ("dt", c_double),
        ("body_force", c_float * 3),
        ("pic_ratio", c_double),
        ("stencil", c_int),
        ("order", c_int),
        ("epilogue", c_int),
        ("extrapolation_layers", c_int),
    ]


class SceneGrad(ctypes.Structure):
    """ctypes image of ``gfsx_scene_grad`` (include/gfsx.h)."""
    _fields_ = [
        ("gx_out", c_void_p), ("gy_out", c_void_p), ("gz_out", c_void_p),
        ("gvx_out", c_void_p), ("gvy_out", c_void_p), ("gvz_out", c_void_p),
        ("gx_in", c_void_p), ("gy_in", c_void_p), ("gz_in", c_void_p),
        ("gvx_in", c_void_p), ("gvy_in", c_void_p), ("gvz_in", c_void_p),
    ]


class StepTape(ctypes.Structure):
    """ctypes image of ``gfsx_step_tape`` (include/gfsx.h)."""
    _fields_ = [
        ("perm", c_void_p), ("start", c_void_p),
        ("xs", c_void_p), ("ys", c_void_p), ("zs", c_void_p),
        ("n", c_int64), ("starts", c_int64),
        ("gbx", c_int), ("gby", c_int),
    ]


class FlipTape:
    """What ``Context.flip_step_record`` keeps of one step for its backward:
    the brick sort of the input positions (``perm``, ``start``, ``xs``) and
    the step's grids (``grids``: u, v, w after the body force; the normalised
    P2G result; the P2G weights), all device tensors."""

    def __init__(self, n, dims, starts, device):
        import torch
        i32, f32 = torch.int32, torch.float32
        self.n = n
        self.dims = tuple(dims)
        self.perm = torch.empty(n, dtype=i32, device=device)
        self.start = torch.empty(starts, dtype=i32, device=device)
        self.xs = torch.empty((3, n), dtype=f32, device=device)
        self.grids = [[torch.empty(sh, dtype=f32, device=device) for sh in mac_shapes(dims)]
                      for _ in range(3)]
        self.gbx = self.gby = 0

    def c_tape(self):
        import torch
        t = StepTape()
        t.perm = _dptr(self.perm, "perm", torch.int32, self.n)
        t.start = _dptr(self.start, "start", torch.int32)
        t.xs, t.ys, t.zs = [_dptr(r, "xs", torch.float32, self.n) for r in self.xs]
        t.n = self.n
        t.starts = int(self.start.numel())
        t.gbx, t.gby = self.gbx, self.gby
        return t

    def set_grids(self, sc):
        """Point a ``Scene``'s grid fields at the tape's grids."""
        names = (("u", "v", "w"), ("u_old", "v_old", "w_old"), ("wu", "wv", "ww"))
        for nms, ts in zip(names, self.grids):
            for nm, t in zip(nms, ts):
                setattr(sc, nm, _dptr(t, nm))


def mac_shapes(dims):
    """(k, j, i) array shapes of the u, v, w faces of an ni*nj*nk grid
    (macvelocityfield.cpp:37-45)."""
    ni, nj, nk = dims
    return [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]


def _np_f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _dptr(t, name, dtype=None, numel=None):
    """Device pointer of a contiguous CUDA tensor (None passes NULL)."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % name)
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if dtype is not None and t.dtype != dtype:
        raise TypeError("%s must have dtype %s" % (name, dtype))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s has %d elements, expected %d" % (name, t.numel(), numel))
    return t.data_ptr()


class Context:
    """A device, a stream and the device workspace of libgfsx
    (``gfsx_create``)."""

    def __init__(self, device=0):
        f = pb.lib_func("gfsx_create", [c_int], c_void_p)
        self._h = pb.execute_lib_func(f, [device])
        self.device = device

    def __del__(self):
        try:
            self.close()
        except Exception:
            # interpreter shutdown may already have torn down ctypes
            pass

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            f = pb.lib_func("gfsx_destroy", [c_void_p])
            pb.execute_lib_func(f, [h])

    def __call__(self):
        if not self._h:
            raise RuntimeError("gfsx context is closed")
        return self._h

    @property
    def device_info(self):
        f = pb.lib_func("gfsx_get_device_info", [c_void_p], ctypes.c_char_p)
        return pb.execute_lib_func(f, [self()]).decode()

    def synchronize(self):
        f = pb.lib_func("gfsx_synchronize", [c_void_p])
        pb.execute_lib_func(f, [self()])

    def set_option(self, name, value):
        """Set a tunable of the fused step (see gfsx_set_option)."""
        f = pb.lib_func("gfsx_set_option", [c_void_p, ctypes.c_char_p, c_int])
        pb.execute_lib_func(f, [self(), name.encode(), int(value)])

    def get_option(self, name):
        f = pb.lib_func("gfsx_get_option", [c_void_p, ctypes.c_char_p], c_int)
        return pb.execute_lib_func(f, [self(), name.encode()])

    def timing_begin(self):
        """Record per-stage HIP events for subsequent flip_step calls."""
        f = pb.lib_func("gfsx_timing_begin", [c_void_p])
        pb.execute_lib_func(f, [self()])

    def timing_end(self):
        """Stop recording; returns (steps, {sort, p2g, g2p, step} total ms)."""
        ms = (c_double * 4)()
        f = pb.lib_func("gfsx_timing_end", [c_void_p, POINTER(c_double)], c_int)
        steps = pb.execute_lib_func(f, [self(), ms])
        return steps, dict(sort=ms[0], p2g=ms[1], g2p=ms[2], step=ms[3])

    def use_torch_stream(self):
        """Enqueue subsequent work on torch's current stream."""
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        f = pb.lib_func("gfsx_set_stream", [c_void_p, c_void_p])
        pb.execute_lib_func(f, [self(), s])

    # ------------------------------------------------------------ host API

    def p2g(self, pos, val, radius, offset, dx, node_dims, S=None, W=None,
            weight=True):
        """Splat onto a node grid (ni, nj, nk); accumulates into S and W
        (created as zeros when None) and returns them, shaped (nk, nj, ni).
        ``val`` None splats the kernel weight (CLScalarField::addPoints)."""
        pos = _np_f32(pos, (-1, 3))
        n = pos.shape[0]
        val = None if val is None else _np_f32(val, (n,))
        ni, nj, nk = node_dims
        S = np.zeros((nk, nj, ni), np.float32) if S is None else S
        if weight and W is None:
            W = np.zeros((nk, nj, ni), np.float32)
        for a, nm in ((S, "S"), (W, "W")):
            if a is not None and (a.dtype != np.float32 or not a.flags.c_contiguous
                                  or a.size != ni * nj * nk):
                raise ValueError("%s must be a C-contiguous float32 array of %d elements"
                                 % (nm, ni * nj * nk))
        off = _np_f32(offset, (3,))
        f = pb.lib_func("gfsx_p2g", [c_void_p, _fp, _fp, c_int64, c_double, _fp,
                                     c_double, c_int, c_int, c_int, _fp, _fp])
        pb.execute_lib_func(f, [self(), _ptr(pos), _ptr(val), n, float(radius),
                                _ptr(off), float(dx), ni, nj, nk, _ptr(S), _ptr(W)])
        return S, W

    def _mac_host(self, u, v, w, dims):
        shapes = mac_shapes(dims)
        return [_np_f32(a, s) for a, s in zip((u, v, w), shapes)]

    def g2p(self, pos, u, v, w, dims, dx, stencil=TRILINEAR):
        pos = _np_f32(pos, (-1, 3))
        u, v, w = self._mac_host(u, v, w, dims)
        out = np.empty_like(pos)
        f = pb.lib_func("gfsx_g2p", [c_void_p, _fp, c_int64, _fp, _fp, _fp, c_int,
                                     c_int, c_int, c_double, c_int, _fp])
        pb.execute_lib_func(f, [self(), _ptr(pos), pos.shape[0], _ptr(u), _ptr(v),
                                _ptr(w), *dims, float(dx), int(stencil), _ptr(out)])
        return out

    def advect(self, pos, u, v, w, dims, dx, dt, order=4, stencil=TRILINEAR):
        pos = _np_f32(pos, (-1, 3))
        u, v, w = self._mac_host(u, v, w, dims)
        out = np.empty_like(pos)
        f = pb.lib_func("gfsx_advect", [c_void_p, c_int, _fp, c_int64, _fp, _fp, _fp,
                                        c_int, c_int, c_int, c_double, c_double,
                                        c_int, _fp])
        pb.execute_lib_func(f, [self(), int(order), _ptr(pos), pos.shape[0],
                                _ptr(u), _ptr(v), _ptr(w), *dims, float(dx),
                                float(dt), int(stencil), _ptr(out)])
        return out

    # ---------------------------------------------------------- device API

    def p2g_device(self, pos, val, radius, offset, dx, node_dims, S, W=None):
        import torch
        ni, nj, nk = node_dims
        n = pos.shape[0]
        off = _np_f32(offset, (3,))
        self.use_torch_stream()
        f = pb.lib_func("gfsx_p2g_device", [c_void_p, c_void_p, c_void_p, c_int64,
                                            c_double, _fp, c_double, c_int, c_int,
                                            c_int, c_void_p, c_void_p])
        pb.execute_lib_func(f, [self(), _dptr(pos, "pos", torch.float32, 3 * n),
                                _dptr(val, "val", torch.float32, n), n,
                                float(radius), _ptr(off), float(dx), ni, nj, nk,
                                _dptr(S, "S", torch.float32, ni * nj * nk),
                                _dptr(W, "W", torch.float32, ni * nj * nk)])
        return S, W

    def p2g_backward_device(self, pos, val, radius, offset, dx, node_dims, gS=None,
                            gW=None, gpos=None, gval=None):
        """Adjoint of ``p2g_device`` (``gfsx_p2g_backward_device``): returns
        (dL/dpos shaped like pos, dL/dval or None) for field adjoints gS, gW
        on the node grid (either may be None)."""
        import torch
        ni, nj, nk = node_dims
        n = pos.shape[0]
        off = _np_f32(offset, (3,))
        gpos = torch.empty_like(pos) if gpos is None else gpos
        if gval is None and val is not None and gS is not None:
            gval = torch.empty_like(val)
        self.use_torch_stream()
        f = pb.lib_func("gfsx_p2g_backward_device",
                        [c_void_p, c_void_p, c_void_p, c_int64, c_double, _fp, c_double,
                         c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p])
        pb.execute_lib_func(f, [self(), _dptr(pos, "pos", torch.float32, 3 * n),
                                _dptr(val, "val", torch.float32, n), n, float(radius),
                                _ptr(off), float(dx), ni, nj, nk,
                                _dptr(gS, "gS", torch.float32, ni * nj * nk),
                                _dptr(gW, "gW", torch.float32, ni * nj * nk),
                                _dptr(gpos, "gpos", torch.float32, 3 * n),
                                _dptr(gval, "gval", torch.float32, n)])
        return gpos, gval

    def _mac_dev(self, u, v, w, dims):
        import torch
        shapes = mac_shapes(dims)
        return [_dptr(a, nm, torch.float32, int(np.prod(s)))
                for a, nm, s in zip((u, v, w), "uvw", shapes)]

    def g2p_device(self, pos, u, v, w, dims, dx, stencil=TRILINEAR, out=None):
        import torch
        n = pos.shape[0]
        out = torch.empty_like(pos) if out is None else out
        self.use_torch_stream()
        f = pb.lib_func("gfsx_g2p_device", [c_void_p, c_void_p, c_int64, c_void_p,
                                            c_void_p, c_void_p, c_int, c_int, c_int,
                                            c_double, c_int, c_void_p])
        pb.execute_lib_func(f, [self(), _dptr(pos, "pos", torch.float32, 3 * n), n,
                                *self._mac_dev(u, v, w, dims), *dims, float(dx),
                                int(stencil), _dptr(out, "out", torch.float32, 3 * n)])
        return out

    def advect_device(self, pos, u, v, w, dims, dx, dt, order=4,
                      stencil=TRILINEAR, out=None):
        import torch
        n = pos.shape[0]
        out = torch.empty_like(pos) if out is None else out
        self.use_torch_stream()
        f = pb.lib_func("gfsx_advect_device", [c_void_p, c_int, c_void_p, c_int64,
                                               c_void_p, c_void_p, c_void_p, c_int,
                                               c_int, c_int, c_double, c_double,
                                               c_int, c_void_p])
        pb.execute_lib_func(f, [self(), int(order),
                                _dptr(pos, "pos", torch.float32, 3 * n), n,
                                *self._mac_dev(u, v, w, dims), *dims, float(dx),
                                float(dt), int(stencil),
                                _dptr(out, "out", torch.float32, 3 * n)])
        return out

    def apply_weight_device(self, S, W):
        import torch
        self.use_torch_stream()
        f = pb.lib_func("gfsx_apply_weight_device", [c_void_p, c_void_p, c_void_p, c_int64])
        pb.execute_lib_func(f, [self(), _dptr(S, "S", torch.float32),
                                _dptr(W, "W", torch.float32, S.numel()), S.numel()])
        return S

    def cell_sort_device(self, x, y, z, dx, dims):
        """Stable counting sort of particle indices by cell (SoA positions).
        Returns (perm[n] int64, cell_start[ni*nj*nk+1] int64)."""
        import torch
        n = x.numel()
        ni, nj, nk = dims
        perm = torch.empty(n, dtype=torch.int32, device=x.device)
        start = torch.empty(ni * nj * nk + 1, dtype=torch.int32, device=x.device)
        self.use_torch_stream()
        f = pb.lib_func("gfsx_cell_sort_device", [c_void_p, c_void_p, c_void_p,
                                                  c_void_p, c_int64, c_double, c_int,
                                                  c_int, c_int, c_void_p, c_void_p])
        pb.execute_lib_func(f, [self(), _dptr(x, "x", torch.float32),
                                _dptr(y, "y", torch.float32, n),
                                _dptr(z, "z", torch.float32, n), n, float(dx),
                                ni, nj, nk, _dptr(perm, "perm"), _dptr(start, "start")])
        return perm.long() & 0xFFFFFFFF, start.long() & 0xFFFFFFFF

    @staticmethod
    def _scene_of(st):
        """The C scene of a FlipState (its current buffers in, the others out)."""
        import torch
        sc = Scene()
        sc.ni, sc.nj, sc.nk = st.dims
        sc.dx = st.dx
        sc.n = st.n
        f32 = torch.float32
        for nm, t in zip(("x_in", "y_in", "z_in", "vx_in", "vy_in", "vz_in"),
                         st.cur_x + st.cur_v):
            setattr(sc, nm, _dptr(t, nm, f32, st.n))
        for nm, t in zip(("x_out", "y_out", "z_out", "vx_out", "vy_out", "vz_out"),
                         st.nxt_x + st.nxt_v):
            setattr(sc, nm, _dptr(t, nm, f32, st.n))
        if st.ids is not None:
            sc.id_in = _dptr(st.ids[st.cur], "id_in", torch.int32, st.n)
            sc.id_out = _dptr(st.ids[1 - st.cur], "id_out", torch.int32, st.n)
        shapes = mac_shapes(st.dims)
        for nm, t, s in zip(("u", "v", "w"), st.grid, shapes):
            setattr(sc, nm, _dptr(t, nm, f32, int(np.prod(s))))
        if st.grid_old is not None:
            for nm, t, s in zip(("u_old", "v_old", "w_old"), st.grid_old, shapes):
                setattr(sc, nm, _dptr(t, nm, f32, int(np.prod(s))))
        if st.weights is not None:
            for nm, t, s in zip(("wu", "wv", "ww"), st.weights, shapes):
                setattr(sc, nm, _dptr(t, nm, f32, int(np.prod(s))))
        if st.solid is not None:
            sc.solid = _dptr(st.solid, "solid", torch.uint8, int(np.prod(st.dims)))
        return sc

    @staticmethod
    def _params(dt, body_force, pic_ratio, stencil, order, epilogue=EPILOGUE_NONE,
                extrapolation_layers=0):
        pr = StepParams()
        pr.dt = float(dt)
        pr.body_force[0], pr.body_force[1], pr.body_force[2] = [float(b) for b in body_force]
        pr.pic_ratio = float(pic_ratio)
        pr.stencil = int(stencil)
        pr.order = int(order)
        pr.epilogue = int(epilogue)
        pr.extrapolation_layers = int(extrapolation_layers)
        return pr

    def flip_step(self, state, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
                  stencil=TRILINEAR, order=ORDER_SORTED, epilogue=EPILOGUE_NONE,
                  extrapolation_layers=0):
        """Advance a ``FlipState`` by one plain step (in its double buffers).
        ``epilogue`` EPILOGUE_FLUIDSIM adds FluidSimulation's face fill and
        velocity extrapolation (``extrapolation_layers``, 0 = 7) around the
        transfer, with the state's solid cells."""
        st = state
        sc = self._scene_of(st)
        pr = self._params(dt, body_force, pic_ratio, stencil, order, epilogue,
                          extrapolation_layers)
        self.use_torch_stream()
        f = pb.lib_func("gfsx_flip_step", [c_void_p, POINTER(Scene), POINTER(StepParams)])
        pb.execute_lib_func(f, [self(), ctypes.byref(sc), ctypes.byref(pr)])
        st.cur = 1 - st.cur

    def flip_step_noalloc(self, state, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
                          stencil=TRILINEAR, order=ORDER_SORTED):
        """``flip_step`` through ``gfsx_flip_step_noalloc``: the same step with
        every workspace buffer carved from the library's static device
        scratch (``NOALLOC_SCRATCH_BYTES``); no device allocation.  One scene,
        no epilogue; a step too large for the scratch raises."""
        st = state
        sc = self._scene_of(st)
        pr = self._params(dt, body_force, pic_ratio, stencil, order)
        self.use_torch_stream()
        f = pb.lib_func("gfsx_flip_step_noalloc",
                        [c_void_p, POINTER(Scene), POINTER(StepParams)])
        pb.execute_lib_func(f, [self(), ctypes.byref(sc), ctypes.byref(pr)])
        st.cur = 1 - st.cur

    def remove_particles(self, state, max_per_cell=100, seed=0):
        """Drop the state's particles in solid cells and keep at most
        ``max_per_cell`` per cell (``gfsx_remove_particles``), advancing its
        buffers; returns the number kept (the state shrinks to it)."""
        import torch
        st = state
        sc = self._scene_of(st)
        f = pb.lib_func("gfsx_remove_particles",
                        [c_void_p, POINTER(Scene), c_int, ctypes.c_uint64], c_int64)
        self.use_torch_stream()
        kept = int(pb.execute_lib_func(f, [self(), ctypes.byref(sc), int(max_per_cell),
                                           int(seed) & 0xFFFFFFFFFFFFFFFF]))
        st.cur = 1 - st.cur
        st.n = kept
        for b in range(2):
            st.x[b] = [t[:kept] for t in st.x[b]]
            st.v[b] = [t[:kept] for t in st.v[b]]
            if st.ids is not None:
                st.ids[b] = st.ids[b][:kept]
        torch.cuda.current_stream().synchronize()
        return kept

    @staticmethod
    def _soa_scene(x, v, dims, dx, x_out=None, v_out=None, grids=None):
        """A C scene over (3, n) SoA tensors: x, v in; x_out, v_out, grids
        (u, v, w) out where given."""
        import torch
        sc = Scene()
        sc.ni, sc.nj, sc.nk = dims
        sc.dx = float(dx)
        n = int(x.shape[1])
        sc.n = n
        f32 = torch.float32
        for t, nm in ((x, "x"), (v, "v"), (x_out, "x_out"), (v_out, "v_out")):
            if t is not None and (t.dim() != 2 or t.shape[0] != 3 or t.shape[1] != n):
                raise ValueError("%s must have shape (3, %d)" % (nm, n))
        names_in = ("x_in", "y_in", "z_in", "vx_in", "vy_in", "vz_in")
        for nm, t in zip(names_in, [x[0], x[1], x[2], v[0], v[1], v[2]]):
            setattr(sc, nm, _dptr(t, nm, f32, n))
        if x_out is not None:
            names_out = ("x_out", "y_out", "z_out", "vx_out", "vy_out", "vz_out")
            for nm, t in zip(names_out, [x_out[0], x_out[1], x_out[2],
                                         v_out[0], v_out[1], v_out[2]]):
                setattr(sc, nm, _dptr(t, nm, f32, n))
        if grids is not None:
            for nm, t, sh in zip(("u", "v", "w"), grids, mac_shapes(dims)):
                setattr(sc, nm, _dptr(t, nm, f32, int(np.prod(sh))))
        return sc

    def flip_step_soa(self, x, v, dims, dx, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
                      x_out=None, v_out=None, grids=None):
        """One plain step (trilinear, output in input order) of particles held
        as (3, n) float32 tensors; returns (x_out, v_out, grids)."""
        import torch
        x_out = torch.empty_like(x) if x_out is None else x_out
        v_out = torch.empty_like(v) if v_out is None else v_out
        if grids is None:
            grids = [torch.empty(sh, dtype=torch.float32, device=x.device)
                     for sh in mac_shapes(dims)]
        sc = self._soa_scene(x, v, dims, dx, x_out, v_out, grids)
        pr = self._params(dt, body_force, pic_ratio, TRILINEAR, ORDER_KEEP)
        self.use_torch_stream()
        f = pb.lib_func("gfsx_flip_step", [c_void_p, POINTER(Scene), POINTER(StepParams)])
        pb.execute_lib_func(f, [self(), ctypes.byref(sc), ctypes.byref(pr)])
        return x_out, v_out, grids

    def flip_step_record(self, x, v, dims, dx, dt, body_force=(0.0, -9.8, 0.0),
                         pic_ratio=0.05):
        """``flip_step_soa`` that also keeps what its backward needs
        (``gfsx_flip_step_record``); returns (x_out, v_out, tape), tape a
        ``FlipTape`` to pass to ``flip_step_backward``."""
        import torch
        n = int(x.shape[1])
        x_out = torch.empty_like(x)
        v_out = torch.empty_like(v)
        f = pb.lib_func("gfsx_step_tape_starts", [c_int, c_int, c_int], c_int64)
        starts = int(pb.execute_lib_func(f, [int(d) for d in dims]))
        tape = FlipTape(n, dims, starts, x.device)
        sc = self._soa_scene(x, v, dims, dx, x_out, v_out)
        tape.set_grids(sc)
        t = tape.c_tape()
        pr = self._params(dt, body_force, pic_ratio, TRILINEAR, ORDER_KEEP)
        self.use_torch_stream()
        f = pb.lib_func("gfsx_flip_step_record",
                        [c_void_p, POINTER(Scene), POINTER(StepParams), POINTER(StepTape)])
        pb.execute_lib_func(f, [self(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(t)])
        tape.gbx, tape.gby = t.gbx, t.gby
        return x_out, v_out, tape

    def flip_step_backward(self, x, v, gx_out, gv_out, dims, dx, dt,
                           body_force=(0.0, -9.8, 0.0), pic_ratio=0.05, gx=None, gv=None,
                           tape=None):
        """Adjoint of ``flip_step_soa`` (``gfsx_flip_step_backward``): from the
        step's inputs x, v and the adjoints of its outputs, all (3, n)
        float32 tensors, returns (dL/dx, dL/dv).  With the ``tape`` of
        ``flip_step_record`` on the same inputs it runs
        ``gfsx_flip_step_backward_tape`` (same bits, no recompute of the
        forward's sort and grids)."""
        import torch
        gx = torch.empty_like(x) if gx is None else gx
        gv = torch.empty_like(v) if gv is None else gv
        n = int(x.shape[1])
        for t, nm in ((gx_out, "gx_out"), (gv_out, "gv_out"), (gx, "gx"), (gv, "gv")):
            if t.dim() != 2 or t.shape[0] != 3 or t.shape[1] != n:
                raise ValueError("%s must have shape (3, %d)" % (nm, n))
        sc = self._soa_scene(x, v, dims, dx)
        pr = self._params(dt, body_force, pic_ratio, TRILINEAR, ORDER_KEEP)
        g = SceneGrad()
        f32 = torch.float32
        for nm, t in zip(("gx_out", "gy_out", "gz_out"), gx_out):
            setattr(g, nm, _dptr(t, nm, f32, n))
        for nm, t in zip(("gvx_out", "gvy_out", "gvz_out"), gv_out):
            setattr(g, nm, _dptr(t, nm, f32, n))
        for nm, t in zip(("gx_in", "gy_in", "gz_in"), gx):
            setattr(g, nm, _dptr(t, nm, f32, n))
        for nm, t in zip(("gvx_in", "gvy_in", "gvz_in"), gv):
            setattr(g, nm, _dptr(t, nm, f32, n))
        self.use_torch_stream()
        if tape is not None:
            if tape.n != n or tape.dims != tuple(dims):
                raise ValueError("the tape was recorded for another step")
            tape.set_grids(sc)
            t = tape.c_tape()
            f = pb.lib_func("gfsx_flip_step_backward_tape",
                            [c_void_p, POINTER(Scene), POINTER(StepParams), POINTER(SceneGrad),
                             POINTER(StepTape)])
            pb.execute_lib_func(f, [self(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(g),
                                    ctypes.byref(t)])
            return gx, gv
        f = pb.lib_func("gfsx_flip_step_backward",
                        [c_void_p, POINTER(Scene), POINTER(StepParams), POINTER(SceneGrad)])
        pb.execute_lib_func(f, [self(), ctypes.byref(sc), ctypes.byref(pr), ctypes.byref(g)])
        return gx, gv

    def capture_batched(self, states, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
                        stencil=TRILINEAR, order=ORDER_SORTED):
        """Capture the batched step of ``states`` into HIP graphs, one per
        direction of the double buffers, and return a function that advances
        all states by one step by replaying the right graph.  Runs one step
        in each direction first (so that the work of the captured calls is
        only kernels: workspace and scene tables exist by then)."""
        import torch
        for _ in range(2):
            self.flip_step_batched(states, dt, body_force, pic_ratio, stencil, order)
        torch.cuda.synchronize()
        graphs = []
        for _ in range(2):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.flip_step_batched(states, dt, body_force, pic_ratio, stencil, order)
            graphs.append(g)
        torch.cuda.synchronize()
        # the captures flipped every state twice: back where it started, and
        # graphs[0] starts from the current buffers
        for st in states:
            if st.cur != states[0].cur:
                raise RuntimeError("capture_batched: the states are in different buffers")
        first = states[0].cur

        def step():
            graphs[0 if states[0].cur == first else 1].replay()
            for st in states:
                st.cur = 1 - st.cur
        return step

    def flip_step_batched(self, states, dt, body_force=(0.0, -9.8, 0.0), pic_ratio=0.05,
                          stencil=TRILINEAR, order=ORDER_SORTED):
        """Advance several ``FlipState`` objects by one plain step each, in
        one launch per kernel (``gfsx_flip_step_batched``).  ``dt``,
        ``body_force`` and ``pic_ratio`` are one value for all states or one
        per state."""
        n = len(states)

        def per(v, k):
            if k == "body_force":
                return v if (len(v) == n and hasattr(v[0], "__len__")) else [v] * n
            return list(v) if hasattr(v, "__len__") else [v] * n
        dts, bfs, prs = per(dt, "dt"), per(body_force, "body_force"), per(pic_ratio, "ratio")
        scenes = (Scene * n)(*[self._scene_of(st) for st in states])
        params = (StepParams * n)(*[self._params(dts[i], bfs[i], prs[i], stencil, order)
                                    for i in range(n)])
        self.use_torch_stream()
        f = pb.lib_func("gfsx_flip_step_batched",
                        [c_void_p, POINTER(Scene), c_int, POINTER(StepParams)])
        pb.execute_lib_func(f, [self(), scenes, n, params])
        for st in states:
            st.cur = 1 - st.cur


class FlipState:
    """Double-buffered particle state (structure of arrays) and MAC grid of
    one scene, on one GPU."""

    def __init__(self, pos, vel, dims, dx, device="cuda", track_ids=False,
                 keep_old=False, keep_weights=False, solid=None):
        import torch
        pos = torch.as_tensor(pos, dtype=torch.float32)
        vel = torch.as_tensor(vel, dtype=torch.float32)
        self.n = int(pos.shape[0])
        self.dims = tuple(int(d) for d in dims)
        self.dx = float(dx)
        dev = torch.device(device)
        self.x = [[pos[:, a].contiguous().to(dev) for a in range(3)],
                  [torch.empty(self.n, dtype=torch.float32, device=dev) for _ in range(3)]]
        self.v = [[vel[:, a].contiguous().to(dev) for a in range(3)],
                  [torch.empty(self.n, dtype=torch.float32, device=dev) for _ in range(3)]]
        self.cur = 0
        shapes = mac_shapes(self.dims)
        self.grid = [torch.zeros(s, dtype=torch.float32, device=dev) for s in shapes]
        self.grid_old = ([torch.zeros(s, dtype=torch.float32, device=dev) for s in shapes]
                         if keep_old else None)
        self.weights = ([torch.zeros(s, dtype=torch.float32, device=dev) for s in shapes]
                        if keep_weights else None)
        # solid cells (nk, nj, ni) uint8, used by the FluidSimulation epilogue
        self.solid = None if solid is None else torch.as_tensor(
            np.ascontiguousarray(solid, np.uint8)).reshape(-1).to(dev)
        self.ids = None
        if track_ids:
            self.ids = [torch.arange(self.n, dtype=torch.int32, device=dev),
                        torch.empty(self.n, dtype=torch.int32, device=dev)]

    @property
    def cur_x(self):
        return self.x[self.cur]

    @property
    def cur_v(self):
        return self.v[self.cur]

    @property
    def nxt_x(self):
        return self.x[1 - self.cur]

    @property
    def nxt_v(self):
        return self.v[1 - self.cur]

    def positions(self):
        import torch
        return torch.stack(self.cur_x, dim=1)

    def velocities(self):
        import torch
        return torch.stack(self.cur_v, dim=1)

    def particle_ids(self):
        return None if self.ids is None else self.ids[self.cur]
