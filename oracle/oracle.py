"""ctypes bindings to the CPU checkers (test infrastructure only).

* ``liboracle.so``          -- the C restatement of the reference CPU path
                                (gfsx_oracle.c), FP32 (``gfo_*``) and FP64
                                (``gfo64_*``) families.
* ``_ref/libgfsx_ref.so``  -- the reference's own CPU sources compiled with a
                                small harness (ref_harness.cpp); present only
                                where it was built from the reference tree.

Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use this module.  The product (``gridfluidsim3d_amd``) never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgfsx_ref.so")
REF_SRC = "/root/reference/src"

TRILINEAR = 0
TRICUBIC = 1

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double


DROPIN_SO = os.path.join(HERE, "_ref", "libgfsx_dropin_check.so")


def build(with_reference=None):
    """Compile liboracle.so, and -- when the reference tree is present (or
    ``with_reference`` is True) -- _ref/libgfsx_ref.so and the drop-in
    checker _ref/libgfsx_dropin_check.so (which links the built libgfsx.so)."""
    if with_reference is None:
        with_reference = os.path.isdir(REF_SRC)
    targets = ["all"] + (["ref", "dropin"] if with_reference else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mac_shapes(dims):
    ni, nj, nk = dims
    return [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]


class CPUPath:
    """One CPU implementation of the path: the restatement (prefix ``gfo``)
    or the reference build (prefix ``ref``).  Arrays are numpy; positions
    and velocities are (n, 3) float32; grids are (k, j, i) float32."""

    def __init__(self, so_path, prefix):
        self.lib = ctypes.CDLL(so_path)
        self.prefix = prefix
        L = self.lib
        p2g = getattr(L, prefix + "_p2g")
        if prefix == "ref":
            p2g.argtypes = [_f32p, _f32p, _i64, _dbl, _f32p, _dbl, _int, _int,
                            _int, _f32p, _f32p, _int]
        else:
            p2g.argtypes = [_f32p, _f32p, _i64, _dbl, _f32p, _dbl, _int, _int,
                            _int, _f32p, _f32p]
        p2g.restype = None
        self._p2g = p2g
        g2p = getattr(L, prefix + "_g2p")
        g2p.argtypes = [_f32p, _i64, _f32p, _f32p, _f32p, _int, _int, _int,
                        _dbl, _int, _f32p]
        g2p.restype = None
        self._g2p = g2p
        adv = getattr(L, prefix + "_advect")
        adv.argtypes = [_f32p, _i64, _f32p, _f32p, _f32p, _int, _int, _int,
                        _dbl, _dbl, _int, _int, _f32p]
        adv.restype = None
        self._advect = adv
        st = getattr(L, prefix + "_flip_step")
        st.argtypes = [_f32p, _f32p, _i64, _int, _int, _int, _dbl, _dbl,
                       _f32p, _dbl, _int] + [_f32p] * 9 + [_u8p]
        st.restype = None
        self._step = st
        fs = getattr(L, prefix + "_flip_step_fs")
        fs.argtypes = [_f32p, _f32p, _i64, _int, _int, _int, _dbl, _dbl, _f32p, _dbl, _int,
                       _u8p, _int] + [_f32p] * 6 + [_u8p]
        fs.restype = None
        self._step_fs = fs
        if prefix == "ref":
            ex = L.ref_extrapolate
            ex.argtypes = [_f32p, _f32p, _f32p, _int, _int, _int, _dbl, _u8p, _int]
            ex.restype = None
            self._extrapolate = lambda u, v, w, ni, nj, nk, dx, m, layers: ex(
                u, v, w, ni, nj, nk, dx, m, layers)
        if prefix == "gfo":
            ex = L.gfo_extrapolate
            ex.argtypes = [_f32p, _f32p, _f32p, _int, _int, _int, _u8p, _int]
            ex.restype = None
            self._extrapolate = lambda u, v, w, ni, nj, nk, dx, m, layers: ex(
                u, v, w, ni, nj, nk, m, layers)
            rk = L.gfo_removal_key
            rk.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
            rk.restype = ctypes.c_uint64
            self.removal_key = rk
            rp = L.gfo_remove_particles
            rp.argtypes = [_f32p, _i64, _dbl, _int, _int, _int, _u8p, _int, ctypes.c_uint64, _u8p]
            rp.restype = _i64
            self._remove = rp
            fc = L.gfo_fill_component
            fc.argtypes = [_f32p, _f32p, _f32p, _int, _u8p, _int, _int, _int]
            fc.restype = None
            self._fill = fc
            aw = L.gfo_apply_weight
            aw.argtypes = [_f32p, _f32p, _i64]
            aw.restype = None
            self._apply_weight = aw
            cs = L.gfo_cell_sort
            cs.argtypes = [_f32p, _i64, _dbl, _int, _int, _int, _u32p, _u32p]
            cs.restype = _i64
            self._cell_sort = cs
            kw = L.gfo_kernel_weight
            kw.argtypes = [_dbl, _dbl]
            kw.restype = _dbl
            self.kernel_weight = kw

    def p2g(self, pos, val, radius, offset, dx, node_dims, weight=True,
            normalize=False):
        """Splat ``val`` (or the kernel weight when ``val`` is None) onto a
        node grid of ``node_dims`` = (ni, nj, nk).  Returns (S, W) with
        shape (nk, nj, ni); W is None when ``weight`` is False."""
        pos = _f32(pos).reshape(-1, 3)
        n = pos.shape[0]
        val = None if val is None else _f32(val).reshape(n)
        ni, nj, nk = node_dims
        S = np.zeros((nk, nj, ni), np.float32)
        W = np.zeros((nk, nj, ni), np.float32) if weight else None
        off = _f32(offset).reshape(3)
        args = [_ptr(pos, _f32p), _ptr(val, _f32p), n, float(radius),
                _ptr(off, _f32p), float(dx), ni, nj, nk, _ptr(S, _f32p),
                _ptr(W, _f32p)]
        if self.prefix == "ref":
            self._p2g(*args, 1 if normalize else 0)
        else:
            self._p2g(*args)
            if normalize:
                self._apply_weight(_ptr(S, _f32p), _ptr(W, _f32p), S.size)
        return S, W

    def g2p(self, pos, u, v, w, dims, dx, stencil=TRILINEAR):
        pos = _f32(pos).reshape(-1, 3)
        u, v, w = _f32(u), _f32(v), _f32(w)
        out = np.zeros_like(pos)
        self._g2p(_ptr(pos, _f32p), pos.shape[0], _ptr(u, _f32p),
                  _ptr(v, _f32p), _ptr(w, _f32p), *dims, float(dx),
                  int(stencil), _ptr(out, _f32p))
        return out

    def advect(self, pos, u, v, w, dims, dx, dt, order=4, stencil=TRILINEAR):
        pos = _f32(pos).reshape(-1, 3)
        u, v, w = _f32(u), _f32(v), _f32(w)
        out = np.zeros_like(pos)
        self._advect(_ptr(pos, _f32p), pos.shape[0], _ptr(u, _f32p),
                     _ptr(v, _f32p), _ptr(w, _f32p), *dims, float(dx),
                     float(dt), int(order), int(stencil), _ptr(out, _f32p))
        return out

    def flip_step(self, pos, vel, dims, dx, dt, body_force=(0.0, -9.8, 0.0),
                  pic_ratio=0.05, stencil=TRILINEAR):
        """The plain step.  Returns a dict with the new pos/vel and the grids
        (u, v, w after the body force; u_old, v_old, w_old; wu, wv, ww) and
        the fluid-cell mask."""
        pos = _f32(pos).reshape(-1, 3).copy()
        vel = _f32(vel).reshape(-1, 3).copy()
        shapes = mac_shapes(dims)
        g = [np.zeros(s, np.float32) for s in shapes]
        o = [np.zeros(s, np.float32) for s in shapes]
        wt = [np.zeros(s, np.float32) for s in shapes]
        ni, nj, nk = dims
        fluid = np.zeros((nk, nj, ni), np.uint8)
        bf = _f32(body_force)
        self._step(_ptr(pos, _f32p), _ptr(vel, _f32p), pos.shape[0], ni, nj,
                   nk, float(dx), float(dt), _ptr(bf, _f32p), float(pic_ratio),
                   int(stencil), *[_ptr(a, _f32p) for a in g + o + wt],
                   _ptr(fluid, _u8p))
        return dict(pos=pos, vel=vel, u=g[0], v=g[1], w=g[2], u_old=o[0],
                    v_old=o[1], w_old=o[2], wu=wt[0], wv=wt[1], ww=wt[2],
                    fluid=fluid)

    def flip_step_fs(self, pos, vel, dims, dx, dt, body_force=(0.0, -9.8, 0.0),
                     pic_ratio=0.05, stencil=TRILINEAR, solid=None, layers=7):
        """The step with FluidSimulation's face fill and extrapolation.
        Returns pos/vel, the grids (u, v, w extrapolated after the body
        force; u_old.. the extrapolated snapshot) and the cell codes
        (0 air, 1 fluid, 2 solid)."""
        pos = _f32(pos).reshape(-1, 3).copy()
        vel = _f32(vel).reshape(-1, 3).copy()
        shapes = mac_shapes(dims)
        g = [np.zeros(s, np.float32) for s in shapes]
        o = [np.zeros(s, np.float32) for s in shapes]
        ni, nj, nk = dims
        mat = np.zeros((nk, nj, ni), np.uint8)
        sol = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(nk, nj, ni)
        bf = _f32(body_force)
        self._step_fs(_ptr(pos, _f32p), _ptr(vel, _f32p), pos.shape[0], ni, nj, nk, float(dx),
                      float(dt), _ptr(bf, _f32p), float(pic_ratio), int(stencil),
                      _ptr(sol, _u8p), int(layers), *[_ptr(a, _f32p) for a in g + o],
                      _ptr(mat, _u8p))
        return dict(pos=pos, vel=vel, u=g[0], v=g[1], w=g[2], u_old=o[0], v_old=o[1],
                    w_old=o[2], mat=mat)

    def extrapolate(self, u, v, w, dims, dx, mat, layers):
        """MACVelocityField::extrapolateVelocityField on copies of u, v, w."""
        u, v, w = (np.array(a, np.float32) for a in (u, v, w))
        ni, nj, nk = dims
        m = np.ascontiguousarray(mat, np.uint8).reshape(nk, nj, ni)
        self._extrapolate(_ptr(u, _f32p), _ptr(v, _f32p), _ptr(w, _f32p), ni, nj, nk,
                          float(dx), _ptr(m, _u8p), int(layers))
        return u, v, w

    def remove_particles(self, pos, dims, dx, solid=None, max_per_cell=100, seed=0):
        """Keep mask (uint8) of the particle removal (restatement only)."""
        pos = _f32(pos).reshape(-1, 3)
        ni, nj, nk = dims
        keep = np.zeros(pos.shape[0], np.uint8)
        sol = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(nk, nj, ni)
        kept = self._remove(_ptr(pos, _f32p), pos.shape[0], float(dx), ni, nj, nk,
                            _ptr(sol, _u8p), int(max_per_cell), int(seed), _ptr(keep, _u8p))
        assert kept == int(keep.sum())
        return keep

    def fill_component(self, ugrid, W, c, dims, mat):
        """The face fill of component c (restatement only)."""
        ugrid = _f32(ugrid)
        W = _f32(W)
        out = np.zeros_like(ugrid)
        ni, nj, nk = dims
        m = np.ascontiguousarray(mat, np.uint8).reshape(nk, nj, ni)
        self._fill(_ptr(out, _f32p), _ptr(ugrid, _f32p), _ptr(W, _f32p), int(c), _ptr(m, _u8p),
                   ni, nj, nk)
        return out

    def cell_sort(self, pos, dx, dims):
        pos = _f32(pos).reshape(-1, 3)
        ni, nj, nk = dims
        perm = np.zeros(pos.shape[0], np.uint32)
        start = np.zeros(ni * nj * nk + 1, np.uint32)
        m = self._cell_sort(_ptr(pos, _f32p), pos.shape[0], float(dx), ni, nj,
                            nk, _ptr(perm, _u32p), _ptr(start, _u32p))
        return perm[:m], start


class Oracle64:
    """The FP64 twin of the restatement."""

    def __init__(self, so_path=ORACLE_SO):
        L = ctypes.CDLL(so_path)
        L.gfo64_p2g.argtypes = [_f64p, _f64p, _i64, _dbl, _f64p, _dbl, _int,
                                _int, _int, _f64p, _f64p]
        L.gfo64_g2p.argtypes = [_f64p, _i64, _f64p, _f64p, _f64p, _int, _int,
                                _int, _dbl, _int, _f64p]
        L.gfo64_advect.argtypes = [_f64p, _i64, _f64p, _f64p, _f64p, _int,
                                   _int, _int, _dbl, _dbl, _int, _int, _f64p]
        L.gfo64_flip_step.argtypes = [_f64p, _f64p, _i64, _int, _int, _int,
                                      _dbl, _dbl, _f64p, _dbl, _int] + [_f64p] * 9
        for f in (L.gfo64_p2g, L.gfo64_g2p, L.gfo64_advect, L.gfo64_flip_step):
            f.restype = None
        self.lib = L

    def p2g(self, pos, val, radius, offset, dx, node_dims, normalize=False):
        pos = _f64(pos).reshape(-1, 3)
        n = pos.shape[0]
        val = None if val is None else _f64(val).reshape(n)
        ni, nj, nk = node_dims
        S = np.zeros((nk, nj, ni))
        W = np.zeros((nk, nj, ni))
        off = _f64(offset).reshape(3)
        self.lib.gfo64_p2g(_ptr(pos, _f64p), _ptr(val, _f64p), n,
                           float(radius), _ptr(off, _f64p), float(dx), ni, nj,
                           nk, _ptr(S, _f64p), _ptr(W, _f64p))
        if normalize:
            S = np.where(W > 0, S / np.where(W > 0, W, 1.0), S)
        return S, W

    def g2p(self, pos, u, v, w, dims, dx, stencil=TRILINEAR):
        pos = _f64(pos).reshape(-1, 3)
        out = np.zeros_like(pos)
        u, v, w = _f64(u), _f64(v), _f64(w)
        self.lib.gfo64_g2p(_ptr(pos, _f64p), pos.shape[0], _ptr(u, _f64p),
                           _ptr(v, _f64p), _ptr(w, _f64p), *dims, float(dx),
                           int(stencil), _ptr(out, _f64p))
        return out

    def advect(self, pos, u, v, w, dims, dx, dt, order=4, stencil=TRILINEAR):
        pos = _f64(pos).reshape(-1, 3)
        out = np.zeros_like(pos)
        u, v, w = _f64(u), _f64(v), _f64(w)
        self.lib.gfo64_advect(_ptr(pos, _f64p), pos.shape[0], _ptr(u, _f64p),
                              _ptr(v, _f64p), _ptr(w, _f64p), *dims,
                              float(dx), float(dt), int(order), int(stencil),
                              _ptr(out, _f64p))
        return out

    def flip_step(self, pos, vel, dims, dx, dt, body_force=(0.0, -9.8, 0.0),
                  pic_ratio=0.05, stencil=TRILINEAR):
        pos = _f64(pos).reshape(-1, 3).copy()
        vel = _f64(vel).reshape(-1, 3).copy()
        shapes = mac_shapes(dims)
        g = [np.zeros(s) for s in shapes]
        o = [np.zeros(s) for s in shapes]
        wt = [np.zeros(s) for s in shapes]
        bf = _f64(np.asarray(body_force, np.float32))  # FP32 inputs, promoted
        ni, nj, nk = dims
        self.lib.gfo64_flip_step(_ptr(pos, _f64p), _ptr(vel, _f64p),
                                 pos.shape[0], ni, nj, nk, float(dx),
                                 float(dt), _ptr(bf, _f64p), float(pic_ratio),
                                 int(stencil),
                                 *[_ptr(a, _f64p) for a in g + o + wt])
        return dict(pos=pos, vel=vel, u=g[0], v=g[1], w=g[2], u_old=o[0],
                    v_old=o[1], w_old=o[2], wu=wt[0], wv=wt[1], ww=wt[2])


def load_oracle():
    if not os.path.exists(ORACLE_SO):
        build(with_reference=False)
    return CPUPath(ORACLE_SO, "gfo")


def load_reference():
    """The reference build, or None where it was never built."""
    if not os.path.exists(REF_SO):
        return None
    return CPUPath(REF_SO, "ref")
