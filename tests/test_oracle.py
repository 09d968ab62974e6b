"""Pins the CPU oracle (oracle/gfsx_oracle.c) before anything is checked
against it:

* bit-for-bit against golden vectors written by the reference's own CPU
  sources (tests/golden/make_golden.py), and against the reference build on
  fresh inputs where oracle/_ref is built;
* against closed forms the mathematics fixes: kernel values and smoothness,
  single-particle splats, brute-force P2G, exactness of trilinear/tricubic
  sampling on affine/quadratic fields, RK steps on constant and linear
  fields, the PIC/FLIP blend on uniform flow, a stable sort.
"""
import math

import numpy as np
import pytest

from conftest import load_golden

TRILINEAR, TRICUBIC = 0, 1


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("case", ["p2g_r1", "p2g_r17"])
def test_p2g_matches_reference_vectors(orc, case):
    g = load_golden(case)
    nd = tuple(int(v) for v in g["node_dims"])
    S, W = orc.p2g(g["pos"], g["val"], float(g["radius"]), g["offset"], float(g["dx"]), nd)
    assert np.array_equal(S, g["S"])
    assert np.array_equal(W, g["W"])
    Sn, _ = orc.p2g(g["pos"], g["val"], float(g["radius"]), g["offset"], float(g["dx"]), nd,
                    normalize=True)
    assert np.array_equal(Sn, g["S_norm"])


def test_addpoints_matches_reference_vectors(orc):
    g = load_golden("addpoints")
    nd = tuple(int(v) for v in g["node_dims"])
    S, _ = orc.p2g(g["pos"], None, float(g["radius"]), g["offset"], float(g["dx"]), nd,
                   weight=False)
    assert np.array_equal(S, g["S"])


@pytest.mark.parametrize("stencil,case", [(TRILINEAR, "g2p_trilinear"),
                                          (TRICUBIC, "g2p_tricubic")])
def test_g2p_and_rk_match_reference_vectors(orc, stencil, case):
    g = load_golden(case)
    dims = tuple(int(v) for v in g["dims"])
    out = orc.g2p(g["pos"], g["u"], g["v"], g["w"], dims, float(g["dx"]), stencil)
    assert np.array_equal(out, g["out"])
    for order in (1, 2, 3, 4):
        adv = orc.advect(g["pos"], g["u"], g["v"], g["w"], dims, float(g["dx"]),
                         float(g["dt"]), order, stencil)
        assert np.array_equal(adv, g["rk%d" % order]), order


@pytest.mark.parametrize("stencil,case", [(TRILINEAR, "step_trilinear"),
                                          (TRICUBIC, "step_tricubic")])
def test_step_matches_reference_vectors(orc, stencil, case):
    g = load_golden(case)
    dims = tuple(int(v) for v in g["dims"])
    r = orc.flip_step(g["pos"], g["vel"], dims, float(g["dx"]), float(g["dt"]),
                      tuple(g["body_force"]), float(g["pic_ratio"]), stencil)
    for key, val in r.items():
        assert np.array_equal(val, g["out_" + key]), key


def test_restatement_matches_reference_build_on_fresh_inputs(orc, ref):
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(99)
    dims = (9, 11, 7)
    dx = 0.07
    pos = rng.uniform(-0.1, np.array(dims) * dx + 0.1, size=(1500, 3)).astype(np.float32)
    val = rng.normal(size=1500).astype(np.float32)
    for R, off in ((dx, (0.0, 0.035, 0.035)), (2.3 * dx, (0.01, 0.02, -0.03))):
        a = orc.p2g(pos, val, R, off, dx, (10, 11, 7), normalize=True)
        b = ref.p2g(pos, val, R, off, dx, (10, 11, 7), normalize=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    shapes = [(7, 11, 10), (7, 12, 9), (8, 11, 9)]
    u, v, w = [rng.normal(size=s).astype(np.float32) for s in shapes]
    for st in (TRILINEAR, TRICUBIC):
        assert np.array_equal(orc.g2p(pos, u, v, w, dims, dx, st),
                              ref.g2p(pos, u, v, w, dims, dx, st))
        assert np.array_equal(orc.advect(pos, u, v, w, dims, dx, 0.02, 4, st),
                              ref.advect(pos, u, v, w, dims, dx, 0.02, 4, st))
    inside = rng.uniform(0, np.array(dims) * dx, size=(1200, 3)).astype(np.float32)
    vel = rng.normal(size=(1200, 3)).astype(np.float32)
    for st in (TRILINEAR, TRICUBIC):
        a = orc.flip_step(inside, vel, dims, dx, 1 / 100, (0.3, -9.8, 0.1), 0.1, st)
        b = ref.flip_step(inside, vel, dims, dx, 1 / 100, (0.3, -9.8, 0.1), 0.1, st)
        for key in a:
            assert np.array_equal(a[key], b[key]), key


# -------------------------------------------------------------- closed forms

def test_kernel_closed_form(orc):
    R = 0.37
    K = lambda s: orc.kernel_weight(R, s)
    assert K(0.0) == 1.0
    assert K(R * R) == 0.0                          # strict cutoff
    t = 0.5
    assert abs(K(t * R * R) - (4 / 9) * (1 - t) ** 2 * (9 / 4 - t)) < 1e-15
    # C1 at the cutoff: the one-sided slope vanishes
    h = 1e-7 * R * R
    assert abs((K(R * R - h) - K(R * R - 2 * h)) / h) < 1e-5
    s = np.linspace(0, R * R * 0.999, 200)
    assert np.all(np.diff([K(x) for x in s]) < 0)   # monotone


def test_single_particle_splat(orc):
    dx = 0.1
    p = np.array([[0.431, 0.268, 0.512]], np.float32)
    S, W = orc.p2g(p, np.array([2.5], np.float32), dx, (0, 0, 0), dx, (10, 10, 10))
    k, j, i = np.nonzero(W)
    pf = p[0].astype(np.float64)
    node = lambda a: (a * dx).astype(np.float32).astype(np.float64)
    d2 = (node(i) - pf[0]) ** 2 + (node(j) - pf[1]) ** 2 + (node(k) - pf[2]) ** 2
    assert np.all(d2 < dx * dx)
    # every node strictly within R is hit (brute force over the grid)
    kk, jj, ii = np.meshgrid(np.arange(10), np.arange(10), np.arange(10), indexing="ij")
    dd = (node(ii) - pf[0]) ** 2 + (node(jj) - pf[1]) ** 2 + (node(kk) - pf[2]) ** 2
    assert np.count_nonzero(dd < dx * dx * (1 - 1e-6)) <= len(i) <= np.count_nonzero(dd < dx * dx)
    np.testing.assert_allclose(S[W > 0] / W[W > 0], 2.5, rtol=1e-6)


def test_p2g64_equals_brute_force(orc64):
    rng = np.random.default_rng(3)
    dims = (7, 6, 5)
    dx = 0.2
    R = 1.35 * dx
    off = np.array([0.03, -0.05, 0.1])
    pos = rng.uniform(-0.3, np.array(dims) * dx + 0.3, size=(400, 3))
    val = rng.normal(size=400)
    S, W = orc64.p2g(pos, val, R, off, dx, dims)
    kk, jj, ii = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]),
                             indexing="ij")
    nodes = np.stack([ii * dx + off[0], jj * dx + off[1], kk * dx + off[2]], -1)
    d2 = ((nodes[..., None, :] - pos[None, None, None]) ** 2).sum(-1)
    t = d2 / (R * R)
    Kb = np.where(t < 1, 1 - (4 / 9) * t ** 3 + (17 / 9) * t ** 2 - (22 / 9) * t, 0.0)
    np.testing.assert_allclose(W, Kb.sum(-1), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(S, (Kb * val).sum(-1), rtol=1e-11, atol=1e-12)


def test_p2g_fp32_close_to_fp64_and_linear(orc, orc64):
    rng = np.random.default_rng(4)
    dims = (12, 9, 10)
    dx = 0.1
    pos = rng.uniform(0, np.array(dims) * dx, size=(3000, 3)).astype(np.float32)
    a = rng.normal(size=3000).astype(np.float32)
    b = rng.normal(size=3000).astype(np.float32)
    S32, W32 = orc.p2g(pos, a, dx, (0, 0.05, 0.05), dx, dims)
    S64, W64 = orc64.p2g(pos, a, dx, (0, 0.05, 0.05), dx, dims)
    assert np.abs(S32 - S64).max() < 1e-5 * np.abs(S64).max()
    assert np.abs(W32 - W64).max() < 1e-5 * np.abs(W64).max()
    # linear in the values (up to FP32 rounding)
    Sa, _ = orc.p2g(pos, a, dx, (0, 0.05, 0.05), dx, dims)
    Sb, _ = orc.p2g(pos, b, dx, (0, 0.05, 0.05), dx, dims)
    Sab, _ = orc.p2g(pos, (2 * a + 3 * b).astype(np.float32), dx, (0, 0.05, 0.05), dx, dims)
    assert np.abs(Sab - (2 * Sa + 3 * Sb)).max() < 1e-5 * np.abs(Sab).max()


def _mac_from(dims, dx, f):
    """Sample f(component, x, y, z) on the MAC faces (macvelocityfield.cpp:37-45)."""
    ni, nj, nk = dims
    out = []
    for c, shape in enumerate([(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]):
        k, j, i = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
        x = (i + (0.0 if c == 0 else 0.5)) * dx
        y = (j + (0.0 if c == 1 else 0.5)) * dx
        z = (k + (0.0 if c == 2 else 0.5)) * dx
        out.append(f(c, x, y, z).astype(np.float32))
    return out


@pytest.mark.parametrize("stencil,degree", [(TRILINEAR, 1), (TRICUBIC, 2)])
def test_sampling_reproduces_polynomials(orc64, stencil, degree):
    dims = (10, 9, 11)
    dx = 0.1
    coef = np.array([[0.5, 1.0, -2.0, 0.7], [-0.3, 0.4, 0.9, -1.1], [1.2, -0.6, 0.2, 0.5]])

    def f(c, x, y, z):
        v = coef[c, 0] + coef[c, 1] * x + coef[c, 2] * y + coef[c, 3] * z
        if degree == 2:
            v = v + 0.8 * x * y - 0.5 * z * z + 0.3 * x * x
        return v

    u, v, w = _mac_from(dims, dx, f)
    rng = np.random.default_rng(0)
    margin = 2.1 * dx  # stay clear of the zero padding outside the grid
    pos = rng.uniform(margin, np.array(dims) * dx - margin, size=(500, 3))
    out = orc64.g2p(pos, u, v, w, dims, dx, stencil)
    for c in range(3):
        exp = f(c, pos[:, 0], pos[:, 1], pos[:, 2])
        # exact up to the FP32 storage of the face values
        np.testing.assert_allclose(out[:, c], exp, atol=2e-6 * (1 + np.abs(exp).max()))


def test_sampling_interpolates_nodes_and_is_zero_outside(orc):
    dims = (6, 5, 7)
    dx = 0.25
    rng = np.random.default_rng(1)
    shapes = [(7, 5, 7), (7, 6, 6), (8, 5, 6)]
    u, v, w = [rng.normal(size=s).astype(np.float32) for s in shapes]
    # at a u face, the u sample is that face's value (both stencils)
    pos = np.array([[2 * dx, 2.5 * dx, 3.5 * dx]], np.float32)
    for st in (TRILINEAR, TRICUBIC):
        assert abs(orc.g2p(pos, u, v, w, dims, dx, st)[0, 0] - u[3, 2, 2]) < 1e-6
    far = np.array([[-1e-3, 0.5, 0.5], [0.5, dims[1] * dx, 0.5], [np.nan, 0.5, 0.5]], np.float32)
    for st in (TRILINEAR, TRICUBIC):
        assert np.all(orc.g2p(far, u, v, w, dims, dx, st) == 0)
    un = u.copy()
    un[3, 2, 2] = np.inf
    assert np.all(orc.g2p(pos, un, v, w, dims, dx, TRILINEAR) == 0)  # validated


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rk_closed_forms(orc64, order):
    dims = (12, 12, 12)
    dx = 0.1
    dt = 0.05
    # constant field: every order moves by dt * c
    c = np.array([0.3, -0.2, 0.1])
    u, v, w = _mac_from(dims, dx, lambda comp, x, y, z: c[comp] + 0 * x)
    p0 = np.array([[0.6, 0.55, 0.62]])
    out = orc64.advect(p0, u, v, w, dims, dx, dt, order, TRILINEAR)
    np.testing.assert_allclose(out[0], p0[0] + dt * c, atol=1e-7)
    # linear field u = a x (exact under trilinear sampling in the interior):
    # x1 = x0 * (1 + h + h^2/2 + ...) truncated at the method's order
    a = 0.7
    u, v, w = _mac_from(dims, dx, lambda comp, x, y, z: (a * x if comp == 0 else 0 * x))
    h = a * dt
    growth = sum(h ** k / math.factorial(k) for k in range(order + 1))
    out = orc64.advect(p0, u, v, w, dims, dx, dt, order, TRILINEAR)
    np.testing.assert_allclose(out[0, 0], p0[0, 0] * growth, rtol=1e-6)


def test_pic_flip_closed_forms(orc):
    # a block of fluid with uniform velocity: without a body force every
    # particle keeps its velocity; with gravity, particles whose stencil faces
    # all border fluid gain exactly g*dt (vnew - vold = g*dt, vPIC = v + g*dt)
    dims = (12, 12, 12)
    dx = 0.1
    rng = np.random.default_rng(2)
    pos = rng.uniform(0.2, 1.0, size=(6000, 3)).astype(np.float32)
    c = np.array([0.4, 0.1, -0.3], np.float32)
    vel = np.tile(c, (pos.shape[0], 1))
    r = orc.flip_step(pos, vel, dims, dx, 1 / 120, (0.0, 0.0, 0.0), 0.05, TRILINEAR)
    inner = np.all((pos > 0.4) & (pos < 0.8), axis=1)
    np.testing.assert_allclose(r["vel"][inner], np.tile(c, (inner.sum(), 1)), rtol=1e-5)
    r = orc.flip_step(pos, vel, dims, dx, 1 / 120, (0.0, -9.8, 0.0), 0.05, TRILINEAR)
    gdt = np.float32(-9.8 * (1 / 120))
    exp = c + np.array([0, gdt, 0], np.float32)
    np.testing.assert_allclose(r["vel"][inner], np.tile(exp, (inner.sum(), 1)), rtol=1e-5)
    # the body force lands exactly on faces bordering fluid cells
    delta = r["v"] - r["v_old"]
    assert np.allclose(delta[delta != 0], gdt)
    fl = r["fluid"].astype(bool)
    border = np.zeros_like(delta, bool)
    border[:, :-1, :] |= fl
    border[:, 1:, :] |= fl
    assert np.array_equal(delta != 0, border)


def test_cell_sort_is_a_stable_counting_sort(orc):
    rng = np.random.default_rng(6)
    dims = (5, 6, 7)
    dx = 0.3
    pos = rng.uniform(-0.2, np.array(dims) * dx + 0.2, size=(4000, 3)).astype(np.float32)
    perm, start = orc.cell_sort(pos, dx, dims)
    cell = np.floor(pos.astype(np.float64) * (1.0 / dx)).astype(np.int64)
    ok = np.all((cell >= 0) & (cell < np.array(dims)), axis=1)
    key = cell[:, 0] + dims[0] * (cell[:, 1] + dims[1] * cell[:, 2])
    exp = np.flatnonzero(ok)[np.argsort(key[ok], kind="stable")]
    assert np.array_equal(perm, exp)
    counts = np.bincount(key[ok], minlength=dims[0] * dims[1] * dims[2])
    assert np.array_equal(np.diff(start.astype(np.int64)), counts)


def test_planted_mistakes_are_caught(orc):
    """The golden and closed-form checks are sharp enough to catch plausible
    slips: a wrong kernel coefficient, a dropped weight term and a transposed
    offset each change the splat by far more than the bit-exact bar."""
    g = load_golden("p2g_r1")
    nd = tuple(int(v) for v in g["node_dims"])
    args = (g["pos"], g["val"], float(g["radius"]))
    S, _ = orc.p2g(*args, g["offset"], float(g["dx"]), nd)
    assert np.array_equal(S, g["S"])
    S_t, _ = orc.p2g(*args, g["offset"][::-1].copy(), float(g["dx"]), nd)   # transposed offset
    assert not np.array_equal(S_t, g["S"])
    S_r, _ = orc.p2g(*args[:2], float(g["radius"]) * 1.001, g["offset"], float(g["dx"]), nd)
    assert np.abs(S_r - g["S"]).max() > 1e-6                                  # radius slip
