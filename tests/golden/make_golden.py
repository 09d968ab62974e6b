"""Writes the golden vectors in this directory from the reference's own CPU
sources (oracle/_ref/libgfsx_ref.so, built by `make -C oracle ref` from
/root/reference/src).  Run where the reference tree is present:

    python tests/golden/make_golden.py

Each .npz holds the inputs and the reference outputs of one case.  The
reference arithmetic comes from ScalarField::addPointValue / addPoint /
applyWeightField (scalarfield.cpp:90-201), MACVelocityField::
evaluateVelocityAtPosition{,Linear} (macvelocityfield.cpp:545-575) and the
FluidSimulation stages composed in oracle/ref_harness.cpp.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402


def uniform_particles(rng, n, dims, dx, margin):
    lo = -margin * dx
    hi = np.array(dims) * dx + margin * dx
    return rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)


def main():
    orc.build(with_reference=True)
    ref = orc.load_reference()
    assert ref is not None, "reference build missing"
    rng = np.random.default_rng(20261202)
    dims = (12, 10, 9)
    dx = 0.1
    cases = {}

    # P2G with weights, R = dx and the u-face offset; R = 1.7 dx, odd offset
    pos = uniform_particles(rng, 3000, dims, dx, 1.5)
    val = rng.normal(size=3000).astype(np.float32)
    for name, R, off, nd in (("p2g_r1", dx, (0.0, 0.05, 0.05), (13, 10, 9)),
                             ("p2g_r17", 1.7 * dx, (0.01, -0.02, 0.03), (12, 10, 9))):
        S, W = ref.p2g(pos, val, R, off, dx, nd, weight=True, normalize=False)
        Sn, Wn = ref.p2g(pos, val, R, off, dx, nd, weight=True, normalize=True)
        cases[name] = dict(pos=pos, val=val, radius=np.float64(R),
                           offset=np.array(off, np.float32), dx=np.float64(dx),
                           node_dims=np.array(nd), S=S, W=W, S_norm=Sn)
    S, _ = ref.p2g(pos, None, 1.3 * dx, (0.02, 0.0, -0.01), dx, (12, 10, 9),
                   weight=False)
    cases["addpoints"] = dict(pos=pos, radius=np.float64(1.3 * dx),
                              offset=np.array((0.02, 0.0, -0.01), np.float32),
                              dx=np.float64(dx), node_dims=np.array((12, 10, 9)), S=S)

    # G2P and RK1..4 on a random MAC field
    shapes = orc.mac_shapes(dims)
    u, v, w = [rng.normal(size=s).astype(np.float32) for s in shapes]
    pos = uniform_particles(rng, 3000, dims, dx, 0.5)
    for st, sname in ((orc.TRILINEAR, "trilinear"), (orc.TRICUBIC, "tricubic")):
        out = ref.g2p(pos, u, v, w, dims, dx, st)
        adv = {("rk%d" % o): ref.advect(pos, u, v, w, dims, dx, 0.01, o, st)
               for o in (1, 2, 3, 4)}
        cases["g2p_" + sname] = dict(pos=pos, u=u, v=v, w=w, dims=np.array(dims),
                                     dx=np.float64(dx), dt=np.float64(0.01),
                                     out=out, **adv)

    # the plain step on a small sphere with Gaussian velocities
    sdims = (16, 14, 12)
    sdx = 0.125
    k, j, i = np.meshgrid(np.arange(sdims[2]), np.arange(sdims[1]),
                          np.arange(sdims[0]), indexing="ij")
    c = np.array(sdims) * sdx / 2
    mask = (((i + 0.5) * sdx - c[0]) ** 2 + ((j + 0.5) * sdx - c[1]) ** 2 +
            ((k + 0.5) * sdx - c[2]) ** 2) < (0.75 * c.min()) ** 2
    cells = np.flatnonzero(mask.reshape(-1))
    ci = cells % sdims[0]
    cj = (cells // sdims[0]) % sdims[1]
    ck = cells // (sdims[0] * sdims[1])
    base = np.stack([ci, cj, ck], 1).astype(np.float64)
    sub = np.array([(a, b, cc) for cc in (0.25, 0.75) for b in (0.25, 0.75)
                    for a in (0.25, 0.75)])
    spos = ((base[:, None, :] + sub[None]) * sdx +
            rng.uniform(-0.025 * sdx, 0.025 * sdx, size=(cells.size, 8, 3)))
    spos = spos.reshape(-1, 3).astype(np.float32)
    svel = rng.normal(size=spos.shape).astype(np.float32)
    for st, sname in ((orc.TRILINEAR, "trilinear"), (orc.TRICUBIC, "tricubic")):
        r = ref.flip_step(spos, svel, sdims, sdx, 1.0 / 120.0, (0.0, -9.8, 0.0),
                          0.05, st)
        cases["step_" + sname] = dict(pos=spos, vel=svel, dims=np.array(sdims),
                                      dx=np.float64(sdx), dt=np.float64(1.0 / 120.0),
                                      body_force=np.array((0.0, -9.8, 0.0), np.float32),
                                      pic_ratio=np.float64(0.05),
                                      **{"out_" + key: val for key, val in r.items()})

    for name, arrays in cases.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
