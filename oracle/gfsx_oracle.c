/*
 * gfsx_oracle.c -- test checker / CPU baseline ONLY (see gfsx_oracle.h).
 *
 * A statement-by-statement restatement of the reference CPU path for the FLIP
 * transfer + advection hot path of GridFluidSim3D.  Every function cites the
 * reference file:line it follows.  The FP32/FP64 mix is kept exactly: the
 * reference stores positions and grid values in FP32 (vmath::vec3 holds
 * floats, vmath.h:53-55) and evaluates kernel weights and interpolation in
 * FP64 (scalarfield.cpp:173-199, macvelocityfield.cpp:350-575).
 *
 * Build with -ffp-contract=off and without -ffast-math: the reference is
 * compiled for plain x86-64 SSE arithmetic (CMakeLists.txt:34), so no fused
 * multiply-adds may appear here either.
 *
 * Parity pins: every gfo_* entry point is checked bit-for-bit against golden
 * vectors written by the reference's own compiled sources (tests/golden), and
 * against closed forms (kernel values, partition of unity, exactness on
 * affine/quadratic fields, RK4 on constant and linear fields) in
 * tests/test_oracle.py.  The gfo64_* twins are pinned by agreeing with gfo_*
 * to FP32 rounding and by the same closed forms.
 */
#include "gfsx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3f;
typedef struct { double x, y, z; } v3d;

static inline int64_t flat(int i, int j, int k, int w, int h) {
    /* Array3d::_getFlatIndex (array3d.h:394-397) */
    return (int64_t)i + (int64_t)w * ((int64_t)j + (int64_t)h * (int64_t)k);
}

static inline int in_range(int i, int j, int k, int w, int h, int d) {
    /* Grid3d::isGridIndexInRange (grid3d.h:153-155) */
    return i >= 0 && j >= 0 && k >= 0 && i < w && j < h && k < d;
}

/* ------------------------------------------------------------------------- */
/* Kernel weight (Wyvill "soft object" polynomial)                           */
/* ------------------------------------------------------------------------- */

typedef struct { double coef1, coef2, coef3; } kcoef;

static kcoef kernel_coefs(double r) {
    /* ScalarField::setPointRadius (scalarfield.cpp:40-46) */
    kcoef c;
    c.coef1 = (4.0 / 9.0) * (1.0 / (r * r * r * r * r * r));
    c.coef2 = (17.0 / 9.0) * (1.0 / (r * r * r * r));
    c.coef3 = (22.0 / 9.0) * (1.0 / (r * r));
    return c;
}

static inline double kernel_eval(const kcoef *c, double s) {
    /* ScalarField::_evaluateTricubicFieldFunctionForRadiusSquared
       (scalarfield.cpp:569-571), same operation order */
    return 1.0 - c->coef1 * s * s * s + c->coef2 * s * s - c->coef3 * s;
}

double gfo_kernel_weight(double radius, double distsq) {
    kcoef c = kernel_coefs(radius);
    if (!(distsq < radius * radius)) return 0.0;
    return kernel_eval(&c, distsq);
}

/* ------------------------------------------------------------------------- */
/* P2G                                                                        */
/* ------------------------------------------------------------------------- */

static void index_bounds(float px, float py, float pz, double r, double dx,
                         int imax, int jmax, int kmax, int g1[3], int g2[3]) {
    /* Grid3d::getGridIndexBounds (grid3d.h:350-371) with
       positionToGridIndex (grid3d.h:58-63) and GridIndexToPosition
       (grid3d.h:83-85): cell index from an FP64 floor, cell corner rounded
       back to FP32, FP32 offset inside the cell. */
    double invdx = 1.0 / dx;
    int ci = (int)floor(px * invdx);
    int cj = (int)floor(py * invdx);
    int ck = (int)floor(pz * invdx);
    float cx = (float)((float)ci * dx);
    float cy = (float)((float)cj * dx);
    float cz = (float)((float)ck * dx);
    float tx = px - cx, ty = py - cy, tz = pz - cz;
    double inv = 1.0 / dx;
    int gimin = ci - (int)fmax(0, ceil((r - tx) * inv));
    int gjmin = cj - (int)fmax(0, ceil((r - ty) * inv));
    int gkmin = ck - (int)fmax(0, ceil((r - tz) * inv));
    int gimax = ci + (int)fmax(0, ceil((r - dx + tx) * inv));
    int gjmax = cj + (int)fmax(0, ceil((r - dx + ty) * inv));
    int gkmax = ck + (int)fmax(0, ceil((r - dx + tz) * inv));
    g1[0] = (int)fmax(gimin, 0);
    g1[1] = (int)fmax(gjmin, 0);
    g1[2] = (int)fmax(gkmin, 0);
    g2[0] = (int)fmin(gimax, imax - 1);
    g2[1] = (int)fmin(gjmax, jmax - 1);
    g2[2] = (int)fmin(gkmax, kmax - 1);
}

void gfo_p2g(const float *pos, const float *val, int64_t n, double radius,
             const float offset[3], double dx, int ni, int nj, int nk,
             float *S, float *W) {
    /* ScalarField::addPointValue (scalarfield.cpp:167-201) applied to every
       point in order, as _addPointValuesNoCL does (clscalarfield.cpp:1527-
       1541) but accumulating into one field, as the OpenCL path does. */
    kcoef c = kernel_coefs(radius);
    double rsq = radius * radius;
    for (int64_t p = 0; p < n; p++) {
        /* p -= _gridOffset (FP32 vec3 subtraction, scalarfield.cpp:168) */
        float px = pos[3 * p + 0] - offset[0];
        float py = pos[3 * p + 1] - offset[1];
        float pz = pos[3 * p + 2] - offset[2];
        double scale = val ? (double)val[p] : 1.0;
        int g1[3], g2[3];
        index_bounds(px, py, pz, radius, dx, ni, nj, nk, g1, g2);
        for (int k = g1[2]; k <= g2[2]; k++) {
            for (int j = g1[1]; j <= g2[1]; j++) {
                for (int i = g1[0]; i <= g2[0]; i++) {
                    /* GridIndexToPosition (grid3d.h:79-81) */
                    float gx = (float)((float)i * dx);
                    float gy = (float)((float)j * dx);
                    float gz = (float)((float)k * dx);
                    float vx = gx - px, vy = gy - py, vz = gz - pz;
                    /* vmath::dot is an FP32 expression (vmath.h:72-74) */
                    float d32 = vx * vx + vy * vy + vz * vz;
                    double distsq = d32;
                    if (distsq < rsq) {
                        double weight = kernel_eval(&c, distsq);
                        int64_t id = flat(i, j, k, ni, nj);
                        /* Array3d<float>::add takes a float (array3d.h:263-271) */
                        S[id] += (float)(weight * scale);
                        if (W) W[id] += (float)weight;
                    }
                }
            }
        }
    }
}

void gfo_apply_weight(float *S, const float *W, int64_t count) {
    /* ScalarField::applyWeightField (scalarfield.cpp:90-106) */
    for (int64_t id = 0; id < count; id++) {
        float weight = W[id];
        if (weight > 0.0) {
            float v = S[id] / weight;
            S[id] = v;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* G2P                                                                        */
/* ------------------------------------------------------------------------- */

typedef struct {
    const void *f[3];  /* u, v, w */
    int isdbl;
    int ni, nj, nk;
    double dx;
} macgrid;

static inline double fetch(const macgrid *g, int c, int i, int j, int k) {
    /* MACVelocityField::U/V/W: out-of-range samples read 0
       (macvelocityfield.cpp:99-121) */
    int w = g->ni + (c == 0), h = g->nj + (c == 1), d = g->nk + (c == 2);
    if (!in_range(i, j, k, w, h, d)) return 0.0;
    int64_t id = flat(i, j, k, w, h);
    return g->isdbl ? ((const double *)g->f[c])[id]
                    : (double)((const float *)g->f[c])[id];
}

static inline int in_grid(double x, double y, double z, double dx, int i,
                          int j, int k) {
    /* Grid3d::isPositionInGrid (grid3d.h:137-139) */
    return x >= 0 && y >= 0 && z >= 0 && x < dx * i && y < dx * j && z < dx * k;
}

static double trilinear(const double p[8], double x, double y, double z) {
    /* Interpolation::trilinearInterpolate (interpolation.cpp:51-60); corner
       order {000,100,010,001,101,011,110,111}, same operation order */
    return p[0] * (1 - x) * (1 - y) * (1 - z) +
           p[1] * x * (1 - y) * (1 - z) +
           p[2] * (1 - x) * y * (1 - z) +
           p[3] * (1 - x) * (1 - y) * z +
           p[4] * x * (1 - y) * z +
           p[5] * (1 - x) * y * z +
           p[6] * x * y * (1 - z) +
           p[7] * x * y * z;
}

static inline double cubic(const double p[4], double x) {
    /* Interpolation::cubicInterpolate (interpolation.cpp:44-46): Catmull-Rom
       (Keys a = -1/2) in Horner form, same operation order */
    return p[1] + 0.5 * x * (p[2] - p[0] +
                             x * (2.0 * p[0] - 5.0 * p[1] + 4.0 * p[2] - p[3] +
                                  x * (3.0 * (p[1] - p[2]) + p[3] - p[0])));
}

static double tricubic(double p[4][4][4], double x, double y, double z) {
    /* Interpolation::tricubicInterpolate / bicubicInterpolate
       (interpolation.cpp:26-42): x first, then y, then z */
    double zarr[4];
    for (int pk = 0; pk < 4; pk++) {
        double yarr[4];
        for (int pj = 0; pj < 4; pj++) yarr[pj] = cubic(p[pk][pj], x);
        zarr[pk] = cubic(yarr, y);
    }
    return cubic(zarr, z);
}

static double interp_component(const macgrid *g, int c, double x, double y,
                               double z, int stencil) {
    /* MACVelocityField::_interpolateU/V/W (macvelocityfield.cpp:350-450) and
       _interpolateLinearU/V/W (macvelocityfield.cpp:452-543) */
    double dx = g->dx;
    if (!in_grid(x, y, z, dx, g->ni, g->nj, g->nk)) return 0.0;
    if (c != 0) x -= 0.5 * dx;
    if (c != 1) y -= 0.5 * dx;
    if (c != 2) z -= 0.5 * dx;
    double invdx = 1.0 / dx;
    int i = (int)floor(x * invdx);
    int j = (int)floor(y * invdx);
    int k = (int)floor(z * invdx);
    double gx = (double)i * dx, gy = (double)j * dx, gz = (double)k * dx;
    double inv_dx = 1 / dx;
    double ix = (x - gx) * inv_dx;
    double iy = (y - gy) * inv_dx;
    double iz = (z - gz) * inv_dx;
    if (stencil == GFO_TRILINEAR) {
        double pts[8];
        pts[0] = fetch(g, c, i, j, k);
        pts[1] = fetch(g, c, i + 1, j, k);
        pts[2] = fetch(g, c, i, j + 1, k);
        pts[3] = fetch(g, c, i, j, k + 1);
        pts[4] = fetch(g, c, i + 1, j, k + 1);
        pts[5] = fetch(g, c, i, j + 1, k + 1);
        pts[6] = fetch(g, c, i + 1, j + 1, k);
        pts[7] = fetch(g, c, i + 1, j + 1, k + 1);
        return trilinear(pts, ix, iy, iz);
    }
    double pts[4][4][4];
    for (int pk = 0; pk < 4; pk++)
        for (int pj = 0; pj < 4; pj++)
            for (int pi = 0; pi < 4; pi++)
                pts[pk][pj][pi] = fetch(g, c, i - 1 + pi, j - 1 + pj, k - 1 + pk);
    return tricubic(pts, ix, iy, iz);
}

static v3d eval_velocity(const macgrid *g, double x, double y, double z,
                         int stencil) {
    /* MACVelocityField::evaluateVelocityAtPosition{,Linear}
       (macvelocityfield.cpp:549-559, 565-575) */
    v3d r = {0.0, 0.0, 0.0};
    if (!in_grid(x, y, z, g->dx, g->ni, g->nj, g->nk)) return r;
    r.x = interp_component(g, 0, x, y, z, stencil);
    r.y = interp_component(g, 1, x, y, z, stencil);
    r.z = interp_component(g, 2, x, y, z, stencil);
    return r;
}

static inline v3f eval_f(const macgrid *g, v3f p, int stencil) {
    /* the vec3 return value rounds each FP64 component to FP32 */
    v3d r = eval_velocity(g, p.x, p.y, p.z, stencil);
    v3f o = {(float)r.x, (float)r.y, (float)r.z};
    return o;
}

static inline v3f validated(v3f v) {
    /* ParticleAdvector::_validateOutput (particleadvector.cpp:1139-1149) */
    if (isinf(v.x) || isnan(v.x) || isinf(v.y) || isnan(v.y) || isinf(v.z) ||
        isnan(v.z)) {
        v3f z = {0.0f, 0.0f, 0.0f};
        return z;
    }
    return v;
}

void gfo_g2p(const float *pos, int64_t n, const float *u, const float *v,
             const float *w, int ni, int nj, int nk, double dx, int stencil,
             float *out) {
    macgrid g = {{u, v, w}, 0, ni, nj, nk, dx};
    for (int64_t p = 0; p < n; p++) {
        v3f q = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        v3f r = validated(eval_f(&g, q, stencil));
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

/* vmath::vec3 operators are component-wise FP32 (vmath.cpp:41-110) */
static inline v3f vadd(v3f a, v3f b) { v3f r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3f vsub(v3f a, v3f b) { v3f r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3f vscale(float s, v3f a) { v3f r = {s * a.x, s * a.y, s * a.z}; return r; }

static v3f rk_step(const macgrid *g, v3f p0, double dt, int order, int stencil) {
    if (order == 4) {
        /* ParticleAdvector::_RK4 (particleadvector.cpp:1045-1054) */
        v3f k1 = eval_f(g, p0, stencil);
        v3f k2 = eval_f(g, vadd(p0, vscale((float)(0.5 * dt), k1)), stencil);
        v3f k3 = eval_f(g, vadd(p0, vscale((float)(0.5 * dt), k2)), stencil);
        v3f k4 = eval_f(g, vadd(p0, vscale((float)dt, k3)), stencil);
        v3f sum = vadd(vadd(vadd(k1, vscale(2.0f, k2)), vscale(2.0f, k3)), k4);
        return vadd(p0, vscale((float)(dt / 6.0f), sum));
    } else if (order == 3) {
        /* ParticleAdvector::_RK3 (particleadvector.cpp:1056-1063) */
        v3f k1 = eval_f(g, p0, stencil);
        v3f k2 = eval_f(g, vadd(p0, vscale((float)(0.5 * dt), k1)), stencil);
        v3f k3 = eval_f(g, vadd(p0, vscale((float)(0.75 * dt), k2)), stencil);
        v3f sum = vadd(vadd(vscale(2.0f, k1), vscale(3.0f, k2)), vscale(4.0f, k3));
        return vadd(p0, vscale((float)(dt / 9.0f), sum));
    } else if (order == 2) {
        /* ParticleAdvector::_RK2 (particleadvector.cpp:1065-1071) */
        v3f k1 = eval_f(g, p0, stencil);
        v3f k2 = eval_f(g, vadd(p0, vscale((float)(0.5 * dt), k1)), stencil);
        return vadd(p0, vscale((float)dt, k2));
    }
    /* ParticleAdvector::_RK1 (particleadvector.cpp:1073-1078) */
    v3f k1 = eval_f(g, p0, stencil);
    return vadd(p0, vscale((float)dt, k1));
}

void gfo_advect(const float *pos, int64_t n, const float *u, const float *v,
                const float *w, int ni, int nj, int nk, double dx, double dt,
                int order, int stencil, float *out) {
    /* _advectParticlesRK{1..4}NoCL (particleadvector.cpp:1080-1122) */
    macgrid g = {{u, v, w}, 0, ni, nj, nk, dx};
    for (int64_t p = 0; p < n; p++) {
        v3f q = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        v3f r = rk_step(&g, q, dt, order, stencil);
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

/* ------------------------------------------------------------------------- */
/* Fluid cells, body force, the plain step                                   */
/* ------------------------------------------------------------------------- */

static inline int cell_of(double x, double y, double z, double dx, int ni,
                          int nj, int nk, int64_t *cell) {
    /* Grid3d::positionToGridIndex (grid3d.h:58-63) */
    double invdx = 1.0 / dx;
    double fi = floor(x * invdx), fj = floor(y * invdx), fk = floor(z * invdx);
    if (!(fi >= 0 && fj >= 0 && fk >= 0 && fi < ni && fj < nj && fk < nk))
        return 0;
    *cell = flat((int)fi, (int)fj, (int)fk, ni, nj);
    return 1;
}

static void apply_body_force_component(float *f, double *fd, int c,
                                       const uint8_t *fluid, int ni, int nj,
                                       int nk, double delta) {
    /* FluidSimulation::_applyConstantBodyForces (fluidsimulation.cpp:2766-
       2805): add body_force[c]*dt to every face that borders a fluid cell
       (FluidMaterialGrid::isFaceBorderingMaterialU/V/W,
       fluidmaterialgrid.cpp:119-143) */
    int w = ni + (c == 0), h = nj + (c == 1), d = nk + (c == 2);
    int naxis = c == 0 ? ni : (c == 1 ? nj : nk);
    for (int k = 0; k < d; k++)
        for (int j = 0; j < h; j++)
            for (int i = 0; i < w; i++) {
                int a = c == 0 ? i : (c == 1 ? j : k);
                int border = 0;
                int ci = i, cj = j, ck = k;
                if (a == naxis) {
                    if (c == 0) ci--; else if (c == 1) cj--; else ck--;
                    border = fluid[flat(ci, cj, ck, ni, nj)];
                } else if (a > 0) {
                    int li = i - (c == 0), lj = j - (c == 1), lk = k - (c == 2);
                    border = fluid[flat(i, j, k, ni, nj)] ||
                             fluid[flat(li, lj, lk, ni, nj)];
                } else {
                    border = fluid[flat(i, j, k, ni, nj)];
                }
                if (border) {
                    int64_t id = flat(i, j, k, w, h);
                    /* MACVelocityField::addU/V/W: Array3d<float>::add of a
                       (float)-cast double (macvelocityfield.cpp:204-226) */
                    if (f) f[id] += (float)delta;
                    if (fd) fd[id] += delta;
                }
            }
}

static void update_particles(float *pos, float *vel, int64_t n, float *const grid[3],
                             float *const old[3], int ni, int nj, int nk, double dx,
                             double dt, double pic_ratio, int stencil) {
    macgrid gnew = {{grid[0], grid[1], grid[2]}, 0, ni, nj, nk, dx};
    macgrid gold = {{old[0], old[1], old[2]}, 0, ni, nj, nk, dx};
    float r_pic = (float)pic_ratio, r_flip = (float)(1 - pic_ratio);
    for (int64_t p = 0; p < n; p++) {
        v3f x0 = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        v3f v0 = {vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
        /* 5. PIC/FLIP (fluidsimulation.cpp:3115-3127) */
        v3f vnew = validated(eval_f(&gnew, x0, stencil));
        v3f vold = validated(eval_f(&gold, x0, stencil));
        v3f vflip = vsub(vadd(v0, vnew), vold);
        v3f vn = vadd(vscale(r_pic, vnew), vscale(r_flip, vflip));
        /* 6. RK4 through the new field (fluidsimulation.cpp:3181-3207,
           particleadvector.cpp:1045-1054) */
        v3f x1 = rk_step(&gnew, x0, dt, 4, stencil);
        vel[3 * p] = vn.x;
        vel[3 * p + 1] = vn.y;
        vel[3 * p + 2] = vn.z;
        pos[3 * p] = x1.x;
        pos[3 * p + 1] = x1.y;
        pos[3 * p + 2] = x1.z;
    }
}

void gfo_flip_step(float *pos, float *vel, int64_t n, int ni, int nj, int nk,
                   double dx, double dt, const float body_force[3],
                   double pic_ratio, int stencil,
                   float *u, float *v, float *w,
                   float *u_old, float *v_old, float *w_old,
                   float *wu, float *wv, float *ww, uint8_t *fluid_out) {
    int64_t ncell = (int64_t)ni * nj * nk;
    int64_t nf[3] = {(int64_t)(ni + 1) * nj * nk, (int64_t)ni * (nj + 1) * nk,
                     (int64_t)ni * nj * (nk + 1)};
    int dims[3][3] = {{ni + 1, nj, nk}, {ni, nj + 1, nk}, {ni, nj, nk + 1}};
    uint8_t *fluid = (uint8_t *)calloc((size_t)ncell, 1);
    float *grid[3], *old[3], *wts[3];
    float *outs[3] = {u, v, w}, *olds[3] = {u_old, v_old, w_old},
          *wouts[3] = {wu, wv, ww};
    float *vals = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));

    /* 1. fluid cells = cells holding a particle (fluidsimulation.cpp:2010-2017) */
    for (int64_t p = 0; p < n; p++) {
        int64_t cell;
        if (cell_of(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], dx, ni, nj, nk,
                    &cell))
            fluid[cell] = 1;
    }

    /* 2. P2G of each velocity component onto its faces with the weight field,
       R = dx and the face offset (fluidsimulation.cpp:2526-2570), then the
       normalisation (fluidsimulation.cpp:2570, scalarfield.cpp:90-106) */
    for (int c = 0; c < 3; c++) {
        grid[c] = (float *)calloc((size_t)nf[c], sizeof(float));
        wts[c] = (float *)calloc((size_t)nf[c], sizeof(float));
        float off[3];
        off[0] = c == 0 ? 0.0f : (float)(0.5 * dx);
        off[1] = c == 1 ? 0.0f : (float)(0.5 * dx);
        off[2] = c == 2 ? 0.0f : (float)(0.5 * dx);
        for (int64_t p = 0; p < n; p++) vals[p] = vel[3 * p + c];
        gfo_p2g(pos, vals, n, dx, off, dx, dims[c][0], dims[c][1], dims[c][2],
                grid[c], wts[c]);
        gfo_apply_weight(grid[c], wts[c], nf[c]);
        /* 3. snapshot (fluidsimulation.cpp:3306) */
        old[c] = (float *)malloc(sizeof(float) * (size_t)nf[c]);
        memcpy(old[c], grid[c], sizeof(float) * (size_t)nf[c]);
        /* 4. constant body force (fluidsimulation.cpp:2766-2805) */
        if (fabs((double)body_force[c]) > 0.0)
            apply_body_force_component(grid[c], NULL, c, fluid, ni, nj, nk,
                                       (double)body_force[c] * dt);
    }

    update_particles(pos, vel, n, grid, old, ni, nj, nk, dx, dt, pic_ratio, stencil);

    for (int c = 0; c < 3; c++) {
        if (outs[c]) memcpy(outs[c], grid[c], sizeof(float) * (size_t)nf[c]);
        if (olds[c]) memcpy(olds[c], old[c], sizeof(float) * (size_t)nf[c]);
        if (wouts[c]) memcpy(wouts[c], wts[c], sizeof(float) * (size_t)nf[c]);
        free(grid[c]);
        free(old[c]);
        free(wts[c]);
    }
    if (fluid_out) memcpy(fluid_out, fluid, (size_t)ncell);
    free(vals);
    free(fluid);
}

/* ------------------------------------------------------------------------- */
/* The step with FluidSimulation's face fill and extrapolation               */
/* ------------------------------------------------------------------------- */

enum { MAT_AIR = 0, MAT_FLUID = 1, MAT_SOLID = 2 };

typedef struct {
    const uint8_t *mat;   /* ni*nj*nk material codes */
    int ni, nj, nk;
} matgrid;

static inline void comp_dims(int c, int ni, int nj, int nk, int d[3]) {
    d[0] = ni + (c == 0);
    d[1] = nj + (c == 1);
    d[2] = nk + (c == 2);
}

/* FluidMaterialGrid::isFaceBorderingMaterialU/V/W (fluidmaterialgrid.cpp:
   119-143) over any per-cell code array, and MACVelocityField::
   _isFaceBorderingGridValueU/V/W (macvelocityfield.h:159-173) over the layer
   grid: the cell on the high side unless the face is the last one, the cell
   on the low side unless the face is the first. */
static int face_borders_u8(const uint8_t *cell, int c, int i, int j, int k, int ni, int nj,
                           int nk, uint8_t value) {
    int a = c == 0 ? i : (c == 1 ? j : k);
    int na = c == 0 ? ni : (c == 1 ? nj : nk);
    int li = i - (c == 0), lj = j - (c == 1), lk = k - (c == 2);
    if (a == na) return cell[flat(li, lj, lk, ni, nj)] == value;
    if (a > 0)
        return cell[flat(i, j, k, ni, nj)] == value || cell[flat(li, lj, lk, ni, nj)] == value;
    return cell[flat(i, j, k, ni, nj)] == value;
}

static int face_borders_i32(const int *cell, int c, int i, int j, int k, int ni, int nj,
                            int nk, int value) {
    int a = c == 0 ? i : (c == 1 ? j : k);
    int na = c == 0 ? ni : (c == 1 ? nj : nk);
    int li = i - (c == 0), lj = j - (c == 1), lk = k - (c == 2);
    if (a == na) return cell[flat(li, lj, lk, ni, nj)] == value;
    if (a > 0)
        return cell[flat(i, j, k, ni, nj)] == value || cell[flat(li, lj, lk, ni, nj)] == value;
    return cell[flat(i, j, k, ni, nj)] == value;
}

/* FluidSimulation::_advectVelocityFieldU/V/W (fluidsimulation.cpp:2597-2730):
   the MAC component is cleared; faces bordering fluid take the normalised
   P2G value where the weight exceeds 1e-9 (isValueSet, :2575-2585), else
   the average of their 26 neighbours (the neighbour test is |value| > 0 for
   U, :2630, but isValueSet for V and W, :2675,2720). */
static void fill_component(float *out, const float *ugrid, const float *W, int c,
                           const matgrid *m) {
    int d[3];
    comp_dims(c, m->ni, m->nj, m->nk, d);
    int64_t nf = (int64_t)d[0] * d[1] * d[2];
    memset(out, 0, sizeof(float) * (size_t)nf);
    for (int k = 0; k < d[2]; k++)
        for (int j = 0; j < d[1]; j++)
            for (int i = 0; i < d[0]; i++) {
                if (!face_borders_u8(m->mat, c, i, j, k, m->ni, m->nj, m->nk, MAT_FLUID))
                    continue;
                int64_t f = flat(i, j, k, d[0], d[1]);
                if ((double)W[f] > 1e-9) {
                    out[f] = ugrid[f];
                    continue;
                }
                double avg = 0.0, weight = 0.0;
                for (int nk2 = k - 1; nk2 <= k + 1; nk2++)
                    for (int nj2 = j - 1; nj2 <= j + 1; nj2++)
                        for (int ni2 = i - 1; ni2 <= i + 1; ni2++) {
                            if (ni2 == i && nj2 == j && nk2 == k) continue;
                            if (!in_range(ni2, nj2, nk2, d[0], d[1], d[2])) continue;
                            int64_t g = flat(ni2, nj2, nk2, d[0], d[1]);
                            int use = c == 0 ? fabs(ugrid[g]) > 0.0 : (double)W[g] > 1e-9;
                            if (use) {
                                avg += ugrid[g];
                                weight += 1.0;
                            }
                        }
                if (weight > 0.0) out[f] = (float)(avg / weight);
            }
}

/* MACVelocityField::extrapolateVelocityField (macvelocityfield.cpp:577-798):
   faces not bordering fluid are cleared; cells get BFS layers (fluid 0,
   then 6-neighbour dilation through non-solid cells); for each layer L the
   faces that border layer L, not layer L-1 and no solid cell take the
   average of their 6 neighbour faces that border layer L-1 (0 if that sum
   is 0). */
static void extrapolate(float *g[3], const matgrid *m, int layers) {
    int ni = m->ni, nj = m->nj, nk = m->nk;
    int64_t ncell = (int64_t)ni * nj * nk;
    int *layer = (int *)malloc(sizeof(int) * (size_t)(ncell > 0 ? ncell : 1));
    for (int c = 0; c < 3; c++) {
        int d[3];
        comp_dims(c, ni, nj, nk, d);
        for (int k = 0; k < d[2]; k++)
            for (int j = 0; j < d[1]; j++)
                for (int i = 0; i < d[0]; i++)
                    if (!face_borders_u8(m->mat, c, i, j, k, ni, nj, nk, MAT_FLUID))
                        g[c][flat(i, j, k, d[0], d[1])] = 0.0f;
    }
    for (int64_t q = 0; q < ncell; q++) layer[q] = m->mat[q] == MAT_FLUID ? 0 : -1;
    for (int L = 1; L <= layers; L++)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    int64_t q = flat(i, j, k, ni, nj);
                    if (layer[q] != L - 1 || m->mat[q] == MAT_SOLID) continue;
                    int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                    {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                    for (int e = 0; e < 6; e++) {
                        if (!in_range(nb[e][0], nb[e][1], nb[e][2], ni, nj, nk)) continue;
                        int64_t r = flat(nb[e][0], nb[e][1], nb[e][2], ni, nj);
                        if (layer[r] == -1 && m->mat[r] != MAT_SOLID) layer[r] = L;
                    }
                }
    for (int L = 1; L <= layers; L++)
        for (int c = 0; c < 3; c++) {
            int d[3];
            comp_dims(c, ni, nj, nk, d);
            for (int k = 0; k < d[2]; k++)
                for (int j = 0; j < d[1]; j++)
                    for (int i = 0; i < d[0]; i++) {
                        if (!(face_borders_i32(layer, c, i, j, k, ni, nj, nk, L) &&
                              !face_borders_i32(layer, c, i, j, k, ni, nj, nk, L - 1) &&
                              !face_borders_u8(m->mat, c, i, j, k, ni, nj, nk, MAT_SOLID)))
                            continue;
                        int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                        {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                        double sum = 0.0, wsum = 0.0;
                        for (int e = 0; e < 6; e++) {
                            int a = nb[e][0], b = nb[e][1], cc = nb[e][2];
                            if (!in_range(a, b, cc, d[0], d[1], d[2])) continue;
                            if (!face_borders_i32(layer, c, a, b, cc, ni, nj, nk, L - 1)) continue;
                            sum += g[c][flat(a, b, cc, d[0], d[1])];
                            wsum++;
                        }
                        double v = sum == 0.0 ? 0.0 : sum / wsum;
                        g[c][flat(i, j, k, d[0], d[1])] = (float)v;
                    }
        }
    free(layer);
}

/* Collision::getLineSegmentVoxelIntersection (collision.cpp:303-394): the
   first solid voxel along the segment p0 -> p1, walking the voxels with the
   reference's error accumulators; the points are scaled by (float)(1/dx) in
   FP32 as vec3 *= float does. */
static int segment_voxel(v3f p0, v3f p1, double dx, const uint8_t *mat, int ni, int nj,
                         int nk, int vox[3]) {
    float invdx = (float)(1.0 / dx);
    p0.x *= invdx; p0.y *= invdx; p0.z *= invdx;
    p1.x *= invdx; p1.y *= invdx; p1.z *= invdx;
    int gx0 = (int)floor(p0.x), gy0 = (int)floor(p0.y), gz0 = (int)floor(p0.z);
    int gx1 = (int)floor(p1.x), gy1 = (int)floor(p1.y), gz1 = (int)floor(p1.z);
    int sx = gx1 > gx0 ? 1 : gx1 < gx0 ? -1 : 0;
    int sy = gy1 > gy0 ? 1 : gy1 < gy0 ? -1 : 0;
    int sz = gz1 > gz0 ? 1 : gz1 < gz0 ? -1 : 0;
    int gx = gx0, gy = gy0, gz = gz0;
    double gxp = gx0 + (gx1 > gx0 ? 1 : 0);
    double gyp = gy0 + (gy1 > gy0 ? 1 : 0);
    double gzp = gz0 + (gz1 > gz0 ? 1 : 0);
    double vx = p1.x == p0.x ? 1 : (double)(p1.x - p0.x);
    double vy = p1.y == p0.y ? 1 : (double)(p1.y - p0.y);
    double vz = p1.z == p0.z ? 1 : (double)(p1.z - p0.z);
    double vxvy = vx * vy, vxvz = vx * vz, vyvz = vy * vz;
    double errx = (gxp - p0.x) * vyvz;
    double erry = (gyp - p0.y) * vxvz;
    double errz = (gzp - p0.z) * vxvy;
    double derrx = sx * vyvz, derry = sy * vxvz, derrz = sz * vxvy;
    /* the walk reaches the end voxel in |dgx| + |dgy| + |dgz| steps */
    long maxsteps = labs((long)gx1 - gx0) + labs((long)gy1 - gy0) + labs((long)gz1 - gz0) + 1;
    for (long it = 0; it <= maxsteps; it++) {
        if (in_range(gx, gy, gz, ni, nj, nk) && mat[flat(gx, gy, gz, ni, nj)] == MAT_SOLID) {
            vox[0] = gx; vox[1] = gy; vox[2] = gz;
            return 1;
        }
        if (gx == gx1 && gy == gy1 && gz == gz1) break;
        double xr = fabs(errx), yr = fabs(erry), zr = fabs(errz);
        if (sx != 0 && (sy == 0 || xr < yr) && (sz == 0 || xr < zr)) {
            gx += sx;
            errx += derrx;
        } else if (sy != 0 && (sz == 0 || yr < zr)) {
            gy += sy;
            erry += derry;
        } else if (sz != 0) {
            gz += sz;
            errz += derrz;
        }
    }
    return 0;
}

/* Collision::rayIntersectsAABB (collision.cpp:404-449), FP32 vec3 maths
   and FP64 slab bounds, with the reference's test of dir.x where dir.z is
   clamped (:414-416). */
static int ray_aabb(v3f p0, v3f dir, v3f bmin, v3f bmax, v3f *hit) {
    double eps = 1e-10;
    if (fabs(dir.x) < eps) dir.x = dir.x < 0 ? (float)-eps : (float)eps;
    if (fabs(dir.y) < eps) dir.y = dir.y < 0 ? (float)-eps : (float)eps;
    if (fabs(dir.x) < eps) dir.z = dir.z < 0 ? (float)-eps : (float)eps;
    v3f inv = {(float)(1.0 / dir.x), (float)(1.0 / dir.y), (float)(1.0 / dir.z)};
    double t1 = (bmin.x - p0.x) * inv.x;
    double t2 = (bmax.x - p0.x) * inv.x;
    double tmin = fmin(t1, t2), tmax = fmax(t1, t2);
    t1 = (bmin.y - p0.y) * inv.y;
    t2 = (bmax.y - p0.y) * inv.y;
    tmin = fmax(tmin, fmin(t1, t2));
    tmax = fmin(tmax, fmax(t1, t2));
    t1 = (bmin.z - p0.z) * inv.z;
    t2 = (bmax.z - p0.z) * inv.z;
    tmin = fmax(tmin, fmin(t1, t2));
    tmax = fmin(tmax, fmax(t1, t2));
    if (tmax > fmax(tmin, 0.0)) {
        *hit = vadd(p0, vscale((float)tmin, dir));
        return 1;
    }
    return 0;
}

static int cell_is_solid(v3f p, double dx, const uint8_t *mat, int ni, int nj, int nk) {
    int64_t cell;
    return cell_of(p.x, p.y, p.z, dx, ni, nj, nk, &cell) && mat[cell] == MAT_SOLID;
}

/* FluidSimulation::_resolveParticleSolidCellCollision (fluidsimulation.cpp:
   3145-3179): back off 0.05 dx along the ray from where it enters the first
   solid voxel; p0 when there is no hit or the backed-off point is solid. */
static v3f resolve_collision(v3f p0, v3f p1, double dx, const uint8_t *mat, int ni, int nj,
                             int nk) {
    int vox[3];
    if (!segment_voxel(p0, p1, dx, mat, ni, nj, nk, vox)) return p0;
    v3f d = vsub(p1, p0);
    float len = (float)sqrt((double)(d.x * d.x + d.y * d.y + d.z * d.z));
    float inv = (float)(1.0 / len);
    v3f raynorm = {d.x * inv, d.y * inv, d.z * inv};
    v3f vpos = {(float)((float)vox[0] * dx), (float)((float)vox[1] * dx),
                (float)((float)vox[2] * dx)};
    v3f vmax = {vpos.x + (float)dx, vpos.y + (float)dx, vpos.z + (float)dx};
    v3f cpoint;
    if (!ray_aabb(p0, raynorm, vpos, vmax, &cpoint)) return p0;
    v3f r = vsub(cpoint, vscale((float)(0.05 * dx), raynorm));
    if (cell_is_solid(r, dx, mat, ni, nj, nk)) return p0;
    return r;
}

void gfo_fill_component(float *out, const float *ugrid, const float *W, int c,
                        const uint8_t *mat, int ni, int nj, int nk) {
    matgrid m = {mat, ni, nj, nk};
    fill_component(out, ugrid, W, c, &m);
}

void gfo_extrapolate(float *u, float *v, float *w, int ni, int nj, int nk,
                     const uint8_t *mat, int layers) {
    matgrid m = {mat, ni, nj, nk};
    float *g[3] = {u, v, w};
    extrapolate(g, &m, layers);
}

void gfo_flip_step_fs(float *pos, float *vel, int64_t n, int ni, int nj, int nk,
                      double dx, double dt, const float body_force[3], double pic_ratio,
                      int stencil, const uint8_t *solid, int layers,
                      float *u, float *v, float *w, float *u_old, float *v_old,
                      float *w_old, uint8_t *mat_out) {
    int64_t ncell = (int64_t)ni * nj * nk;
    int64_t nf[3] = {(int64_t)(ni + 1) * nj * nk, (int64_t)ni * (nj + 1) * nk,
                     (int64_t)ni * nj * (nk + 1)};
    int dims[3][3] = {{ni + 1, nj, nk}, {ni, nj + 1, nk}, {ni, nj, nk + 1}};
    uint8_t *mat = (uint8_t *)calloc((size_t)(ncell > 0 ? ncell : 1), 1);
    uint8_t *fluid = (uint8_t *)calloc((size_t)(ncell > 0 ? ncell : 1), 1);
    float *grid[3], *old[3];
    float *outs[3] = {u, v, w}, *olds[3] = {u_old, v_old, w_old};
    float *vals = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    float *S = NULL, *W = NULL;

    /* solid cells, then fluid cells = non-solid cells holding a particle
       (fluidsimulation.cpp:1990-2017; the reference removes particles from
       solid cells beforehand, :1991) */
    for (int64_t q = 0; q < ncell; q++) mat[q] = solid && solid[q] ? MAT_SOLID : MAT_AIR;
    for (int64_t p = 0; p < n; p++) {
        int64_t cell;
        if (cell_of(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], dx, ni, nj, nk, &cell) &&
            mat[cell] != MAT_SOLID)
            mat[cell] = MAT_FLUID;
    }
    for (int64_t q = 0; q < ncell; q++) fluid[q] = mat[q] == MAT_FLUID;
    matgrid m = {mat, ni, nj, nk};

    for (int c = 0; c < 3; c++) {
        grid[c] = (float *)calloc((size_t)nf[c], sizeof(float));
        S = (float *)calloc((size_t)nf[c], sizeof(float));
        W = (float *)calloc((size_t)nf[c], sizeof(float));
        float off[3];
        off[0] = c == 0 ? 0.0f : (float)(0.5 * dx);
        off[1] = c == 1 ? 0.0f : (float)(0.5 * dx);
        off[2] = c == 2 ? 0.0f : (float)(0.5 * dx);
        for (int64_t p = 0; p < n; p++) vals[p] = vel[3 * p + c];
        gfo_p2g(pos, vals, n, dx, off, dx, dims[c][0], dims[c][1], dims[c][2], S, W);
        gfo_apply_weight(S, W, nf[c]);
        fill_component(grid[c], S, W, c, &m);
        free(S);
        free(W);
        /* snapshot (fluidsimulation.cpp:3306) */
        old[c] = (float *)malloc(sizeof(float) * (size_t)nf[c]);
        memcpy(old[c], grid[c], sizeof(float) * (size_t)nf[c]);
    }
    /* the snapshot is extrapolated (:3307), the field gets the body force
       (:2766-2805) and is extrapolated after the (omitted) pressure solve
       (:3334), with ceil(CFL + 2) layers (:3067-3070) */
    extrapolate(old, &m, layers);
    for (int c = 0; c < 3; c++)
        if (fabs((double)body_force[c]) > 0.0)
            apply_body_force_component(grid[c], NULL, c, fluid, ni, nj, nk,
                                       (double)body_force[c] * dt);
    extrapolate(grid, &m, layers);

    float *x0 = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? 3 * n : 1));
    memcpy(x0, pos, sizeof(float) * (size_t)(3 * n));
    update_particles(pos, vel, n, grid, old, ni, nj, nk, dx, dt, pic_ratio, stencil);
    /* a particle moved into a solid cell is pulled back to where its path
       enters the solid (fluidsimulation.cpp:3181-3207) */
    if (solid)
        for (int64_t p = 0; p < n; p++) {
            v3f a = {x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]};
            v3f b = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
            if (!cell_is_solid(b, dx, mat, ni, nj, nk)) continue;
            v3f r = resolve_collision(a, b, dx, mat, ni, nj, nk);
            pos[3 * p] = r.x;
            pos[3 * p + 1] = r.y;
            pos[3 * p + 2] = r.z;
        }
    free(x0);

    for (int c = 0; c < 3; c++) {
        if (outs[c]) memcpy(outs[c], grid[c], sizeof(float) * (size_t)nf[c]);
        if (olds[c]) memcpy(olds[c], old[c], sizeof(float) * (size_t)nf[c]);
        free(grid[c]);
        free(old[c]);
    }
    if (mat_out) memcpy(mat_out, mat, (size_t)ncell);
    free(vals);
    free(fluid);
    free(mat);
}

/* ------------------------------------------------------------------------- */
/* Particle removal                                                           */
/* ------------------------------------------------------------------------- */

/* SplitMix64 (Steele, Lea, Flood 2014; constants of Vigna's splitmix64.c):
   a bijection of i for a fixed seed, so keys never tie. */
uint64_t gfo_removal_key(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* FluidSimulation::_removeMarkerParticlesInSolidCells and
   _removeMarkerParticles (fluidsimulation.cpp:1985-1988, 3213-3243): drop
   particles in solid cells, then keep at most max_per_cell per cell.  The
   reference keeps the first ones of a rand() shuffle (:3213-3221); here a
   cell keeps its max_per_cell particles of smallest gfo_removal_key(seed,
   index).  Particles outside the grid are kept.  Returns the number kept;
   keep[p] is set to 0/1. */
int64_t gfo_remove_particles(const float *pos, int64_t n, double dx, int ni, int nj, int nk,
                             const uint8_t *solid, int max_per_cell, uint64_t seed,
                             uint8_t *keep) {
    int64_t ncell = (int64_t)ni * nj * nk;
    int64_t *cellof = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t kept = 0;
    for (int64_t p = 0; p < n; p++) {
        int64_t c;
        cellof[p] = cell_of(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], dx, ni, nj, nk, &c) ? c : -1;
    }
    /* brute force: a particle's rank among the non-solid particles of its cell */
    int64_t *count = (int64_t *)calloc((size_t)(ncell > 0 ? ncell : 1), sizeof(int64_t));
    for (int64_t p = 0; p < n; p++)
        if (cellof[p] >= 0) count[cellof[p]]++;
    for (int64_t p = 0; p < n; p++) {
        int64_t c = cellof[p];
        int k = 1;
        if (c >= 0 && solid && solid[c]) {
            k = 0;
        } else if (c >= 0 && count[c] > max_per_cell) {
            uint64_t kp = gfo_removal_key(seed, (uint64_t)p);
            int64_t rank = 0;
            for (int64_t q = 0; q < n; q++)
                if (cellof[q] == c && gfo_removal_key(seed, (uint64_t)q) < kp) rank++;
            k = rank < max_per_cell;
        }
        keep[p] = (uint8_t)k;
        kept += k;
    }
    free(count);
    free(cellof);
    return kept;
}

int64_t gfo_cell_sort(const float *pos, int64_t n, double dx, int ni, int nj,
                      int nk, uint32_t *perm, uint32_t *cell_start) {
    /* Stable counting sort keyed by Grid3d::positionToGridIndex
       (grid3d.h:58-63).  This is the build's binning (the reference bins
       into work-group chunks instead, clscalarfield.cpp:882-942); it is
       pinned against a brute-force sort in the tests. */
    int64_t ncell = (int64_t)ni * nj * nk;
    memset(cell_start, 0, sizeof(uint32_t) * (size_t)(ncell + 1));
    for (int64_t p = 0; p < n; p++) {
        int64_t cell;
        if (cell_of(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], dx, ni, nj, nk,
                    &cell))
            cell_start[cell + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) cell_start[c + 1] += cell_start[c];
    uint32_t *cursor = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(ncell > 0 ? ncell : 1));
    memcpy(cursor, cell_start, sizeof(uint32_t) * (size_t)ncell);
    for (int64_t p = 0; p < n; p++) {
        int64_t cell;
        if (cell_of(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], dx, ni, nj, nk,
                    &cell))
            perm[cursor[cell]++] = (uint32_t)p;
    }
    free(cursor);
    return (int64_t)cell_start[ncell];
}

/* ------------------------------------------------------------------------- */
/* FP64 twins                                                                 */
/* ------------------------------------------------------------------------- */

void gfo64_p2g(const double *pos, const double *val, int64_t n, double radius,
               const double offset[3], double dx, int ni, int nj, int nk,
               double *S, double *W) {
    /* scalarfield.cpp:167-201 evaluated in FP64 throughout */
    kcoef c = kernel_coefs(radius);
    double rsq = radius * radius;
    for (int64_t p = 0; p < n; p++) {
        double px = pos[3 * p] - offset[0];
        double py = pos[3 * p + 1] - offset[1];
        double pz = pos[3 * p + 2] - offset[2];
        double scale = val ? val[p] : 1.0;
        double invdx = 1.0 / dx;
        int ci = (int)floor(px * invdx), cj = (int)floor(py * invdx),
            ck = (int)floor(pz * invdx);
        double tx = px - ci * dx, ty = py - cj * dx, tz = pz - ck * dx;
        int g1[3], g2[3];
        g1[0] = (int)fmax(ci - (int)fmax(0, ceil((radius - tx) * invdx)), 0);
        g1[1] = (int)fmax(cj - (int)fmax(0, ceil((radius - ty) * invdx)), 0);
        g1[2] = (int)fmax(ck - (int)fmax(0, ceil((radius - tz) * invdx)), 0);
        g2[0] = (int)fmin(ci + (int)fmax(0, ceil((radius - dx + tx) * invdx)), ni - 1);
        g2[1] = (int)fmin(cj + (int)fmax(0, ceil((radius - dx + ty) * invdx)), nj - 1);
        g2[2] = (int)fmin(ck + (int)fmax(0, ceil((radius - dx + tz) * invdx)), nk - 1);
        for (int k = g1[2]; k <= g2[2]; k++)
            for (int j = g1[1]; j <= g2[1]; j++)
                for (int i = g1[0]; i <= g2[0]; i++) {
                    double vx = i * dx - px, vy = j * dx - py, vz = k * dx - pz;
                    double distsq = vx * vx + vy * vy + vz * vz;
                    if (distsq < rsq) {
                        double weight = kernel_eval(&c, distsq);
                        int64_t id = flat(i, j, k, ni, nj);
                        S[id] += weight * scale;
                        if (W) W[id] += weight;
                    }
                }
    }
}

void gfo64_g2p(const double *pos, int64_t n, const double *u, const double *v,
               const double *w, int ni, int nj, int nk, double dx, int stencil,
               double *out) {
    macgrid g = {{u, v, w}, 1, ni, nj, nk, dx};
    for (int64_t p = 0; p < n; p++) {
        v3d r = eval_velocity(&g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], stencil);
        if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) r.x = r.y = r.z = 0.0;
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

static v3d rk_step64(const macgrid *g, v3d p0, double dt, int order, int stencil) {
    v3d k1 = eval_velocity(g, p0.x, p0.y, p0.z, stencil), k2, k3, k4, r;
    if (order == 1) {
        r.x = p0.x + dt * k1.x; r.y = p0.y + dt * k1.y; r.z = p0.z + dt * k1.z;
        return r;
    }
    double h = 0.5 * dt;
    k2 = eval_velocity(g, p0.x + h * k1.x, p0.y + h * k1.y, p0.z + h * k1.z, stencil);
    if (order == 2) {
        r.x = p0.x + dt * k2.x; r.y = p0.y + dt * k2.y; r.z = p0.z + dt * k2.z;
        return r;
    }
    if (order == 3) {
        double q = 0.75 * dt;
        k3 = eval_velocity(g, p0.x + q * k2.x, p0.y + q * k2.y, p0.z + q * k2.z, stencil);
        double s = dt / 9.0;
        r.x = p0.x + s * (2.0 * k1.x + 3.0 * k2.x + 4.0 * k3.x);
        r.y = p0.y + s * (2.0 * k1.y + 3.0 * k2.y + 4.0 * k3.y);
        r.z = p0.z + s * (2.0 * k1.z + 3.0 * k2.z + 4.0 * k3.z);
        return r;
    }
    k3 = eval_velocity(g, p0.x + h * k2.x, p0.y + h * k2.y, p0.z + h * k2.z, stencil);
    k4 = eval_velocity(g, p0.x + dt * k3.x, p0.y + dt * k3.y, p0.z + dt * k3.z, stencil);
    double s = dt / 6.0;
    r.x = p0.x + s * (k1.x + 2.0 * k2.x + 2.0 * k3.x + k4.x);
    r.y = p0.y + s * (k1.y + 2.0 * k2.y + 2.0 * k3.y + k4.y);
    r.z = p0.z + s * (k1.z + 2.0 * k2.z + 2.0 * k3.z + k4.z);
    return r;
}

void gfo64_advect(const double *pos, int64_t n, const double *u,
                  const double *v, const double *w, int ni, int nj, int nk,
                  double dx, double dt, int order, int stencil, double *out) {
    macgrid g = {{u, v, w}, 1, ni, nj, nk, dx};
    for (int64_t p = 0; p < n; p++) {
        v3d q = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        v3d r = rk_step64(&g, q, dt, order, stencil);
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

void gfo64_flip_step(double *pos, double *vel, int64_t n, int ni, int nj,
                     int nk, double dx, double dt, const double body_force[3],
                     double pic_ratio, int stencil,
                     double *u, double *v, double *w,
                     double *u_old, double *v_old, double *w_old,
                     double *wu, double *wv, double *ww) {
    int64_t ncell = (int64_t)ni * nj * nk;
    int64_t nf[3] = {(int64_t)(ni + 1) * nj * nk, (int64_t)ni * (nj + 1) * nk,
                     (int64_t)ni * nj * (nk + 1)};
    int dims[3][3] = {{ni + 1, nj, nk}, {ni, nj + 1, nk}, {ni, nj, nk + 1}};
    uint8_t *fluid = (uint8_t *)calloc((size_t)ncell, 1);
    double *grid[3], *old[3], *wts[3];
    double *outs[3] = {u, v, w}, *olds[3] = {u_old, v_old, w_old},
           *wouts[3] = {wu, wv, ww};
    double *vals = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t p = 0; p < n; p++) {
        int64_t cell;
        if (cell_of(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], dx, ni, nj, nk, &cell))
            fluid[cell] = 1;
    }
    for (int c = 0; c < 3; c++) {
        grid[c] = (double *)calloc((size_t)nf[c], sizeof(double));
        wts[c] = (double *)calloc((size_t)nf[c], sizeof(double));
        double off[3];
        off[0] = c == 0 ? 0.0 : 0.5 * dx;
        off[1] = c == 1 ? 0.0 : 0.5 * dx;
        off[2] = c == 2 ? 0.0 : 0.5 * dx;
        for (int64_t p = 0; p < n; p++) vals[p] = vel[3 * p + c];
        gfo64_p2g(pos, vals, n, dx, off, dx, dims[c][0], dims[c][1], dims[c][2],
                  grid[c], wts[c]);
        for (int64_t id = 0; id < nf[c]; id++)
            if (wts[c][id] > 0.0) grid[c][id] = grid[c][id] / wts[c][id];
        old[c] = (double *)malloc(sizeof(double) * (size_t)nf[c]);
        memcpy(old[c], grid[c], sizeof(double) * (size_t)nf[c]);
        if (fabs(body_force[c]) > 0.0)
            apply_body_force_component(NULL, grid[c], c, fluid, ni, nj, nk,
                                       body_force[c] * dt);
    }
    macgrid gnew = {{grid[0], grid[1], grid[2]}, 1, ni, nj, nk, dx};
    macgrid gold = {{old[0], old[1], old[2]}, 1, ni, nj, nk, dx};
    for (int64_t p = 0; p < n; p++) {
        v3d x0 = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        v3d vnew = eval_velocity(&gnew, x0.x, x0.y, x0.z, stencil);
        v3d vold = eval_velocity(&gold, x0.x, x0.y, x0.z, stencil);
        double r_flip = 1 - pic_ratio;
        vel[3 * p] = pic_ratio * vnew.x + r_flip * (vel[3 * p] + vnew.x - vold.x);
        vel[3 * p + 1] = pic_ratio * vnew.y + r_flip * (vel[3 * p + 1] + vnew.y - vold.y);
        vel[3 * p + 2] = pic_ratio * vnew.z + r_flip * (vel[3 * p + 2] + vnew.z - vold.z);
        v3d x1 = rk_step64(&gnew, x0, dt, 4, stencil);
        pos[3 * p] = x1.x;
        pos[3 * p + 1] = x1.y;
        pos[3 * p + 2] = x1.z;
    }
    for (int c = 0; c < 3; c++) {
        if (outs[c]) memcpy(outs[c], grid[c], sizeof(double) * (size_t)nf[c]);
        if (olds[c]) memcpy(olds[c], old[c], sizeof(double) * (size_t)nf[c]);
        if (wouts[c]) memcpy(wouts[c], wts[c], sizeof(double) * (size_t)nf[c]);
        free(grid[c]);
        free(old[c]);
        free(wts[c]);
    }
    free(vals);
    free(fluid);
}
