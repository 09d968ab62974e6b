// ref_harness.cpp -- test checker ONLY.
//
// A small C ABI over the reference's own CPU classes, compiled together with
// the reference sources where they lie (oracle/Makefile, target _ref).  It
// exists to pin the C restatement in gfsx_oracle.c: every ref_* function has
// the same signature and meaning as the gfo_* function of the same name, but
// the arithmetic is done by the reference's ScalarField (scalarfield.cpp),
// MACVelocityField (macvelocityfield.cpp), Interpolation (interpolation.cpp),
// FluidMaterialGrid (fluidmaterialgrid.cpp) and vmath (vmath.cpp).
//
// The reference's ParticleAdvector (particleadvector.cpp) cannot be built
// here (it needs <CL/cl.hpp>, absent from this image), so the ten-line RK
// drivers ParticleAdvector::_RK1.._RK4 (particleadvector.cpp:1045-1078) are
// restated around the reference's MACVelocityField, and the three
// FluidSimulation stages of the plain step are composed from the reference
// classes in the order fluidsimulation.cpp:3262-3356 runs them.

#include <cmath>
#include <cstdint>
#include <cstring>

#include "fluidmaterialgrid.h"
#include "macvelocityfield.h"
#include "scalarfield.h"
#include "vmath.h"

namespace {

void fill_array(Array3d<float> *a, const float *src) {
    for (int k = 0; k < a->depth; k++)
        for (int j = 0; j < a->height; j++)
            for (int i = 0; i < a->width; i++)
                a->set(i, j, k, src[(int64_t)i + (int64_t)a->width * ((int64_t)j + (int64_t)a->height * k)]);
}

void read_array(Array3d<float> *a, float *dst, bool accumulate) {
    for (int k = 0; k < a->depth; k++)
        for (int j = 0; j < a->height; j++)
            for (int i = 0; i < a->width; i++) {
                int64_t id = (int64_t)i + (int64_t)a->width * ((int64_t)j + (int64_t)a->height * k);
                if (accumulate) dst[id] += a->get(i, j, k);
                else dst[id] = a->get(i, j, k);
            }
}

vmath::vec3 eval(MACVelocityField &m, vmath::vec3 p, int stencil) {
    return stencil ? m.evaluateVelocityAtPosition(p)
                   : m.evaluateVelocityAtPositionLinear(p);
}

vmath::vec3 validated(vmath::vec3 v) {
    // ParticleAdvector::_validateOutput (particleadvector.cpp:1139-1149)
    if (std::isinf(v.x) || std::isnan(v.x) || std::isinf(v.y) ||
        std::isnan(v.y) || std::isinf(v.z) || std::isnan(v.z))
        return vmath::vec3(0.0, 0.0, 0.0);
    return v;
}

vmath::vec3 rk(MACVelocityField &m, vmath::vec3 p0, double dt, int order,
               int stencil) {
    if (order == 4) {
        vmath::vec3 k1 = eval(m, p0, stencil);
        vmath::vec3 k2 = eval(m, p0 + (float)(0.5 * dt) * k1, stencil);
        vmath::vec3 k3 = eval(m, p0 + (float)(0.5 * dt) * k2, stencil);
        vmath::vec3 k4 = eval(m, p0 + (float)dt * k3, stencil);
        return p0 + (float)(dt / 6.0f) * (k1 + 2.0f * k2 + 2.0f * k3 + k4);
    }
    if (order == 3) {
        vmath::vec3 k1 = eval(m, p0, stencil);
        vmath::vec3 k2 = eval(m, p0 + (float)(0.5 * dt) * k1, stencil);
        vmath::vec3 k3 = eval(m, p0 + (float)(0.75 * dt) * k2, stencil);
        return p0 + (float)(dt / 9.0f) * (2.0f * k1 + 3.0f * k2 + 4.0f * k3);
    }
    if (order == 2) {
        vmath::vec3 k1 = eval(m, p0, stencil);
        vmath::vec3 k2 = eval(m, p0 + (float)(0.5 * dt) * k1, stencil);
        return p0 + (float)dt * k2;
    }
    vmath::vec3 k1 = eval(m, p0, stencil);
    return p0 + (float)dt * k1;
}

void load_mac(MACVelocityField &m, const float *u, const float *v, const float *w) {
    fill_array(m.getArray3dU(), u);
    fill_array(m.getArray3dV(), v);
    fill_array(m.getArray3dW(), w);
}

}  // namespace

extern "C" {

void ref_p2g(const float *pos, const float *val, int64_t n, double radius,
             const float offset[3], double dx, int ni, int nj, int nk,
             float *S, float *W, int apply_weight) {
    ScalarField f(ni, nj, nk, dx);
    if (W) f.enableWeightField();
    f.setPointRadius(radius);
    f.setOffset(vmath::vec3(offset[0], offset[1], offset[2]));
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 q(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        if (val) f.addPointValue(q, val[p]);
        else f.addPoint(q);
    }
    if (apply_weight) f.applyWeightField();
    read_array(f.getPointerToScalarField(), S, !apply_weight);
    if (W) read_array(f.getPointerToWeightField(), W, !apply_weight);
}

void ref_g2p(const float *pos, int64_t n, const float *u, const float *v,
             const float *w, int ni, int nj, int nk, double dx, int stencil,
             float *out) {
    MACVelocityField m(ni, nj, nk, dx);
    load_mac(m, u, v, w);
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 r = validated(eval(m, vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]), stencil));
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

void ref_advect(const float *pos, int64_t n, const float *u, const float *v,
                const float *w, int ni, int nj, int nk, double dx, double dt,
                int order, int stencil, float *out) {
    MACVelocityField m(ni, nj, nk, dx);
    load_mac(m, u, v, w);
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 r = rk(m, vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]), dt, order, stencil);
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

// The plain step, stage by stage, with the reference classes.
void ref_flip_step(float *pos, float *vel, int64_t n, int ni, int nj, int nk,
                   double dx, double dt, const float body_force[3],
                   double pic_ratio, int stencil,
                   float *u, float *v, float *w,
                   float *u_old, float *v_old, float *w_old,
                   float *wu, float *wv, float *ww, uint8_t *fluid_out) {
    // fluid cells (fluidsimulation.cpp:2010-2017)
    FluidMaterialGrid mat(ni, nj, nk);
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 q(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        GridIndex g = Grid3d::positionToGridIndex(q, dx);
        if (Grid3d::isGridIndexInRange(g, ni, nj, nk)) mat.setFluid(g);
    }

    // P2G per component (fluidsimulation.cpp:2526-2586)
    MACVelocityField mac(ni, nj, nk, dx);
    Array3d<float> *grids[3] = {mac.getArray3dU(), mac.getArray3dV(), mac.getArray3dW()};
    float *wouts[3] = {wu, wv, ww};
    std::vector<float> values((size_t)(n > 0 ? n : 1));
    for (int c = 0; c < 3; c++) {
        Array3d<float> *field = grids[c];
        ScalarField grid(field->width, field->height, field->depth, dx);
        grid.enableWeightField();
        grid.setPointRadius(dx);
        vmath::vec3 offset;
        if (c == 0) offset = vmath::vec3(0.0, 0.5 * dx, 0.5 * dx);
        else if (c == 1) offset = vmath::vec3(0.5 * dx, 0.0, 0.5 * dx);
        else offset = vmath::vec3(0.5 * dx, 0.5 * dx, 0.0);
        grid.setOffset(offset);
        for (int64_t p = 0; p < n; p++) {
            vmath::vec3 q(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
            grid.addPointValue(q, vel[3 * p + c]);
        }
        grid.applyWeightField();
        Array3d<float> *sf = grid.getPointerToScalarField();
        for (int k = 0; k < field->depth; k++)
            for (int j = 0; j < field->height; j++)
                for (int i = 0; i < field->width; i++)
                    field->set(i, j, k, sf->get(i, j, k));
        if (wouts[c]) read_array(grid.getPointerToWeightField(), wouts[c], false);
    }

    // snapshot (fluidsimulation.cpp:3306)
    MACVelocityField saved = mac;

    // constant body force (fluidsimulation.cpp:2766-2805)
    vmath::vec3 bf(body_force[0], body_force[1], body_force[2]);
    if (fabs(bf.x) > 0.0)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni + 1; i++)
                    if (mat.isFaceBorderingFluidU(i, j, k)) mac.addU(i, j, k, bf.x * dt);
    if (fabs(bf.y) > 0.0)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj + 1; j++)
                for (int i = 0; i < ni; i++)
                    if (mat.isFaceBorderingFluidV(i, j, k)) mac.addV(i, j, k, bf.y * dt);
    if (fabs(bf.z) > 0.0)
        for (int k = 0; k < nk + 1; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    if (mat.isFaceBorderingFluidW(i, j, k)) mac.addW(i, j, k, bf.z * dt);

    // PIC/FLIP (fluidsimulation.cpp:3104-3128) then RK4 (3181-3207)
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 x0(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        vmath::vec3 v0(vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]);
        vmath::vec3 vnew = validated(eval(mac, x0, stencil));
        vmath::vec3 vold = validated(eval(saved, x0, stencil));
        vmath::vec3 vPIC = vnew;
        vmath::vec3 vFLIP = v0 + vnew - vold;
        vmath::vec3 vn = (float)pic_ratio * vPIC + (float)(1 - pic_ratio) * vFLIP;
        vmath::vec3 x1 = rk(mac, x0, dt, 4, stencil);
        vel[3 * p] = vn.x;
        vel[3 * p + 1] = vn.y;
        vel[3 * p + 2] = vn.z;
        pos[3 * p] = x1.x;
        pos[3 * p + 1] = x1.y;
        pos[3 * p + 2] = x1.z;
    }

    if (u) read_array(mac.getArray3dU(), u, false);
    if (v) read_array(mac.getArray3dV(), v, false);
    if (w) read_array(mac.getArray3dW(), w, false);
    if (u_old) read_array(saved.getArray3dU(), u_old, false);
    if (v_old) read_array(saved.getArray3dV(), v_old, false);
    if (w_old) read_array(saved.getArray3dW(), w_old, false);
    if (fluid_out)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    fluid_out[(int64_t)i + (int64_t)ni * ((int64_t)j + (int64_t)nj * k)] =
                        mat.isCellFluid(i, j, k) ? 1 : 0;
}

}  // extern "C"
