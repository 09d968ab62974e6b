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

#include "collision.h"
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

}  // extern "C"

namespace {

// FluidSimulation::_advectVelocityFieldU/V/W (fluidsimulation.cpp:2597-2730)
// restated around the reference's classes (FluidSimulation itself does not
// compile here): the component is cleared, faces bordering fluid take the
// P2G value where the weight exceeds 1e-9, else the average of their 26
// neighbours (|value| > 0 for U, isValueSet for V and W).
void fill_face_values(MACVelocityField &mac, FluidMaterialGrid &mat, int c,
                      Array3d<float> &ugrid, Array3d<float> &weight) {
    const double eps = 1e-9;
    Array3d<bool> isValueSet(ugrid.width, ugrid.height, ugrid.depth, false);
    for (int k = 0; k < ugrid.depth; k++)
        for (int j = 0; j < ugrid.height; j++)
            for (int i = 0; i < ugrid.width; i++)
                if (weight(i, j, k) > eps) isValueSet.set(i, j, k, true);
    if (c == 0) mac.clearU();
    else if (c == 1) mac.clearV();
    else mac.clearW();
    std::vector<GridIndex> todo;
    for (int k = 0; k < ugrid.depth; k++)
        for (int j = 0; j < ugrid.height; j++)
            for (int i = 0; i < ugrid.width; i++) {
                bool border = c == 0 ? mat.isFaceBorderingFluidU(i, j, k)
                            : c == 1 ? mat.isFaceBorderingFluidV(i, j, k)
                                     : mat.isFaceBorderingFluidW(i, j, k);
                if (!border) continue;
                if (!isValueSet(i, j, k)) {
                    todo.push_back(GridIndex(i, j, k));
                } else if (c == 0) {
                    mac.setU(i, j, k, ugrid(i, j, k));
                } else if (c == 1) {
                    mac.setV(i, j, k, ugrid(i, j, k));
                } else {
                    mac.setW(i, j, k, ugrid(i, j, k));
                }
            }
    GridIndex nb[26];
    for (size_t t = 0; t < todo.size(); t++) {
        GridIndex g = todo[t];
        Grid3d::getNeighbourGridIndices26(g, nb);
        double avg = 0.0, wsum = 0.0;
        for (int e = 0; e < 26; e++) {
            GridIndex q = nb[e];
            if (!ugrid.isIndexInRange(q)) continue;
            bool use = c == 0 ? fabs(ugrid(q)) > 0.0 : isValueSet(q);
            if (use) {
                avg += ugrid(q);
                wsum += 1.0;
            }
        }
        if (wsum > 0.0) {
            if (c == 0) mac.setU(g, avg / wsum);
            else if (c == 1) mac.setV(g, avg / wsum);
            else mac.setW(g, avg / wsum);
        }
    }
}

void step_impl(float *pos, float *vel, int64_t n, int ni, int nj, int nk, double dx, double dt,
               const float body_force[3], double pic_ratio, int stencil, bool fs,
               const uint8_t *solid, int layers, float *u, float *v, float *w, float *u_old,
               float *v_old, float *w_old, float *wu, float *wv, float *ww, uint8_t *mat_out);

// FluidSimulation::_resolveParticleSolidCellCollision (fluidsimulation.cpp:
// 3145-3179) restated around the reference's own Collision functions; a
// backed-off point outside the grid counts as not solid (the reference's
// domain is walled by solid cells, so it never gets there).
vmath::vec3 resolve_collision(FluidMaterialGrid &mat, double dx, vmath::vec3 p0, vmath::vec3 p1,
                              int ni, int nj, int nk) {
    GridIndex voxel;
    if (!Collision::getLineSegmentVoxelIntersection(p0, p1, dx, mat, &voxel)) return p0;
    vmath::vec3 raynorm = vmath::normalize(p1 - p0);
    vmath::vec3 vpos = Grid3d::GridIndexToPosition(voxel, dx);
    AABB bbox(vpos, dx, dx, dx);
    vmath::vec3 cpoint;
    if (!Collision::rayIntersectsAABB(p0, raynorm, bbox, &cpoint)) return p0;
    vmath::vec3 r = cpoint - (float)(0.05 * dx) * raynorm;
    GridIndex gr = Grid3d::positionToGridIndex(r, dx);
    if (Grid3d::isGridIndexInRange(gr, ni, nj, nk) && mat.isCellSolid(gr)) return p0;
    return r;
}

}  // namespace

extern "C" {

// The plain step, stage by stage, with the reference classes.
void ref_flip_step(float *pos, float *vel, int64_t n, int ni, int nj, int nk,
                   double dx, double dt, const float body_force[3],
                   double pic_ratio, int stencil,
                   float *u, float *v, float *w,
                   float *u_old, float *v_old, float *w_old,
                   float *wu, float *wv, float *ww, uint8_t *fluid_out) {
    step_impl(pos, vel, n, ni, nj, nk, dx, dt, body_force, pic_ratio, stencil, false, nullptr,
              0, u, v, w, u_old, v_old, w_old, wu, wv, ww, fluid_out);
}

// The step with FluidSimulation's face fill and extrapolation (snapshot
// extrapolated, body force, field extrapolated: fluidsimulation.cpp:3304-3334
// without the pressure solve) using the reference's own
// MACVelocityField::extrapolateVelocityField and FluidMaterialGrid.
void ref_flip_step_fs(float *pos, float *vel, int64_t n, int ni, int nj, int nk, double dx,
                      double dt, const float body_force[3], double pic_ratio, int stencil,
                      const uint8_t *solid, int layers, float *u, float *v, float *w,
                      float *u_old, float *v_old, float *w_old, uint8_t *mat_out) {
    step_impl(pos, vel, n, ni, nj, nk, dx, dt, body_force, pic_ratio, stencil, true, solid,
              layers, u, v, w, u_old, v_old, w_old, nullptr, nullptr, nullptr, mat_out);
}

// MACVelocityField::extrapolateVelocityField alone, on the given field and
// cell codes (0 air, 1 fluid, 2 solid), in place.
void ref_extrapolate(float *u, float *v, float *w, int ni, int nj, int nk, double dx,
                     const uint8_t *mat_codes, int layers) {
    FluidMaterialGrid mat(ni, nj, nk);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                uint8_t m = mat_codes[(int64_t)i + (int64_t)ni * ((int64_t)j + (int64_t)nj * k)];
                if (m == 1) mat.setFluid(i, j, k);
                else if (m == 2) mat.setSolid(i, j, k);
            }
    MACVelocityField mac(ni, nj, nk, dx);
    load_mac(mac, u, v, w);
    mac.extrapolateVelocityField(mat, layers);
    read_array(mac.getArray3dU(), u, false);
    read_array(mac.getArray3dV(), v, false);
    read_array(mac.getArray3dW(), w, false);
}

}  // extern "C"

namespace {

void step_impl(float *pos, float *vel, int64_t n, int ni, int nj, int nk, double dx, double dt,
               const float body_force[3], double pic_ratio, int stencil, bool fs,
               const uint8_t *solid, int layers, float *u, float *v, float *w, float *u_old,
               float *v_old, float *w_old, float *wu, float *wv, float *ww, uint8_t *fluid_out) {
    // solid cells, then fluid cells (fluidsimulation.cpp:1990-2017)
    FluidMaterialGrid mat(ni, nj, nk);
    if (solid)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    if (solid[(int64_t)i + (int64_t)ni * ((int64_t)j + (int64_t)nj * k)])
                        mat.setSolid(i, j, k);
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 q(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        GridIndex g = Grid3d::positionToGridIndex(q, dx);
        if (Grid3d::isGridIndexInRange(g, ni, nj, nk) && !mat.isCellSolid(g)) mat.setFluid(g);
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
        if (fs) {
            fill_face_values(mac, mat, c, *sf, *grid.getPointerToWeightField());
        } else {
            for (int k = 0; k < field->depth; k++)
                for (int j = 0; j < field->height; j++)
                    for (int i = 0; i < field->width; i++)
                        field->set(i, j, k, sf->get(i, j, k));
        }
        if (wouts[c]) read_array(grid.getPointerToWeightField(), wouts[c], false);
    }

    // snapshot (fluidsimulation.cpp:3306), extrapolated in the epilogue form (:3307)
    MACVelocityField saved = mac;
    if (fs) saved.extrapolateVelocityField(mat, layers);

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

    // the field is extrapolated after the (omitted) pressure solve (:3334)
    if (fs) mac.extrapolateVelocityField(mat, layers);

    // PIC/FLIP (fluidsimulation.cpp:3104-3128) then RK4 (3181-3207), with
    // the solid-cell collision fix of _advanceRangeOfMarkerParticles
    for (int64_t p = 0; p < n; p++) {
        vmath::vec3 x0(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        vmath::vec3 v0(vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]);
        vmath::vec3 vnew = validated(eval(mac, x0, stencil));
        vmath::vec3 vold = validated(eval(saved, x0, stencil));
        vmath::vec3 vPIC = vnew;
        vmath::vec3 vFLIP = v0 + vnew - vold;
        vmath::vec3 vn = (float)pic_ratio * vPIC + (float)(1 - pic_ratio) * vFLIP;
        vmath::vec3 x1 = rk(mac, x0, dt, 4, stencil);
        if (fs && solid) {
            GridIndex g = Grid3d::positionToGridIndex(x1, dx);
            if (Grid3d::isGridIndexInRange(g, ni, nj, nk) && mat.isCellSolid(g))
                x1 = resolve_collision(mat, dx, x0, x1, ni, nj, nk);
        }
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
                        mat.isCellFluid(i, j, k) ? 1 : (mat.isCellSolid(i, j, k) ? 2 : 0);
}

}  // namespace
