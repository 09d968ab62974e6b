// dropin_harness.cpp -- test checker ONLY.
//
// Drives the drop-in CLScalarField / ParticleAdvector (gridfluidsim3d_amd/
// dropin, on the GPU through libgfsx.so) with the reference's own
// ScalarField and MACVelocityField objects, and computes the reference CPU
// answer for the same call next to it:
//   * P2G: ScalarField::addPointValue over every point (scalarfield.cpp:
//     167-201), as _addPointValuesNoCL does (clscalarfield.cpp:1527-1541);
//   * G2P: MACVelocityField::evaluateVelocityAtPosition{,Linear}
//     (macvelocityfield.cpp:545-575) with _validateOutput's NaN/Inf rule
//     (particleadvector.cpp:1139-1149);
//   * RK4: ParticleAdvector::_RK4 (particleadvector.cpp:1045-1054), restated
//     around the reference's MACVelocityField (the reference's
//     particleadvector.cpp needs <CL/cl.hpp> and is not built here).
// Built by oracle/Makefile (target dropin) into oracle/_ref/.
#include <cmath>
#include <stdexcept>
#include <cstdint>
#include <algorithm>
#include <vector>

#include "clscalarfield.h"
#include "gfsx.h"
#include "macvelocityfield.h"
#include "particleadvector.h"
#include "scalarfield.h"

namespace {
void copy_out(Array3d<float> *a, float *dst) {
    const float *src = a->getRawArray();
    for (int64_t q = 0; q < (int64_t)a->width * a->height * a->depth; q++) dst[q] = src[q];
}

void fill(Array3d<float> *a, const float *src) {
    float *dst = a->getRawArray();
    for (int64_t q = 0; q < (int64_t)a->width * a->height * a->depth; q++) dst[q] = src[q];
}

vmath::vec3 eval(MACVelocityField &m, vmath::vec3 p, int stencil) {
    return stencil == GFSX_TRICUBIC ? m.evaluateVelocityAtPosition(p)
                                    : m.evaluateVelocityAtPositionLinear(p);
}
}  // namespace

extern "C" {

// The accelerated-path toggles of both classes (clscalarfield.cpp:499-509,
// particleadvector.cpp:189-199): enabled by default, reported as set, and a
// call while disabled fails with std::runtime_error before any device work
// (this drop-in has no CPU path).  No device is needed.  Returns a bit mask
// of the checks that held (all ten: 0x3FF).
int dropin_toggles(void) {
    int ok = 0, bit = 0;
    auto mark = [&](bool c) { ok |= (c ? 1 : 0) << bit++; };
    CLScalarField acc;
    mark(acc.isOpenCLEnabled());
    acc.disableOpenCL();
    mark(!acc.isOpenCLEnabled());
    std::vector<vmath::vec3> points(4, vmath::vec3(0.1f, 0.1f, 0.1f));
    std::vector<float> values(4, 1.0f);
    Array3d<float> field(4, 4, 4, 0.0f);
    try {
        acc.addPointValues(points, values, 0.1, vmath::vec3(), 0.1, &field);
        mark(false);
    } catch (const std::runtime_error &) {
        mark(true);
    }
    mark(!acc.isUsingGPU());   // nothing ran, nothing initialised
    acc.enableOpenCL();
    mark(acc.isOpenCLEnabled());

    ParticleAdvector adv;
    mark(adv.isOpenCLEnabled() && adv.stencil() == GFSX_TRICUBIC);
    adv.disableOpenCL();
    mark(!adv.isOpenCLEnabled());
    MACVelocityField mac(4, 4, 4, 0.1);
    std::vector<vmath::vec3> out;
    try {
        adv.tricubicInterpolate(points, &mac, out);
        mark(false);
    } catch (const std::runtime_error &) {
        mark(true);
    }
    try {
        adv.setStencil(7);
        mark(false);
    } catch (const std::invalid_argument &) {
        mark(adv.stencil() == GFSX_TRICUBIC);
    }
    adv.enableOpenCL();
    mark(adv.isOpenCLEnabled() && !adv.isUsingGPU());
    return ok;
}

int dropin_p2g(const float *pos, const float *val, int64_t n, double radius,
               const float off[3], double dx, int ni, int nj, int nk,
               float *S_gpu, float *W_gpu, float *S_cpu, float *W_cpu) {
    std::vector<vmath::vec3> points((size_t)n);
    std::vector<float> values(val, val + n);
    for (int64_t p = 0; p < n; p++) points[p] = vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    const vmath::vec3 offset(off[0], off[1], off[2]);

    CLScalarField acc;
    if (!acc.initialize()) return 0;
    ScalarField gpu(ni, nj, nk, dx);
    gpu.enableWeightField();
    gpu.setPointRadius(radius);
    gpu.setOffset(offset);
    acc.addPointValues(points, values, gpu);           // the FluidSimulation call form
    copy_out(gpu.getPointerToScalarField(), S_gpu);
    copy_out(gpu.getPointerToWeightField(), W_gpu);

    ScalarField cpu(ni, nj, nk, dx);
    cpu.enableWeightField();
    cpu.setPointRadius(radius);
    cpu.setOffset(offset);
    for (int64_t p = 0; p < n; p++) cpu.addPointValue(points[p], values[p]);
    copy_out(cpu.getPointerToScalarField(), S_cpu);
    copy_out(cpu.getPointerToWeightField(), W_cpu);
    return 1;
}

// The isotropic mesher's call form (isotropicparticlemesher.cpp:332-391):
// the threshold set, then addPoints(positions, field) on batches of `batch`
// points into one ScalarField; against the mesher's CPU form, addPoint per
// point (:325-329).  threshold < 0 leaves it unset.
int dropin_add_points(const float *pos, int64_t n, double radius, const float off[3],
                      double dx, int ni, int nj, int nk, float threshold, int batch,
                      float *S_gpu, float *S_cpu) {
    std::vector<vmath::vec3> points((size_t)n);
    for (int64_t p = 0; p < n; p++) points[p] = vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    const vmath::vec3 offset(off[0], off[1], off[2]);

    CLScalarField acc;
    if (!acc.initialize()) return 0;
    if (threshold >= 0.0f) acc.setMaxScalarFieldValueThreshold(threshold);
    ScalarField gpu(ni, nj, nk, dx);
    gpu.setPointRadius(radius);
    gpu.setOffset(offset);
    std::vector<vmath::vec3> part;
    for (int64_t s = 0; s < n; s += batch) {
        part.assign(points.begin() + s, points.begin() + std::min<int64_t>(n, s + batch));
        acc.addPoints(part, gpu);
    }
    copy_out(gpu.getPointerToScalarField(), S_gpu);

    ScalarField cpu(ni, nj, nk, dx);
    cpu.setPointRadius(radius);
    cpu.setOffset(offset);
    for (int64_t p = 0; p < n; p++) cpu.addPoint(points[p]);
    copy_out(cpu.getPointerToScalarField(), S_cpu);
    return 1;
}

// ParticleAdvector::advectParticlesRK1..3 in the call form of
// DiffuseParticleSimulation (diffuseparticlesimulation.cpp:570-582, RK2
// into a separate output), against ParticleAdvector::_RK1.._RK3
// (particleadvector.cpp:1056-1078) restated around the reference's
// MACVelocityField.
int dropin_advect(const float *pos, int64_t n, const float *u, const float *v, const float *w,
                  int ni, int nj, int nk, double dx, double dt, int order, int stencil,
                  float *out_gpu, float *out_cpu) {
    MACVelocityField mac(ni, nj, nk, dx);
    fill(mac.getArray3dU(), u);
    fill(mac.getArray3dV(), v);
    fill(mac.getArray3dW(), w);
    std::vector<vmath::vec3> points((size_t)n);
    for (int64_t p = 0; p < n; p++) points[p] = vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    ParticleAdvector adv;
    if (!adv.initialize()) return 0;
    adv.setStencil(stencil);
    std::vector<vmath::vec3> moved;
    if (order == 1) adv.advectParticlesRK1(points, &mac, dt, moved);
    else if (order == 2) adv.advectParticlesRK2(points, &mac, dt, moved);
    else if (order == 3) adv.advectParticlesRK3(points, &mac, dt, moved);
    else return 0;
    if ((int64_t)moved.size() != n) return 0;
    for (int64_t p = 0; p < n; p++) {
        out_gpu[3 * p] = moved[p].x;
        out_gpu[3 * p + 1] = moved[p].y;
        out_gpu[3 * p + 2] = moved[p].z;
        const vmath::vec3 p0 = points[p];
        const vmath::vec3 k1 = eval(mac, p0, stencil);
        vmath::vec3 p1;
        if (order == 1) {
            p1 = p0 + (float)dt * k1;
        } else if (order == 2) {
            const vmath::vec3 k2 = eval(mac, p0 + (float)(0.5 * dt) * k1, stencil);
            p1 = p0 + (float)dt * k2;
        } else {
            const vmath::vec3 k2 = eval(mac, p0 + (float)(0.5 * dt) * k1, stencil);
            const vmath::vec3 k3 = eval(mac, p0 + (float)(0.75 * dt) * k2, stencil);
            p1 = p0 + (float)(dt / 9.0f) * (2.0f * k1 + 3.0f * k2 + 4.0f * k3);
        }
        out_cpu[3 * p] = p1.x;
        out_cpu[3 * p + 1] = p1.y;
        out_cpu[3 * p + 2] = p1.z;
    }
    return 1;
}

int dropin_g2p_rk4(const float *pos, int64_t n, const float *u, const float *v,
                   const float *w, int ni, int nj, int nk, double dx, double dt,
                   int stencil, float *g2p_gpu, float *g2p_cpu, float *rk4_gpu,
                   float *rk4_cpu) {
    MACVelocityField mac(ni, nj, nk, dx);
    fill(mac.getArray3dU(), u);
    fill(mac.getArray3dV(), v);
    fill(mac.getArray3dW(), w);
    std::vector<vmath::vec3> points((size_t)n);
    for (int64_t p = 0; p < n; p++) points[p] = vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);

    ParticleAdvector adv;
    if (!adv.initialize()) return 0;
    adv.setStencil(stencil);
    std::vector<vmath::vec3> out;
    adv.tricubicInterpolate(points, &mac, out);
    std::vector<vmath::vec3> moved;
    adv.advectParticlesRK4(points, &mac, dt, moved);
    if ((int64_t)out.size() != n || (int64_t)moved.size() != n) return 0;
    for (int64_t p = 0; p < n; p++) {
        g2p_gpu[3 * p] = out[p].x;
        g2p_gpu[3 * p + 1] = out[p].y;
        g2p_gpu[3 * p + 2] = out[p].z;
        rk4_gpu[3 * p] = moved[p].x;
        rk4_gpu[3 * p + 1] = moved[p].y;
        rk4_gpu[3 * p + 2] = moved[p].z;
        vmath::vec3 r = eval(mac, points[p], stencil);
        if (std::isinf(r.x) || std::isnan(r.x) || std::isinf(r.y) || std::isnan(r.y) ||
            std::isinf(r.z) || std::isnan(r.z))
            r = vmath::vec3(0.0, 0.0, 0.0);
        g2p_cpu[3 * p] = r.x;
        g2p_cpu[3 * p + 1] = r.y;
        g2p_cpu[3 * p + 2] = r.z;
        const vmath::vec3 p0 = points[p];
        const vmath::vec3 k1 = eval(mac, p0, stencil);
        const vmath::vec3 k2 = eval(mac, p0 + (float)(0.5 * dt) * k1, stencil);
        const vmath::vec3 k3 = eval(mac, p0 + (float)(0.5 * dt) * k2, stencil);
        const vmath::vec3 k4 = eval(mac, p0 + (float)dt * k3, stencil);
        const vmath::vec3 p1 = p0 + (float)(dt / 6.0f) * (k1 + 2.0f * k2 + 2.0f * k3 + k4);
        rk4_cpu[3 * p] = p1.x;
        rk4_cpu[3 * p + 1] = p1.y;
        rk4_cpu[3 * p + 2] = p1.z;
    }
    return 1;
}

}  // extern "C"
