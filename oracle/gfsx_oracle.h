/*
 * gfsx_oracle.h -- CPU restatement of GridFluidSim3D's FLIP transfer + advection
 * path, used ONLY as a test checker and as the CPU baseline of bench.py.
 * Nothing in the product library links or calls this code.
 *
 * Two families:
 *   gfo_*   : the reference's own mixed precision (FP32 storage and vec3 maths,
 *             FP64 kernel weights / interpolation, FP32 accumulation in particle
 *             order), restated statement by statement.
 *   gfo64_* : the same algorithms evaluated entirely in FP64 (inputs promoted),
 *             used to report the error of both the GPU path and gfo_* against
 *             a higher-precision answer.
 *
 * Layouts follow the reference: positions/velocities are AoS float triples
 * (std::vector<vmath::vec3>, vmath.h:33-57); grids are flat i-fastest arrays
 * (Array3d::_getFlatIndex, array3d.h:394-397) with MAC dimensions
 * u:(ni+1,nj,nk) v:(ni,nj+1,nk) w:(ni,nj,nk+1) (macvelocityfield.cpp:37-45).
 */
#ifndef GFSX_ORACLE_H
#define GFSX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GFO_TRILINEAR = 0, GFO_TRICUBIC = 1 };

/* ScalarField::addPointValue over all points in order, weight field on
 * (scalarfield.cpp:167-201).  val == NULL means ScalarField::addPoint
 * (scalarfield.cpp:126-160): the splatted value is the kernel weight.
 * S and W (W nullable) are accumulated into, like the OpenCL path
 * (clscalarfield.cpp:1427-1455). */
void gfo_p2g(const float *pos, const float *val, int64_t n, double radius,
             const float offset[3], double dx, int ni, int nj, int nk,
             float *S, float *W);

/* ScalarField::applyWeightField (scalarfield.cpp:90-106). */
void gfo_apply_weight(float *S, const float *W, int64_t count);

/* ParticleAdvector::_tricubicInterpolateNoCL + _validateOutput
 * (particleadvector.cpp:1124-1149), with the TRILINEAR stencil taken from
 * MACVelocityField::evaluateVelocityAtPositionLinear (macvelocityfield.cpp:561-575). */
void gfo_g2p(const float *pos, int64_t n, const float *u, const float *v,
             const float *w, int ni, int nj, int nk, double dx, int stencil,
             float *out);

/* ParticleAdvector::_RK1.._RK4 (particleadvector.cpp:1045-1078) through the
 * NoCL drivers (particleadvector.cpp:1080-1122). order in 1..4. */
void gfo_advect(const float *pos, int64_t n, const float *u, const float *v,
                const float *w, int ni, int nj, int nk, double dx, double dt,
                int order, int stencil, float *out);

/* The canonical plain FLIP step (P2G u,v,w with weights -> normalise ->
 * snapshot -> constant body force on fluid-bordering faces -> PIC/FLIP ->
 * RK4), composed from fluidsimulation.cpp:2526-2595, 2766-2805, 2010-2017,
 * 3104-3129, 3181-3209.  pos/vel are updated in place (AoS).  Grid outputs
 * (u,v,w after the body force; u_old.. the snapshot; wu.. the weights) are
 * written when non-NULL.  fluid (ni*nj*nk bytes) receives the fluid-cell
 * mask when non-NULL. */
void gfo_flip_step(float *pos, float *vel, int64_t n, int ni, int nj, int nk,
                   double dx, double dt, const float body_force[3],
                   double pic_ratio, int stencil,
                   float *u, float *v, float *w,
                   float *u_old, float *v_old, float *w_old,
                   float *wu, float *wv, float *ww, uint8_t *fluid);

/* The plain step with FluidSimulation's epilogue stages: the face fill of
 * _advectVelocityFieldU/V/W (fluidsimulation.cpp:2597-2730: faces bordering
 * fluid keep the P2G value where the weight exceeds 1e-9, else the average
 * of their 26 neighbours; other faces 0), the snapshot, extrapolation of the
 * snapshot, the body force, extrapolation of the field
 * (MACVelocityField::extrapolateVelocityField, macvelocityfield.cpp:577-798,
 * `layers` layers; fluidsimulation.cpp:3304-3334 without the pressure
 * solve), then PIC/FLIP and RK4.  solid (ni*nj*nk bytes, nonzero = solid)
 * may be NULL.  mat_out (optional) receives the cell codes: 0 air, 1 fluid,
 * 2 solid. */
void gfo_flip_step_fs(float *pos, float *vel, int64_t n, int ni, int nj, int nk,
                      double dx, double dt, const float body_force[3], double pic_ratio,
                      int stencil, const uint8_t *solid, int layers,
                      float *u, float *v, float *w, float *u_old, float *v_old,
                      float *w_old, uint8_t *mat_out);

/* The two epilogue stages alone, on cell codes mat (0 air, 1 fluid,
 * 2 solid): the face fill of one component c (out <- ugrid, W) and the
 * extrapolation of u, v, w in place. */
void gfo_fill_component(float *out, const float *ugrid, const float *W, int c,
                        const uint8_t *mat, int ni, int nj, int nk);
void gfo_extrapolate(float *u, float *v, float *w, int ni, int nj, int nk,
                     const uint8_t *mat, int layers);

/* Particle removal: drop particles in solid cells, keep at most
 * max_per_cell per cell, chosen by the smallest gfo_removal_key(seed,
 * index) (SplitMix64) where the reference shuffles with rand()
 * (fluidsimulation.cpp:1985-1988, 3213-3243).  Particles outside the grid
 * are kept.  keep[p] receives 0/1; returns the number kept.  O(n * cell)
 * brute force: small inputs only. */
uint64_t gfo_removal_key(uint64_t seed, uint64_t i);
int64_t gfo_remove_particles(const float *pos, int64_t n, double dx, int ni, int nj, int nk,
                             const uint8_t *solid, int max_per_cell, uint64_t seed,
                             uint8_t *keep);

/* Stable counting sort of particle indices by the cell that
 * Grid3d::positionToGridIndex (grid3d.h:58-63) assigns.  Particles outside
 * [0,ni)x[0,nj)x[0,nk) get no cell and are left out of perm.
 * cell_start has ni*nj*nk+1 entries; returns the number of binned particles. */
int64_t gfo_cell_sort(const float *pos, int64_t n, double dx, int ni, int nj,
                      int nk, uint32_t *perm, uint32_t *cell_start);

/* FP64 twins. */
void gfo64_p2g(const double *pos, const double *val, int64_t n, double radius,
               const double offset[3], double dx, int ni, int nj, int nk,
               double *S, double *W);
void gfo64_g2p(const double *pos, int64_t n, const double *u, const double *v,
               const double *w, int ni, int nj, int nk, double dx, int stencil,
               double *out);
void gfo64_advect(const double *pos, int64_t n, const double *u,
                  const double *v, const double *w, int ni, int nj, int nk,
                  double dx, double dt, int order, int stencil, double *out);
void gfo64_flip_step(double *pos, double *vel, int64_t n, int ni, int nj,
                     int nk, double dx, double dt, const double body_force[3],
                     double pic_ratio, int stencil,
                     double *u, double *v, double *w,
                     double *u_old, double *v_old, double *w_old,
                     double *wu, double *wv, double *ww);

/* Wyvill kernel weight for a squared distance (scalarfield.cpp:40-46,569-571). */
double gfo_kernel_weight(double radius, double distsq);

#ifdef __cplusplus
}
#endif
#endif
