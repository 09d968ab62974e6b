/*
 * gfsx.h -- C ABI of the MI355X (gfx950) FLIP transfer + advection library.
 *
 * Conventions follow the reference's own C bindings
 * (src/c_bindings/cbindings.cpp:9-19, cbindings.h:25-131): every entry point
 * takes a trailing `int *err` that receives GFSX_SUCCESS (1) or GFSX_FAIL (0),
 * and the text of the last failure on the calling thread is returned by
 * gfsx_get_error_message().  No entry point aborts or throws.
 *
 * Layouts are the reference's: positions/velocities for the drop-in entry
 * points are AoS float triples (std::vector<vmath::vec3>, vmath.h:33-57);
 * grids are flat i-fastest float arrays (array3d.h:394-397); MAC grids have
 * u:(ni+1)*nj*nk, v:ni*(nj+1)*nk, w:ni*nj*(nk+1) entries
 * (macvelocityfield.cpp:37-45).
 *
 * Host-pointer entry points copy their inputs to the GPU, run, copy back and
 * return when the result is in host memory (the reference's calls are
 * synchronous too, clscalarfield.cpp:1388-1393).  *_device entry points take
 * device pointers, enqueue on the context's stream and return without waiting.
 */
#ifndef GFSX_H
#define GFSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFSX_SUCCESS 1
#define GFSX_FAIL 0

/* G2P stencils.  TRICUBIC is the reference's ParticleAdvector stencil
 * (tricubicinterpolate.cl:46-67, macvelocityfield.cpp:350-450); TRILINEAR is
 * MACVelocityField::evaluateVelocityAtPositionLinear (macvelocityfield.cpp:
 * 452-575). */
#define GFSX_TRILINEAR 0
#define GFSX_TRICUBIC 1

/* Output order of gfsx_flip_step. */
#define GFSX_ORDER_SORTED 0 /* particles come out grouped by brick of 4x4x4 cells */
#define GFSX_ORDER_KEEP 1   /* particle i stays at index i */

typedef struct gfsx_ctx gfsx_ctx;

/* Last error on this thread; replaces the reference's
 * CBindings_get_error_message (cbindings.cpp:11-19). */
const char *gfsx_get_error_message(void);

/* Number of visible HIP devices (0 on a machine without a GPU: *err is
 * GFSX_FAIL then, with the runtime's message). */
int gfsx_get_device_count(int *err);

/* Context: device, stream and device workspace.  Replaces
 * CLScalarField::initialize / ParticleAdvector::initialize
 * (clscalarfield.cpp:31-66, particleadvector.cpp:30-58): returns NULL and
 * GFSX_FAIL instead of aborting when no device is usable. */
gfsx_ctx *gfsx_create(int device, int *err);
void gfsx_destroy(gfsx_ctx *ctx, int *err);

/* Run on the caller's stream (a hipStream_t) from now on; NULL selects the
 * default (null) stream.  A new context runs on a stream of its own. */
void gfsx_set_stream(gfsx_ctx *ctx, void *stream, int *err);
void gfsx_synchronize(gfsx_ctx *ctx, int *err);

/* Device name / library build text, for CLScalarField::getDeviceInfo-style
 * reporting (clscalarfield.cpp:580-637). */
const char *gfsx_get_device_info(gfsx_ctx *ctx, int *err);

/* ------------------------------------------------------------------------ */
/* P2G: CLScalarField::addPointValues / addPoints.                          */
/* ------------------------------------------------------------------------ */
/* Splat n points (pos: n*3 floats) onto the node grid ni*nj*nk whose node
 * (i,j,k) sits at offset + (i,j,k)*dx, with the Wyvill kernel of the given
 * radius (scalarfield.cpp:40-46,167-201, scalarfield.cl:202-268).
 * S[node] += sum_p val[p]*K and, when W is not NULL, W[node] += sum_p K.
 * val == NULL splats the weight itself (CLScalarField::addPoints,
 * clscalarfield.cpp:68-122, scalarfield.cl:79-138).  Accumulates into S/W
 * like the reference (clscalarfield.cpp:1448-1449); the sum for each node is
 * taken in a fixed order, so results are bit-identical run to run.
 * Replaces CLScalarField::addPointValues (clscalarfield.cpp:143-301). */
void gfsx_p2g(gfsx_ctx *ctx, const float *pos, const float *val, int64_t n,
              double radius, const float offset[3], double dx, int ni, int nj,
              int nk, float *S, float *W, int *err);
void gfsx_p2g_device(gfsx_ctx *ctx, const float *pos, const float *val,
                     int64_t n, double radius, const float offset[3],
                     double dx, int ni, int nj, int nk, float *S, float *W,
                     int *err);

/* ScalarField::applyWeightField (scalarfield.cpp:90-106): S = S/W where W>0. */
void gfsx_apply_weight_device(gfsx_ctx *ctx, float *S, const float *W,
                              int64_t count, int *err);

/* ------------------------------------------------------------------------ */
/* G2P: ParticleAdvector::tricubicInterpolate.                              */
/* ------------------------------------------------------------------------ */
/* out[p] = velocity of the MAC field (u,v,w on an ni*nj*nk grid of spacing
 * dx) at pos[p]; (0,0,0) outside [0,ni*dx)x[0,nj*dx)x[0,nk*dx); NaN/Inf
 * results are zeroed (particleadvector.cpp:1139-1149).  out may alias pos
 * (the reference's in-place overload, particleadvector.cpp:457-465).
 * Replaces ParticleAdvector::tricubicInterpolate (particleadvector.cpp:
 * 401-455). */
void gfsx_g2p(gfsx_ctx *ctx, const float *pos, int64_t n, const float *u,
              const float *v, const float *w, int ni, int nj, int nk,
              double dx, int stencil, float *out, int *err);
void gfsx_g2p_device(gfsx_ctx *ctx, const float *pos, int64_t n,
                     const float *u, const float *v, const float *w, int ni,
                     int nj, int nk, double dx, int stencil, float *out,
                     int *err);

/* ------------------------------------------------------------------------ */
/* Advection: ParticleAdvector::advectParticlesRK1..RK4.                    */
/* ------------------------------------------------------------------------ */
/* out[p] = one Runge-Kutta step of the given order (1..4) from pos[p]
 * through the MAC field (particleadvector.cpp:209-399, 1045-1078).
 * Replaces ParticleAdvector::advectParticlesRK{1,2,3,4}. */
void gfsx_advect(gfsx_ctx *ctx, int order, const float *pos, int64_t n,
                 const float *u, const float *v, const float *w, int ni,
                 int nj, int nk, double dx, double dt, int stencil, float *out,
                 int *err);
void gfsx_advect_device(gfsx_ctx *ctx, int order, const float *pos, int64_t n,
                        const float *u, const float *v, const float *w, int ni,
                        int nj, int nk, double dx, double dt, int stencil,
                        float *out, int *err);

/* ------------------------------------------------------------------------ */
/* The fused plain FLIP step on device memory.                              */
/* ------------------------------------------------------------------------ */
/* One step of: fluid cells from particle positions -> P2G of u,v,w with the
 * weight field, R = dx, face offsets -> normalisation -> snapshot -> constant
 * body force on faces bordering fluid -> PIC/FLIP velocity update -> RK4
 * (the FluidSimulation stages at fluidsimulation.cpp:2010-2017, 2526-2595,
 * 3306, 2766-2805, 3104-3129, 3181-3209).  Particles are structure-of-arrays
 * device buffers; the *_in arrays are read, the *_out arrays written (they
 * must not alias the inputs).  With GFSX_ORDER_SORTED the outputs come out
 * grouped by brick of 4x4x4 cells, in the order of a stable sort (the
 * reference reorders its particles every step as well,
 * fluidsimulation.cpp:3211-3223); id_in/id_out (optional) carry a particle
 * label through that permutation.  u,v,w receive the grid after the body
 * force; the snapshot (u_old..) and the weights (wu..) are written when not
 * NULL. */
typedef struct gfsx_scene {
    int ni, nj, nk;
    double dx;
    int64_t n;
    const float *x_in, *y_in, *z_in;
    const float *vx_in, *vy_in, *vz_in;
    float *x_out, *y_out, *z_out;
    float *vx_out, *vy_out, *vz_out;
    const uint32_t *id_in; /* optional */
    uint32_t *id_out;      /* optional */
    float *u, *v, *w;
    float *u_old, *v_old, *w_old; /* optional */
    float *wu, *wv, *ww;          /* optional */
    const uint8_t *solid;         /* optional: ni*nj*nk cells, nonzero = solid
                                     (GFSX_EPILOGUE_FLUIDSIM only) */
} gfsx_scene;

typedef struct gfsx_step_params {
    double dt;
    float body_force[3]; /* e.g. (0, -9.8, 0) */
    double pic_ratio;    /* 0.05 in the reference (fluidsimulation.h:1161) */
    int stencil;         /* GFSX_TRILINEAR or GFSX_TRICUBIC */
    int order;           /* GFSX_ORDER_SORTED or GFSX_ORDER_KEEP */
    int epilogue;        /* GFSX_EPILOGUE_NONE or GFSX_EPILOGUE_FLUIDSIM */
    int extrapolation_layers; /* GFSX_EPILOGUE_FLUIDSIM: layers (0 = 7) */
} gfsx_step_params;

/* Epilogues of the P2G stage.  NONE: the transfer alone (normalised values,
 * snapshot, body force on faces bordering fluid).  FLUIDSIM: the stages
 * FluidSimulation runs around it, without the pressure solve: the face fill
 * of _advectVelocityFieldU/V/W (fluidsimulation.cpp:2597-2730; faces
 * bordering fluid keep the P2G value where the weight exceeds 1e-9, else the
 * average of their 26 neighbours; other faces 0), the snapshot and its
 * extrapolation, the body force, and the extrapolation of the field
 * (MACVelocityField::extrapolateVelocityField, macvelocityfield.cpp:577-798,
 * with the scene's solid cells; fluidsimulation.cpp:3304-3334).  One scene
 * per call (not batched); u_old.. receive the extrapolated snapshot, wu..
 * the P2G weights. */
#define GFSX_EPILOGUE_NONE 0
#define GFSX_EPILOGUE_FLUIDSIM 1

void gfsx_flip_step(gfsx_ctx *ctx, const gfsx_scene *scene,
                    const gfsx_step_params *params, int *err);

/* The same step for nscenes independent scenes (each with its own grid,
 * particles and parameters; one stencil and one output order for all) in
 * one launch per kernel: the sort, P2G and G2P kernels take every scene's
 * cells, bricks and particles together.  Each scene's results are
 * bit-identical to gfsx_flip_step on that scene alone.  Scenes with a NULL
 * u_old/v_old/w_old share one workspace for their snapshots.  The call
 * enqueues work only (no allocation) once the same scenes have run once, so
 * it can be captured into a HIP graph on the context's stream.  Batching of
 * GridFluidSim3D's step has no counterpart in the reference; it serves many
 * small simulations on one device. */
void gfsx_flip_step_batched(gfsx_ctx *ctx, const gfsx_scene *scenes, int nscenes,
                            const gfsx_step_params *params, int *err);

/* Bytes of the static device scratch of gfsx_flip_step_noalloc (64 MiB,
 * held on every device the library runs on until the process ends). */
#define GFSX_NOALLOC_SCRATCH_BYTES 67108864

/* gfsx_flip_step with no allocation: the same kernels and the same bits,
 * every workspace buffer (sort, P2G tiles, staged particles, the snapshot
 * when u_old.. are NULL, the scene table) carved from a static scratch of
 * GFSX_NOALLOC_SCRATCH_BYTES per device; no hipMalloc, hipFree or
 * allocating library call, no host-to-device copy, no timing events.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Limits: one scene, p2g_variant 4, no
 * epilogue; a step whose workspace exceeds the scratch (about 60 bytes per
 * particle plus 80 per cell) fails with a message giving the size it needs,
 * before any work is enqueued.  Every call on a device shares that device's
 * scratch and leaves its contents behind: calls that overlap on one device
 * (other streams, contexts or threads) race. */
void gfsx_flip_step_noalloc(gfsx_ctx *ctx, const gfsx_scene *scene,
                            const gfsx_step_params *params, int *err);

/* Tunables of the fused step, by name (an unknown name or value fails).
 * Every P2G and G2P variant is deterministic; they differ in speed, and the
 * P2G variants in the rounding of the sums (within FP32 rounding of each
 * other).
 *   "p2g_variant": 0 = brick-staged LDS gather, candidates streamed in
 *                      chunks,
 *                  1 = direct gather from global memory,
 *                  2 = brick-staged, candidates range by range,
 *                  3 = brick-staged, candidates as one flattened stream,
 *                  4 = scatter of each brick's own particles into a tile of
 *                      fixed-point sums in LDS, then a pass adding the
 *                      tiles that hold each face (default).
 *   "g2p_variant": 0 = grid boxes staged in LDS per brick (default),
 *                  1 = direct sampling from global memory.
 *   "g2p_brick":   cells per edge of the G2P brick of g2p_variant 0 with
 *                  p2g_variant 4: 8 (default) or 4.
 *   "particle_gather" (p2g_variant 4, g2p_variant 0, GFSX_ORDER_SORTED):
 *                  0 = the sort writes sorted copies of the particles,
 *                      P2G and G2P read them,
 *                  1 = the sort writes only the permutation; P2G and G2P
 *                      read each sort slot's particle from the *_in arrays
 *                      through it (when the *_in arrays are the previous
 *                      step's sorted output the reads are nearly
 *                      consecutive),
 *                  2 = as 1, but P2G writes the sorted copies for G2P
 *                      (default).
 *                  Results are the same bits in every mode.
 *   "fp32_cells":  1 = when 1/dx is a power of two (in every scene of the
 *                      call), the brick G2P computes cell indices and
 *                      fractions in FP32, where x/dx is exact: the same
 *                      bits as FP64 (default),
 *                  0 = always FP64.
 *   "adjoint_variant" (gfsx_flip_step_backward; results equal up to FP32
 *                  rounding):
 *                  0 = bricks of sorted particles sum the grid adjoints in
 *                      LDS in 64-bit fixed point, then add them once per
 *                      face (default; deterministic),
 *                  1 = one thread per particle, every grid adjoint added
 *                      to global memory with float atomics. */
void gfsx_set_option(gfsx_ctx *ctx, const char *name, int value, int *err);
int gfsx_get_option(gfsx_ctx *ctx, const char *name, int *err);

/* Per-stage timing of gfsx_flip_step with HIP events recorded on the
 * context's stream.  After gfsx_timing_begin every step records four events:
 * before the cell sort, after the sort, after P2G and after the
 * G2P/PIC-FLIP/RK4 kernel.  gfsx_timing_end waits for the stream, writes the
 * summed milliseconds of the three stages to stage_ms[0..2] and of whole
 * steps to stage_ms[3], stops recording and returns the number of steps
 * timed.  (The reference only has wall-clock StopWatch timers per stage,
 * fluidsimulation.cpp:3273-3389.) */
void gfsx_timing_begin(gfsx_ctx *ctx, int *err);
int gfsx_timing_end(gfsx_ctx *ctx, double stage_ms[4], int *err);

/* ------------------------------------------------------------------------ */
/* Adjoints (differentiable form).                                          */
/* ------------------------------------------------------------------------ */
/* Adjoint of gfsx_p2g_device: given dL/dS (gS) and/or dL/dW (gW) on the
 * node grid, writes dL/dpos (gpos, n*3 floats) and, when gval is not NULL,
 * dL/dval (gval, n floats) for S = sum K(|node - pos|^2) val and W = sum K.
 * One thread per point gathers over the nodes in its kernel radius (no
 * atomics; bit-identical run to run).  The reference has no adjoint; this
 * differentiates the splat of CLScalarField::addPointValues / addPoints
 * (clscalarfield.cpp:68-122,143-301, scalarfield.cpp:167-201). */
void gfsx_p2g_backward_device(gfsx_ctx *ctx, const float *pos, const float *val,
                              int64_t n, double radius, const float offset[3],
                              double dx, int ni, int nj, int nk, const float *gS,
                              const float *gW, float *gpos, float *gval, int *err);

/* Adjoints of one step's outputs (d(loss)/d(x_out..), d(loss)/d(vx_out..))
 * and, written by the call, of its inputs.  SoA device arrays of n floats;
 * the written arrays must not alias the scene's inputs (they may alias the
 * output adjoints). */
typedef struct gfsx_scene_grad {
    const float *gx_out, *gy_out, *gz_out;
    const float *gvx_out, *gvy_out, *gvz_out;
    float *gx_in, *gy_in, *gz_in;
    float *gvx_in, *gvy_in, *gvz_in;
} gfsx_scene_grad;

/* Adjoint of gfsx_flip_step for the trilinear stencil with
 * GFSX_ORDER_KEEP (output particle p is input particle p).  Reads the
 * step's inputs from scene (x_in.., vx_in..; the scene's output and grid
 * fields are not used), recomputes the forward grids (P2G, weights,
 * snapshot, body force) in the context's workspace, then back-propagates
 * through RK4 (stage positions recomputed from x_in), the PIC/FLIP blend,
 * the trilinear samples and the P2G normalisation (quotient rule:
 * dL/dS = dL/dn / W, dL/dW = -dL/dn * n / W).  The fluid mask and cell
 * indices are piecewise constant, so they contribute nothing.  By default
 * the grid adjoints are summed as 64-bit fixed-point integers, so results
 * are bit-identical run to run (adjoint_variant 1 adds floats atomically and
 * is not).  The reference has no adjoint; this differentiates the step
 * gfsx_flip_step replaces (fluidsimulation.cpp:2526-2595, 3104-3129,
 * 3181-3209; particleadvector.cpp:1045-1054). */
void gfsx_flip_step_backward(gfsx_ctx *ctx, const gfsx_scene *scene,
                             const gfsx_step_params *params,
                             const gfsx_scene_grad *grad, int *err);

/* What a forward step saves for its backward pass, so that the backward
 * need not recompute the brick sort and P2G: the sort of the step's input
 * positions.  Device arrays the caller allocates: perm and xs, ys, zs of n
 * entries, start of gfsx_step_tape_starts(ni, nj, nk) entries.  The caller
 * sets n and starts; gfsx_flip_step_record fills the arrays and gbx, gby. */
typedef struct gfsx_step_tape {
    uint32_t *perm;       /* sort slot -> input index */
    uint32_t *start;      /* first sort slot of each brick bin, then the end */
    float *xs, *ys, *zs;  /* input positions in sort order */
    int64_t n;
    int64_t starts;       /* entries of start */
    int32_t gbx, gby;     /* bin grid of 8^3-cell groups along x and y */
} gfsx_step_tape;

/* Entries of gfsx_step_tape.start for an ni x nj x nk grid. */
int64_t gfsx_step_tape_starts(int ni, int nj, int nk, int *err);

/* gfsx_flip_step (trilinear, GFSX_ORDER_KEEP, no epilogue: the step
 * gfsx_flip_step_backward differentiates) that also fills tape.  Every grid
 * field of scene (u, v, w, u_old, v_old, w_old, wu, wv, ww) is required: the
 * taped backward reads them.  Outputs and grids are bit-identical to
 * gfsx_flip_step with the default P2G variant.  No counterpart in the
 * reference, which has no adjoint; for gradient callers that keep each
 * step's state (autograd rollouts). */
void gfsx_flip_step_record(gfsx_ctx *ctx, const gfsx_scene *scene,
                           const gfsx_step_params *params, gfsx_step_tape *tape, int *err);

/* gfsx_flip_step_backward from a step recorded by gfsx_flip_step_record:
 * scene holds that step's inputs and grids (unchanged since), tape its sort.
 * Results are bit-identical to gfsx_flip_step_backward on the same inputs;
 * the forward's sort and P2G are not recomputed. */
void gfsx_flip_step_backward_tape(gfsx_ctx *ctx, const gfsx_scene *scene,
                                  const gfsx_step_params *params,
                                  const gfsx_scene_grad *grad, const gfsx_step_tape *tape,
                                  int *err);

/* Particle removal after a step: FluidSimulation::
 * _removeMarkerParticlesInSolidCells and _removeMarkerParticles
 * (fluidsimulation.cpp:1985-1988, 3213-3243; the reference's cap is 100,
 * fluidsimulation.h:1167).  Reads n particles from scene's *_in arrays,
 * drops those in solid cells (scene->solid, optional) and keeps at most
 * max_per_cell per cell, writing the kept ones, in input order, to the
 * *_out arrays (and id_out: id_in[p], or p when id_in is NULL).  Where the
 * reference keeps the first particles of a rand() shuffle, a cell here
 * keeps the max_per_cell particles with the smallest SplitMix64 key of
 * (seed, input index): the same result on every run.  Particles outside
 * the grid are kept.  The scene's grid fields are not used.  Returns the
 * number of particles kept (and waits for the result). */
int64_t gfsx_remove_particles(gfsx_ctx *ctx, const gfsx_scene *scene, int max_per_cell,
                              uint64_t seed, int *err);

/* Stable counting sort of particle indices by cell (the binning the P2G
 * gather uses), exposed for testing: perm[n] receives the indices ordered by
 * cell, then by index; cell_start[ni*nj*nk+1] the cell offsets.  Particles
 * outside the grid are placed after cell_start[ni*nj*nk], in index order.
 * pos is SoA (x, y, z device arrays). */
void gfsx_cell_sort_device(gfsx_ctx *ctx, const float *x, const float *y,
                           const float *z, int64_t n, double dx, int ni,
                           int nj, int nk, uint32_t *perm,
                           uint32_t *cell_start, int *err);

#ifdef __cplusplus
}
#endif
#endif /* GFSX_H */
