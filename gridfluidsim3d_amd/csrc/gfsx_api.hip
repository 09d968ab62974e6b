// gfsx_api.hip -- the C ABI declared in include/gfsx.h.
//
// Error convention of the reference's C bindings (cbindings.cpp:9-19): a
// trailing int* err receives GFSX_SUCCESS/GFSX_FAIL and the message of the
// last failure is kept per thread.  Nothing here aborts: HIP errors and bad
// arguments become GFSX_FAIL with a message, where the reference would call
// FLUIDSIM_ASSERT (clscalarfield.cpp:519-524, particleadvector.cpp:467-472).

#include "gfsx.h"
#include "gfsx_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>

using gfsx::Error;

// Size of the static scratch of gfsx_flip_step_noalloc (GFSX_NOALLOC_SCRATCH_BYTES).
constexpr size_t kNoallocScratchBytes = (size_t)GFSX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

thread_local std::string g_error = "";

void fail(int *err, const std::string &msg) {
    g_error = msg;
    if (err) *err = GFSX_FAIL;
}

void ok(int *err) {
    if (err) *err = GFSX_SUCCESS;
}

enum Slot {
    S_COUNT, S_START, S_KEY, S_SLOT, S_PERM, S_PERM_TMP, S_HRANK, S_SOCT, S_BFLAG, S_BIG, S_BIGN,
    S_BOUT, S_BOUTS, S_SCAN, S_OCT, S_POCT,
    S_OLD0, S_XS0, S_XS1, S_XS2, S_VS0, S_VS1, S_VS2, S_DENSE, S_DENSEN,
    S_D_POS, S_D_VAL, S_D_S, S_D_W, S_D_U, S_D_V, S_D_WW, S_D_OUT,
    S_B_GRID, S_B_OLD, S_B_WGT, S_B_LG, S_B_XO, S_B_VO, S_B_FX, S_B_MAX,
    S_E_N, S_E_W, S_E_LAYER, S_E_LIST, S_R_KEEP, S_R_DST, S_R_SCAN, S_R_N, S_R_FULL, S_R_NFULL,
    S_P_TILE, S_P_TFLAG,
    S_K_COUNT, S_K_START, S_K_PSTART, S_K_RSTART, S_K_RUNREC, S_K_NRUNS, S_K_LIST, S_K_LIST2,
    S_K_RUNBASE, S_K_MID, S_K_HUGE, S_K_LCOUNT, S_K_SCAN, S_K_FLUID, S_K_SSLOT,
    S_K_SCOUNT, S_K_JR, S_K_RLEN, S_K_NCHUNK,
    NSLOT
};

void require(bool cond, const char *msg) {
    if (!cond) throw Error(msg);
}

}  // namespace

struct gfsx_ctx {
    int device = -1;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string info;
    int num_cu = 1;
    void *buf[NSLOT] = {};
    size_t bytes[NSLOT] = {};
    int p2g_variant = 4;
    int g2p_variant = 0;
    int g2p_brick = 8;
    int particle_gather = 2;
    int fp32_cells = 1;
    int adjoint_variant = 0;
    bool timing = false;
    std::vector<std::array<hipEvent_t, 4>> events;
    size_t events_used = 0;

    hipEvent_t *next_events() {
        if (events_used == events.size()) {
            std::array<hipEvent_t, 4> e{};
            for (auto &x : e) GFSX_HIP(hipEventCreate(&x));
            events.push_back(e);
        }
        return events[events_used++].data();
    }

    // Device workspace, grown on demand and kept until gfsx_destroy.
    void *get(int slot, size_t nbytes) {
        nbytes = std::max<size_t>(nbytes, 256);
        if (bytes[slot] < nbytes) {
            if (buf[slot]) GFSX_HIP(hipFree(buf[slot]));
            buf[slot] = nullptr;
            bytes[slot] = 0;
            GFSX_HIP(hipMalloc(&buf[slot], nbytes));
            bytes[slot] = nbytes;
        }
        return buf[slot];
    }

    uint32_t *u32(int slot, int64_t count) {
        return (uint32_t *)get(slot, sizeof(uint32_t) * (size_t)std::max<int64_t>(count, 1));
    }
    float *f32(int slot, int64_t count) {
        return (float *)get(slot, sizeof(float) * (size_t)std::max<int64_t>(count, 1));
    }

    // Scene tables on the device, kept by content: a simulation loop
    // alternates between a few tables (its buffers swap every step), so
    // after the first steps nothing is uploaded, and a step captured into a
    // graph holds no copy.
    // Each device table holds the n descriptors followed by the n combine
    // wave offsets (r0) as a dense array, which the combine searches.
    struct Table {
        std::vector<gfsx::SceneDesc> host;
        std::vector<int64_t> r0;
        gfsx::SceneDesc *dev = nullptr;
        size_t cap = 0;
        uint64_t used = 0;
        hipStream_t stream = nullptr;   // the stream of the last work that read it
    };
    // Tables of at most this many scenes are written by kernels (the
    // entries travel as kernel arguments), ordered after every earlier
    // reader on the stream: replacing one needs no host wait, so a loop
    // whose buffers change every step (an autograd rollout) keeps queueing.
    static constexpr size_t kPutScenesMax = 16;
    std::array<Table, 4> tables;
    uint64_t table_clock = 0;

    const gfsx::SceneDesc *table(const std::vector<gfsx::SceneDesc> &t) {
        const size_t bytes = sizeof(gfsx::SceneDesc) * t.size();
        for (auto &e : tables) {
            if (e.dev && e.host.size() == t.size() &&
                std::memcmp(e.host.data(), t.data(), bytes) == 0) {
                e.used = ++table_clock;
                e.stream = stream;
                return e.dev;
            }
        }
        hipStreamCaptureStatus cap_status = hipStreamCaptureStatusNone;
        GFSX_HIP(hipStreamIsCapturing(stream, &cap_status));
        require(cap_status == hipStreamCaptureStatusNone,
                "gfsx: the scenes of a captured step must have run once before capture "
                "(their table is uploaded then)");
        Table *slot = &tables[0];
        for (auto &e : tables)
            if (e.used < slot->used) slot = &e;
        if (slot->cap < t.size()) {
            // the old table may still be read by queued work
            GFSX_HIP(hipStreamSynchronize(stream));
            if (slot->used && slot->stream != stream) GFSX_HIP(hipStreamSynchronize(slot->stream));
            if (slot->dev) GFSX_HIP(hipFree(slot->dev));
            slot->dev = nullptr;
            slot->cap = 0;
            GFSX_HIP(hipMalloc(&slot->dev, std::max<size_t>(bytes + sizeof(int64_t) * t.size(), 1)));
            slot->cap = t.size();
        } else if (slot->used) {
            // queued work may still read the table being replaced: work on
            // this stream reads it before kernels written after it run;
            // anything else is waited for
            if (t.size() > kPutScenesMax || slot->stream != stream)
                GFSX_HIP(hipStreamSynchronize(slot->stream));
        }
        slot->host = t;
        slot->r0.resize(t.size());
        for (size_t i = 0; i < t.size(); i++) slot->r0[i] = t[i].r0;
        if (t.size() <= kPutScenesMax) {
            gfsx::put_scenes(slot->dev, slot->host.data(), (int)t.size(), stream,
                             reinterpret_cast<int64_t *>(slot->dev + t.size()));
        } else {
            GFSX_HIP(hipMemcpyAsync(slot->dev, slot->host.data(), bytes, hipMemcpyHostToDevice, stream));
            GFSX_HIP(hipMemcpyAsync(slot->dev + t.size(), slot->r0.data(), sizeof(int64_t) * t.size(),
                                    hipMemcpyHostToDevice, stream));
        }
        slot->used = ++table_clock;
        slot->stream = stream;
        return slot->dev;
    }

    gfsx::SortBuffers sort_buffers(const std::vector<gfsx::SceneDesc> &t, int64_t n,
                                   uint32_t nbin_total, uint32_t nblk) {
        gfsx::SortBuffers sb;
        sb.scenes = table(t);
        sb.nscene = (int)t.size();
        sb.nbin_total = nbin_total;
        sb.nblk_total = nblk;
        const int64_t nb = nblk;
        const uint32_t nbin = nbin_total;
        sb.count = u32(S_COUNT, (int64_t)nbin);
        sb.start = u32(S_START, (int64_t)nbin + 1);
        sb.key = u32(S_KEY, n);
        sb.slot = u32(S_SLOT, n);
        sb.perm = u32(S_PERM, n);
        sb.perm_tmp = u32(S_PERM_TMP, n);
        sb.hrank = u32(S_HRANK, n);
        sb.soct = (uint8_t *)get(S_SOCT, (size_t)std::max<int64_t>(n, 1));
        sb.brick_flag = nullptr;
        // every big bin holds more than kSmallBin particles
        sb.big_cap = (uint32_t)std::min<int64_t>((int64_t)nbin, n / 16 + 1);
        sb.big_list = u32(S_BIG, sb.big_cap);
        sb.big_count = u32(S_BIGN, 2);
        sb.block_out = u32(S_BOUT, nb);
        sb.block_out_scan = u32(S_BOUTS, nb + 1);
        int64_t words = std::max(gfsx::scan_scratch_words((int64_t)nbin),
                                 gfsx::scan_scratch_words(nb));
        sb.scan_scratch = u32(S_SCAN, words);
        sb.oct = (uint8_t *)get(S_OCT, (size_t)std::max<int64_t>(n, 1));
        sb.poct = (uint8_t *)get(S_POCT, (size_t)std::max<int64_t>(n, 1));
        return sb;
    }

    // Teardown cannot report failures; the runtime's status is discarded.
    ~gfsx_ctx() {
        (void)hipSetDevice(device);
        for (int s = 0; s < NSLOT; s++)
            if (buf[s]) (void)hipFree(buf[s]);
        for (auto &e : events)
            for (auto &x : e) (void)hipEventDestroy(x);
        for (auto &t : tables)
            if (t.dev) (void)hipFree(t.dev);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

// The static scratch of gfsx_flip_step_noalloc: every device buffer of that
// entry point's step (sort, tiles, staged particles, snapshot, scene table)
// is carved from it.  Each device holds its own copy from the load of the
// library to the end of the process.  In host code `scratch` only names the
// symbol: its device address comes from hipGetSymbolAddress, once per device
// (with that device current) in gfsx_create, and is kept in g_scratch_dev.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device memory handed out front to back from one region (the static
// scratch of gfsx_flip_step_noalloc).  `used` counts what a layout needs even
// past the end, so the caller can refuse it, with the full size, before
// anything is written.
struct Carve {
    char *base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    void *take(size_t bytes) {
        used = (used + 255) & ~(size_t)255;
        void *p = base + used;
        used += std::max<size_t>(bytes, 1);
        return p;
    }
};

// Workspace of the brick sort for a step of n particles in nblk binning
// blocks over nbin brick bins (overflow bins included); ws(slot, bytes)
// hands out device memory.
template <class WS>
gfsx::BrickSort brick_sort(WS &&ws, int64_t n, int64_t nblk, uint32_t nbin, bool want_perm) {
    gfsx::BrickSort b;
    const int64_t nwave = nblk * 16;   // a binning block is 4 x 256 particles
    const size_t u64 = sizeof(unsigned long long), u32 = sizeof(uint32_t);
    const size_t nn = (size_t)std::max<int64_t>(n, 1), nw = (size_t)std::max<int64_t>(nwave, 1);
    b.nwave = nwave;
    b.nbb_total = nbin;
    b.count = (unsigned long long *)ws(S_K_COUNT, u64 * nbin);
    b.start = (unsigned long long *)ws(S_K_START, u64 * ((size_t)nbin + 1));
    b.pstart = (uint32_t *)ws(S_K_PSTART, u32 * ((size_t)nbin + 1));
    b.rstart = (uint32_t *)ws(S_K_RSTART, u32 * ((size_t)nbin + 1));
    b.runrec = (uint2 *)ws(S_K_RUNREC, sizeof(uint2) * nw * 64);
    b.nruns = (uint8_t *)ws(S_K_NRUNS, nw);
    b.jr = (uint32_t *)ws(S_K_JR, u32 * nw * 64);
    b.rlen = (uint32_t *)ws(S_K_RLEN, u32 * nw * 64);
    b.nchunk = (uint32_t *)ws(S_K_NCHUNK, u32 * (size_t)std::max<int64_t>(nblk, 1));
    b.list = (uint32_t *)ws(S_K_LIST, u32 * nn);
    b.list2 = (uint32_t *)ws(S_K_LIST2, u32 * nn);
    b.runbase = (uint32_t *)ws(S_K_RUNBASE, u32 * nw * 64);
    b.mid = (uint32_t *)ws(S_K_MID, u32 * nbin);
    b.huge = (uint32_t *)ws(S_K_HUGE, u32 * nbin);
    b.lcount = (uint32_t *)ws(S_K_LCOUNT, u32 * 2);
    b.scan_scratch = (unsigned long long *)ws(
        S_K_SCAN, u64 * (size_t)std::max<int64_t>(gfsx::scan64_scratch_words(nbin), 1));
    b.perm = want_perm ? (uint32_t *)ws(S_PERM, u32 * nn) : nullptr;
    b.fluid = (unsigned long long *)ws(S_K_FLUID, u64 * nbin);
    b.slow_slots = (uint32_t *)ws(S_K_SSLOT, u32 * nn);
    b.slow_count = (uint32_t *)ws(S_K_SCOUNT, u32);
    return b;
}

template <class F>
void guarded(gfsx_ctx *ctx, int *err, F &&f) {
    try {
        if (!ctx) throw Error("gfsx: null context");
        GFSX_HIP(hipSetDevice(ctx->device));
        f();
        ok(err);
    } catch (const std::exception &e) {
        fail(err, e.what());
    } catch (...) {
        fail(err, "gfsx: unknown error");
    }
}

void check_grid(int ni, int nj, int nk, double dx) {
    require(ni > 0 && nj > 0 && nk > 0, "gfsx: grid dimensions must be positive");
    require(std::isfinite(dx) && dx > 0.0, "gfsx: dx must be positive and finite");
    int64_t faces = (int64_t)(ni + 1) * (nj + 1) * (nk + 1);
    require(faces < (int64_t)1 << 31, "gfsx: grid too large (more than 2^31 entries per array)");
}

void check_particles(int64_t n, const void *p) {
    require(n >= 0, "gfsx: negative particle count");
    require(n < ((int64_t)1 << 32) - 1, "gfsx: more than 2^32-2 particles in one call");
    require(n == 0 || p != nullptr, "gfsx: null particle array");
}

gfsx::MacGrid mac_grid(const float *u, const float *v, const float *w, int ni,
                       int nj, int nk, double dx) {
    gfsx::MacGrid g;
    g.f[0] = u;
    g.f[1] = v;
    g.f[2] = w;
    g.ni = ni;
    g.nj = nj;
    g.nk = nk;
    g.dx = dx;
    g.inv_dx = 1.0 / dx;
    gfsx::set_limits(g);
    return g;
}

int64_t mac_count(int c, int ni, int nj, int nk) {
    return (int64_t)(ni + (c == 0)) * (nj + (c == 1)) * (nk + (c == 2));
}

// P2G on device arrays: bin by node cell in the frame shifted by `offset`,
// then gather.  Bins cover every node cell a contributing point can lie in:
// node n gathers cells [n + floor(-R/dx), n + ceil(R/dx) - 1].
void p2g_device_impl(gfsx_ctx *c, const float *pos, const float *val,
                     int64_t n, double radius, const float offset[3],
                     double dx, int ni, int nj, int nk, float *S, float *W) {
    check_grid(ni, nj, nk, dx);
    check_particles(n, pos);
    require(std::isfinite(radius) && radius > 0.0, "gfsx: radius must be positive and finite");
    require(S != nullptr, "gfsx: null scalar field");
    require(offset != nullptr, "gfsx: null offset");
    const double r = radius / dx;
    require(r <= 64.0, "gfsx: radius larger than 64 cells");
    const int pad = (int)(-std::floor(-r));
    const int up = (int)std::ceil(r) - 1;
    gfsx::BinGrid bg;
    gfsx::NodeGrid ng;
    const int dims[3] = {ni, nj, nk};
    int64_t nbin = 1;
    for (int a = 0; a < 3; a++) {
        bg.b[a] = dims[a] + pad + up;
        bg.pad[a] = pad;
        bg.off[a] = offset[a];
        ng.n[a] = dims[a];
        ng.off[a] = offset[a];
        ng.dlo[a] = (int)std::floor(pad - r);
        ng.dhi[a] = (int)std::ceil(pad + r) - 1;
        nbin *= bg.b[a];
    }
    require(nbin < ((int64_t)1 << 32) - 2, "gfsx: too many bins");
    bg.inv_dx = 1.0 / dx;
    bg.nbin = (uint32_t)nbin;
    ng.dx_f = (float)dx;
    std::vector<gfsx::SceneDesc> t(1);
    std::memset(t.data(), 0, sizeof(gfsx::SceneDesc));
    t[0].aos = pos;
    t[0].n = n;
    t[0].bg = bg;
    gfsx::SortBuffers sb = c->sort_buffers(t, n, bg.nbin + 1, (uint32_t)gfsx::sort_blocks(n));
    gfsx::cell_sort(sb, n, c->stream, false);
    gfsx::p2g_gather(pos, val, n, bg, ng, radius, dx, sb, S, W, c->stream);
}

// Checks one scene of a fused step and fills its table entry.
gfsx::SceneDesc scene_desc(const gfsx_scene *sc, const gfsx_step_params *pr, const char *who) {
    auto req = [&](bool c, const char *what) {
        if (!c) throw gfsx::Error(std::string(who) + ": " + what);
    };
    check_grid(sc->ni, sc->nj, sc->nk, sc->dx);
    check_particles(sc->n, sc->x_in);
    req(pr->stencil == GFSX_TRILINEAR || pr->stencil == GFSX_TRICUBIC, "unknown stencil");
    req(pr->order == GFSX_ORDER_SORTED || pr->order == GFSX_ORDER_KEEP, "unknown output order");
    req(std::isfinite(pr->dt), "dt must be finite");
    req(pr->epilogue == GFSX_EPILOGUE_NONE || pr->epilogue == GFSX_EPILOGUE_FLUIDSIM,
        "unknown epilogue");
    req(pr->extrapolation_layers >= 0 && pr->extrapolation_layers <= 100,
        "extrapolation_layers must be 0..100");
    req(!sc->solid || pr->epilogue == GFSX_EPILOGUE_FLUIDSIM,
        "solid cells are used by the FluidSimulation epilogue only");
    req(sc->u && sc->v && sc->w, "null grid output");
    if (sc->n > 0) {
        req(sc->x_in && sc->y_in && sc->z_in && sc->vx_in && sc->vy_in && sc->vz_in,
            "null particle input");
        req(sc->x_out && sc->y_out && sc->z_out && sc->vx_out && sc->vy_out && sc->vz_out,
            "null particle output");
        const void *ins[6] = {sc->x_in, sc->y_in, sc->z_in, sc->vx_in, sc->vy_in, sc->vz_in};
        const void *outs[6] = {sc->x_out, sc->y_out, sc->z_out, sc->vx_out, sc->vy_out, sc->vz_out};
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) req(ins[a] != outs[b], "outputs must not alias inputs");
    }
    const int64_t nbin = (int64_t)(sc->ni + 2) * (sc->nj + 2) * (sc->nk + 2);
    req(nbin < ((int64_t)1 << 31), "too many cells");
    gfsx::SceneDesc d;
    std::memset(&d, 0, sizeof d);   // padding too: tables are compared bytewise
    d.n = sc->n;
    d.xin[0] = sc->x_in;
    d.xin[1] = sc->y_in;
    d.xin[2] = sc->z_in;
    d.vin[0] = sc->vx_in;
    d.vin[1] = sc->vy_in;
    d.vin[2] = sc->vz_in;
    d.xout[0] = sc->x_out;
    d.xout[1] = sc->y_out;
    d.xout[2] = sc->z_out;
    d.vout[0] = sc->vx_out;
    d.vout[1] = sc->vy_out;
    d.vout[2] = sc->vz_out;
    d.id_in = sc->id_in;
    d.id_out = sc->id_out;
    d.ni = sc->ni;
    d.nj = sc->nj;
    d.nk = sc->nk;
    d.dx = sc->dx;
    float *grids[3] = {sc->u, sc->v, sc->w};
    float *olds[3] = {sc->u_old, sc->v_old, sc->w_old};
    float *wts[3] = {sc->wu, sc->wv, sc->ww};
    for (int c = 0; c < 3; c++) {
        d.grid[c] = grids[c];
        d.grid_old[c] = olds[c];
        d.wgt[c] = wts[c];
        d.grav_on[c] = std::fabs((double)pr->body_force[c]) > 0.0;
        d.grav_delta[c] = (float)((double)pr->body_force[c] * pr->dt);
    }
    d.h_half = (float)(0.5 * pr->dt);
    d.h_full = (float)pr->dt;
    d.h_sixth = (float)(pr->dt / 6.0f);
    d.r_pic = (float)pr->pic_ratio;
    d.r_flip = (float)(1 - pr->pic_ratio);
    return d;
}

// Runs the step; returns the scenes as laid out (bin grids, offsets).  With
// carve, every device buffer (and the scene table) comes from that region
// and nothing is allocated: only the brick pipeline, one scene, no
// epilogue and no timing events.
std::vector<gfsx::SceneDesc> flip_step_impl(gfsx_ctx *ctx, const gfsx_scene *scenes,
                                            const gfsx_step_params *params, int nscenes,
                                            const char *who, bool grids_only = false,
                                            Carve *carve = nullptr) {
    auto ws = [&](int slot, size_t bytes) -> void * {
        return carve ? carve->take(bytes) : ctx->get(slot, bytes);
    };
    auto wsf = [&](int slot, int64_t count) -> float * {
        return (float *)ws(slot, sizeof(float) * (size_t)std::max<int64_t>(count, 1));
    };
    gfsx::SceneDesc *table_dev = nullptr;
    if (carve) {
        require(nscenes == 1 && ctx->p2g_variant == 4 && params[0].epilogue == GFSX_EPILOGUE_NONE,
                "the no-allocation step takes one scene, p2g_variant 4 and no epilogue");
        table_dev = (gfsx::SceneDesc *)carve->take(sizeof(gfsx::SceneDesc));
    }
    gfsx::StepBatch s;
    s.scenes.resize(nscenes);
    for (int i = 0; i < nscenes; i++) {
        s.scenes[i] = scene_desc(&scenes[i], &params[i], who);
        if (params[i].stencil != params[0].stencil || params[i].order != params[0].order)
            throw gfsx::Error(std::string(who) + ": every scene needs the same stencil and order");
    }
    const bool fs = params[0].epilogue == GFSX_EPILOGUE_FLUIDSIM;
    if (fs && nscenes != 1)
        throw gfsx::Error(std::string(who) + ": the FluidSimulation epilogue takes one scene");
    gfsx::layout_scenes(s.scenes.data(), nscenes, s.tot);
    require(s.tot.n < ((int64_t)1 << 32) - 1, "gfsx: too many particles in one step");
    require(s.tot.nbin < ((int64_t)1 << 32) - 2, "gfsx: too many cells in one step");
    require(s.tot.nbb < ((int64_t)1 << gfsx::kBrickBinBits), "gfsx: too many bricks in one step");
    require(s.tot.nrm < ((int64_t)1 << 32), "gfsx: too many bricks in one step");
    require(s.tot.nblk * 1024 < ((int64_t)1 << 32) - 1024, "gfsx: too many particles in one step");
    require(nscenes == 1 || (ctx->p2g_variant != 1 && ctx->g2p_variant != 1),
            "gfsx: the direct P2G and G2P variants take one scene");
    // snapshots the caller does not keep live in one workspace buffer
    int64_t old_faces = 0;
    for (auto &d : s.scenes)
        for (int c = 0; c < 3; c++)
            if (!d.grid_old[c]) old_faces += mac_count(c, d.ni, d.nj, d.nk);
    float *old_ws = wsf(S_OLD0, old_faces);
    for (auto &d : s.scenes)
        for (int c = 0; c < 3; c++)
            if (!d.grid_old[c]) {
                d.grid_old[c] = old_ws;
                old_ws += mac_count(c, d.ni, d.nj, d.nk);
            }
    // input already grouped by brick (the previous step's sorted output):
    // the sort writes no copies; P2G reads each sort slot's particle through
    // the permutation, at nearly consecutive addresses, and so does G2P
    // (particle_gather 1), or P2G writes the sorted copies for G2P (2)
    s.gather = ctx->particle_gather != 0 && ctx->p2g_variant == 4 && ctx->g2p_variant == 0 &&
               params[0].order == GFSX_ORDER_SORTED;
    const bool no_copies = s.gather && ctx->particle_gather == 1;
    for (int c = 0; c < 3; c++) {
        s.xs[c] = no_copies ? nullptr : wsf(S_XS0 + c, s.tot.n);
        s.vs[c] = no_copies ? nullptr : wsf(S_VS0 + c, s.tot.n);
    }
    if (ctx->p2g_variant == 4) {
        // one tile per brick bin (the overflow bins' stay unused)
        s.tiles = (float2 *)wsf(S_P_TILE, 2 * (int64_t)gfsx::kP2gTileNodes * s.tot.nbb);
        s.tile_rows = (unsigned long long *)ws(
            S_P_TFLAG, sizeof(unsigned long long) * (size_t)std::max<int64_t>(s.tot.nrm, 1));
    } else {
        s.dense = ctx->u32(S_DENSE, s.tot.np2g);
        s.dense_count = ctx->u32(S_DENSEN, 2);
    }
    s.num_cu = ctx->num_cu;
    gfsx::Sorted so;
    for (int c = 0; c < 3; c++) {
        so.x[c] = s.xs[c];
        so.v[c] = s.vs[c];
    }
    for (auto &d : s.scenes) d.p2g = gfsx::p2g_view(d, so);
    gfsx::FsEpilogue epi;
    if (fs) {
        // P2G writes the bare normalised values and weights; the epilogue
        // kernels fill, snapshot, add the body force and extrapolate
        gfsx::SceneDesc &d = s.scenes[0];
        int64_t nf[3], tot = 0;
        for (int c = 0; c < 3; c++) tot += (nf[c] = mac_count(c, d.ni, d.nj, d.nk));
        float *nbuf = ctx->f32(S_E_N, tot);
        float *wbuf = ctx->f32(S_E_W, tot);
        std::memset(&epi, 0, sizeof epi);
        epi.ni = d.ni;
        epi.nj = d.nj;
        epi.nk = d.nk;
        epi.solid = scenes[0].solid;
        epi.layer = (int8_t *)ctx->get(S_E_LAYER, (size_t)d.ni * d.nj * d.nk);
        epi.layers = params[0].extrapolation_layers > 0 ? params[0].extrapolation_layers : 7;
        epi.fclass = (uint16_t *)ctx->get(S_E_LIST, sizeof(uint16_t) * (size_t)tot);
        epi.dx = d.dx;
        epi.np = d.n;
        for (int c = 0; c < 3; c++) {
            // the input of output index i: itself (keep), sorted slot i, or
            // the input at the permutation's entry i (no sorted copies)
            epi.x0[c] = params[0].order == GFSX_ORDER_KEEP || !s.xs[0] ? d.xin[c] : s.xs[c];
            epi.x1[c] = d.xout[c];
        }
        int64_t at = 0;
        for (int c = 0; c < 3; c++) {
            float *w = d.wgt[c] ? d.wgt[c] : wbuf + at;
            epi.n[c] = nbuf + at;
            epi.w[c] = w;
            epi.grid[c] = d.grid[c];
            epi.old[c] = d.grid_old[c];
            epi.grav_delta[c] = d.grav_delta[c];
            epi.grav_on[c] = d.grav_on[c];
            d.p2g.grid[c] = nbuf + at;
            d.p2g.grid_old[c] = nullptr;
            d.p2g.wgt[c] = w;
            d.p2g.grav_on[c] = 0;
            at += nf[c];
        }
    }
    s.stencil = params[0].stencil;
    s.keep_order = params[0].order == GFSX_ORDER_KEEP;
    s.p2g_variant = ctx->p2g_variant;
    s.g2p_variant = ctx->g2p_variant;
    s.g2p_brick = ctx->g2p_brick;
    s.grids_only = grids_only;
    // 1/dx a power of two (within FP32's normal range) in every scene
    s.p2 = ctx->fp32_cells != 0;
    for (const auto &d : s.scenes) {
        int e = 0;
        const double m = std::frexp(d.bg.inv_dx, &e);
        s.p2 = s.p2 && m == 0.5 && e > -100 && e < 100;
    }
    hipEvent_t *ev = ctx->timing && !carve ? ctx->next_events() : nullptr;
    if (ctx->p2g_variant == 4) {
        bool ids = false;
        for (auto &d : s.scenes) ids = ids || d.id_out != nullptr;
        gfsx::BrickSort bs = brick_sort(ws, s.tot.n, s.tot.nblk, (uint32_t)s.tot.nbb,
                                        s.keep_order || ids || s.gather);
        if (carve) {
            if (carve->used > carve->cap) {
                std::ostringstream m;
                m << who << ": the step needs " << carve->used << " bytes of scratch, the static "
                  << "buffer holds " << carve->cap << " (fewer particles or a smaller grid)";
                throw gfsx::Error(m.str());
            }
            gfsx::put_scenes(table_dev, s.scenes.data(), nscenes, ctx->stream);
            s.scenes_dev = table_dev;
        } else {
            s.scenes_dev = ctx->table(s.scenes);
            s.r0s = reinterpret_cast<const int64_t *>(s.scenes_dev + nscenes);
        }
        s.bs = &bs;
        if (fs) {
            epi.x0_perm = s.xs[0] ? nullptr : bs.perm;
            epi.fluid = bs.fluid;
            epi.kb0 = s.scenes[0].kb0;
            epi.gbx = s.scenes[0].gbx;
            epi.gby = s.scenes[0].gby;
            s.epi = &epi;
        }
        gfsx::flip_step_brick(s, bs, ctx->stream, ev);
        return s.scenes;
    }
    gfsx::SortBuffers sb = ctx->sort_buffers(s.scenes, s.tot.n, (uint32_t)s.tot.nbin,
                                             (uint32_t)s.tot.nblk);
    s.scenes_dev = sb.scenes;
    // P2G brick flags, one byte per 4x4x4-cell brick of every scene
    sb.brick_flag = (uint8_t *)ctx->get(S_BFLAG, (size_t)s.tot.np2g);
    if (fs) {
        epi.start = sb.start;
        for (int a = 0; a < 3; a++) epi.b[a] = s.scenes[0].bg.b[a];
        s.epi = &epi;
    }
    gfsx::flip_step(s, sb, ctx->stream, ev);
    return s.scenes;
}

void copy_in(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes) GFSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}

void copy_out(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes) GFSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
}

// The step the adjoint differentiates, by the scatter P2G (whose sort the
// adjoint reads) whatever P2G variant the context has selected.
std::vector<gfsx::SceneDesc> scatter_step(gfsx_ctx *ctx, const gfsx_scene *sc,
                                          const gfsx_step_params *pr, const char *who,
                                          bool grids_only) {
    const int p2g_variant = ctx->p2g_variant;
    ctx->p2g_variant = 4;
    try {
        std::vector<gfsx::SceneDesc> laid = flip_step_impl(ctx, sc, pr, 1, who, grids_only);
        ctx->p2g_variant = p2g_variant;
        return laid;
    } catch (...) {
        ctx->p2g_variant = p2g_variant;
        throw;
    }
}

// entries of the brick sort's bin starts for one scene: 8 bins per G2P brick
// of 8^3 cells, the overflow bin, the end
int64_t tape_starts(int ni, int nj, int nk) {
    return 8 * (int64_t)((ni + 7) / 8) * ((nj + 7) / 8) * ((nk + 7) / 8) + 2;
}

void check_step_kind(const char *who, const gfsx_step_params *pr) {
    auto req = [&](bool c, const char *what) {
        if (!c) throw Error(std::string(who) + ": " + what);
    };
    req(pr->stencil == GFSX_TRILINEAR, "the adjoint covers the trilinear stencil only");
    req(pr->order == GFSX_ORDER_KEEP,
        "the adjoints refer to particles in input order (GFSX_ORDER_KEEP)");
    req(pr->epilogue == GFSX_EPILOGUE_NONE,
        "the adjoint covers the transfer without the FluidSimulation epilogue");
}

// the particle arrays of a backward call
struct AdjointIo {
    const float *ins[6], *aout[6];
    float *ain[6];
};

AdjointIo check_adjoint(const char *who, const gfsx_scene *sc, const gfsx_step_params *pr,
                        const gfsx_scene_grad *gr) {
    auto req = [&](bool c, const char *what) {
        if (!c) throw Error(std::string(who) + ": " + what);
    };
    req(sc && pr && gr, "null scene, params or gradients");
    check_step_kind(who, pr);
    check_grid(sc->ni, sc->nj, sc->nk, sc->dx);
    check_particles(sc->n, sc->x_in);
    AdjointIo io = {{sc->x_in, sc->y_in, sc->z_in, sc->vx_in, sc->vy_in, sc->vz_in},
                    {gr->gx_out, gr->gy_out, gr->gz_out, gr->gvx_out, gr->gvy_out, gr->gvz_out},
                    {gr->gx_in, gr->gy_in, gr->gz_in, gr->gvx_in, gr->gvy_in, gr->gvz_in}};
    if (sc->n > 0) {
        for (int a = 0; a < 6; a++) {
            req(io.ins[a] != nullptr, "null particle input");
            req(io.aout[a] != nullptr, "null output adjoint");
            req(io.ain[a] != nullptr, "null input adjoint");
            for (int b = 0; b < 6; b++)
                req((const float *)io.ain[a] != io.ins[b],
                    "input adjoints must not alias the inputs");
        }
    }
    return io;
}

// a recorded step: every grid of the scene, and a tape sized for it
void check_taped(const char *who, const gfsx_scene *sc, const gfsx_step_params *pr,
                 const gfsx_step_tape *tape) {
    auto req = [&](bool c, const char *what) {
        if (!c) throw Error(std::string(who) + ": " + what);
    };
    req(sc && pr && tape, "null scene, params or tape");
    check_step_kind(who, pr);
    check_grid(sc->ni, sc->nj, sc->nk, sc->dx);
    check_particles(sc->n, sc->x_in);
    const float *grids[9] = {sc->u, sc->v, sc->w, sc->u_old, sc->v_old, sc->w_old,
                             sc->wu, sc->wv, sc->ww};
    for (const float *g : grids)
        req(g != nullptr, "every grid of the scene (u, v, w, *_old, wu, wv, ww) is required");
    req(tape->n == sc->n, "the tape holds a different particle count");
    req(tape->starts == tape_starts(sc->ni, sc->nj, sc->nk),
        "tape.starts must be gfsx_step_tape_starts(ni, nj, nk)");
    req(tape->start != nullptr, "null tape.start");
    req(sc->n == 0 || (tape->perm && tape->xs && tape->ys && tape->zs),
        "null tape array (perm, xs, ys, zs)");
}

// The adjoint passes of one step whose grids (u, v, w: after the body force;
// *_old: the normalised P2G result; wu, wv, ww: the P2G weights) are those
// of `grids` and whose brick sort is `so`.
void run_adjoint(gfsx_ctx *ctx, const gfsx_scene *sc, const gfsx_step_params *pr,
                 const AdjointIo &io, const gfsx_scene &grids, gfsx::AdjointSort so) {
    const int ni = sc->ni, nj = sc->nj, nk = sc->nk;
    int64_t nf[3], tot = 0;
    for (int c = 0; c < 3; c++) tot += (nf[c] = mac_count(c, ni, nj, nk));
    const int64_t at[3] = {0, nf[0], nf[0] + nf[1]};
    // dL/dg then dL/dn, one block (the LDS-brick adjoint converts its
    // fixed-point sums into it in one pass)
    float *lg = ctx->f32(S_B_LG, 2 * tot), *ln = lg + tot;
    so.fx = (unsigned long long *)ctx->get(S_B_FX, sizeof(unsigned long long) * 2 * tot);
    so.bound = ctx->u32(S_B_MAX, 3);
    const float *g[3] = {grids.u, grids.v, grids.w};
    const float *old[3] = {grids.u_old, grids.v_old, grids.w_old};
    const float *wgt[3] = {grids.wu, grids.wv, grids.ww};
    gfsx::StepAdjoint a;
    std::memset(&a, 0, sizeof a);
    a.n = sc->n;
    for (int c = 0; c < 3; c++) {
        a.x[c] = io.ins[c];
        a.v[c] = io.ins[3 + c];
        a.ax[c] = io.aout[c];
        a.av[c] = io.aout[3 + c];
        a.gx[c] = io.ain[c];
        a.gv[c] = io.ain[3 + c];
        a.g[c] = g[c];
        a.old[c] = old[c];
        a.w[c] = wgt[c];
        a.lg[c] = lg + at[c];
        a.ln[c] = ln + at[c];
    }
    a.mac = mac_grid(nullptr, nullptr, nullptr, ni, nj, nk, sc->dx);
    a.r_pic = (float)pr->pic_ratio;
    a.r_flip = (float)(1 - pr->pic_ratio);
    a.h_half = (float)(0.5 * pr->dt);
    a.h_full = (float)pr->dt;
    a.h_sixth = (float)(pr->dt / 6.0f);
    a.half_dx = (float)(0.5 * sc->dx);
    a.inv_r2 = (float)(1.0 / (sc->dx * sc->dx));
    gfsx::flip_step_adjoint(a, ctx->adjoint_variant == 0 ? &so : nullptr, ctx->stream);
}

}  // namespace

extern "C" {

const char *gfsx_get_error_message(void) { return g_error.c_str(); }

int gfsx_get_device_count(int *err) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        fail(err, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
        return 0;
    }
    ok(err);
    return n;
}

gfsx_ctx *gfsx_create(int device, int *err) {
    gfsx_ctx *c = nullptr;
    try {
        int n = 0;
        GFSX_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, "gfsx_create: no such device");
        c = new gfsx_ctx();
        c->device = device;
        GFSX_HIP(hipSetDevice(device));
        GFSX_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
        c->stream = c->own_stream;
        hipDeviceProp_t prop;
        GFSX_HIP(hipGetDeviceProperties(&prop, device));
        std::ostringstream ss;
        ss << "device " << device << ": " << prop.name << " (" << prop.gcnArchName
           << "), " << prop.multiProcessorCount << " CUs, "
           << (prop.totalGlobalMem >> 20) << " MiB";
        c->info = ss.str();
        c->num_cu = std::max(prop.multiProcessorCount, 1);
        // the device address of this device's copy of the static scratch,
        // looked up once per device with that device current
        if (device < kMaxDevices) {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            if (!g_scratch_dev[device]) {
                void *p = nullptr;
                GFSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
                require(p != nullptr, "gfsx_create: no device address for the static scratch");
                g_scratch_dev[device] = p;
            }
        }
        ok(err);
        return c;
    } catch (const std::exception &e) {
        delete c;
        fail(err, e.what());
        return nullptr;
    }
}

void gfsx_destroy(gfsx_ctx *ctx, int *err) {
    if (!ctx) {
        ok(err);
        return;
    }
    try {
        GFSX_HIP(hipSetDevice(ctx->device));
        hipError_t e = hipStreamSynchronize(ctx->stream);
        delete ctx;
        if (e != hipSuccess)
            throw Error(std::string("gfsx_destroy: ") + hipGetErrorString(e));
        ok(err);
    } catch (const std::exception &e) {
        fail(err, e.what());
    }
}

void gfsx_set_stream(gfsx_ctx *ctx, void *stream, int *err) {
    guarded(ctx, err, [&] {
        ctx->stream = (hipStream_t)stream;
    });
}

void gfsx_synchronize(gfsx_ctx *ctx, int *err) {
    guarded(ctx, err, [&] { GFSX_HIP(hipStreamSynchronize(ctx->stream)); });
}

const char *gfsx_get_device_info(gfsx_ctx *ctx, int *err) {
    if (!ctx) {
        fail(err, "gfsx: null context");
        return "";
    }
    ok(err);
    return ctx->info.c_str();
}

// ---------------------------------------------------------------- P2G

void gfsx_p2g_device(gfsx_ctx *ctx, const float *pos, const float *val,
                     int64_t n, double radius, const float offset[3],
                     double dx, int ni, int nj, int nk, float *S, float *W,
                     int *err) {
    guarded(ctx, err, [&] {
        p2g_device_impl(ctx, pos, val, n, radius, offset, dx, ni, nj, nk, S, W);
    });
}

void gfsx_p2g(gfsx_ctx *ctx, const float *pos, const float *val, int64_t n,
              double radius, const float offset[3], double dx, int ni, int nj,
              int nk, float *S, float *W, int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, pos);
        require(S != nullptr, "gfsx: null scalar field");
        const int64_t nn = (int64_t)ni * nj * nk;
        hipStream_t s = ctx->stream;
        float *dpos = ctx->f32(S_D_POS, 3 * n);
        float *dval = val ? ctx->f32(S_D_VAL, n) : nullptr;
        float *dS = ctx->f32(S_D_S, nn);
        float *dW = W ? ctx->f32(S_D_W, nn) : nullptr;
        copy_in(dpos, pos, sizeof(float) * 3 * n, s);
        if (val) copy_in(dval, val, sizeof(float) * n, s);
        copy_in(dS, S, sizeof(float) * nn, s);
        if (W) copy_in(dW, W, sizeof(float) * nn, s);
        p2g_device_impl(ctx, dpos, dval, n, radius, offset, dx, ni, nj, nk, dS, dW);
        copy_out(S, dS, sizeof(float) * nn, s);
        if (W) copy_out(W, dW, sizeof(float) * nn, s);
        GFSX_HIP(hipStreamSynchronize(s));
    });
}

void gfsx_apply_weight_device(gfsx_ctx *ctx, float *S, const float *W,
                              int64_t count, int *err) {
    guarded(ctx, err, [&] {
        require(count >= 0, "gfsx: negative count");
        require(count == 0 || (S && W), "gfsx: null field");
        gfsx::apply_weight(S, W, count, ctx->stream);
    });
}

// ---------------------------------------------------------------- G2P

void gfsx_g2p_device(gfsx_ctx *ctx, const float *pos, int64_t n,
                     const float *u, const float *v, const float *w, int ni,
                     int nj, int nk, double dx, int stencil, float *out,
                     int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, pos);
        require(stencil == GFSX_TRILINEAR || stencil == GFSX_TRICUBIC, "gfsx: unknown stencil");
        require(u && v && w, "gfsx: null velocity grid");
        require(n == 0 || out, "gfsx: null output");
        gfsx::g2p(pos, n, mac_grid(u, v, w, ni, nj, nk, dx), stencil, out, ctx->stream);
    });
}

void gfsx_g2p(gfsx_ctx *ctx, const float *pos, int64_t n, const float *u,
              const float *v, const float *w, int ni, int nj, int nk,
              double dx, int stencil, float *out, int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, pos);
        require(stencil == GFSX_TRILINEAR || stencil == GFSX_TRICUBIC, "gfsx: unknown stencil");
        require(u && v && w, "gfsx: null velocity grid");
        require(n == 0 || out, "gfsx: null output");
        hipStream_t s = ctx->stream;
        float *dpos = ctx->f32(S_D_POS, 3 * n);
        float *dout = ctx->f32(S_D_OUT, 3 * n);
        float *dg[3];
        const float *hg[3] = {u, v, w};
        const int slots[3] = {S_D_U, S_D_V, S_D_WW};
        for (int c = 0; c < 3; c++) {
            int64_t m = mac_count(c, ni, nj, nk);
            dg[c] = ctx->f32(slots[c], m);
            copy_in(dg[c], hg[c], sizeof(float) * m, s);
        }
        copy_in(dpos, pos, sizeof(float) * 3 * n, s);
        gfsx::g2p(dpos, n, mac_grid(dg[0], dg[1], dg[2], ni, nj, nk, dx), stencil, dout, s);
        copy_out(out, dout, sizeof(float) * 3 * n, s);
        GFSX_HIP(hipStreamSynchronize(s));
    });
}

// ---------------------------------------------------------------- RK

void gfsx_advect_device(gfsx_ctx *ctx, int order, const float *pos, int64_t n,
                        const float *u, const float *v, const float *w, int ni,
                        int nj, int nk, double dx, double dt, int stencil,
                        float *out, int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, pos);
        require(order >= 1 && order <= 4, "gfsx: Runge-Kutta order must be 1..4");
        require(stencil == GFSX_TRILINEAR || stencil == GFSX_TRICUBIC, "gfsx: unknown stencil");
        require(std::isfinite(dt), "gfsx: dt must be finite");
        require(u && v && w, "gfsx: null velocity grid");
        require(n == 0 || out, "gfsx: null output");
        gfsx::advect(pos, n, mac_grid(u, v, w, ni, nj, nk, dx), order, dt, stencil, out,
                     ctx->stream);
    });
}

void gfsx_advect(gfsx_ctx *ctx, int order, const float *pos, int64_t n,
                 const float *u, const float *v, const float *w, int ni,
                 int nj, int nk, double dx, double dt, int stencil, float *out,
                 int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, pos);
        require(order >= 1 && order <= 4, "gfsx: Runge-Kutta order must be 1..4");
        require(stencil == GFSX_TRILINEAR || stencil == GFSX_TRICUBIC, "gfsx: unknown stencil");
        require(std::isfinite(dt), "gfsx: dt must be finite");
        require(u && v && w, "gfsx: null velocity grid");
        require(n == 0 || out, "gfsx: null output");
        hipStream_t s = ctx->stream;
        float *dpos = ctx->f32(S_D_POS, 3 * n);
        float *dout = ctx->f32(S_D_OUT, 3 * n);
        float *dg[3];
        const float *hg[3] = {u, v, w};
        const int slots[3] = {S_D_U, S_D_V, S_D_WW};
        for (int c = 0; c < 3; c++) {
            int64_t m = mac_count(c, ni, nj, nk);
            dg[c] = ctx->f32(slots[c], m);
            copy_in(dg[c], hg[c], sizeof(float) * m, s);
        }
        copy_in(dpos, pos, sizeof(float) * 3 * n, s);
        gfsx::advect(dpos, n, mac_grid(dg[0], dg[1], dg[2], ni, nj, nk, dx), order, dt,
                     stencil, dout, s);
        copy_out(out, dout, sizeof(float) * 3 * n, s);
        GFSX_HIP(hipStreamSynchronize(s));
    });
}

// ---------------------------------------------------------------- step

void gfsx_flip_step(gfsx_ctx *ctx, const gfsx_scene *sc,
                    const gfsx_step_params *pr, int *err) {
    guarded(ctx, err, [&] {
        require(sc && pr, "gfsx_flip_step: null scene or params");
        flip_step_impl(ctx, sc, pr, 1, "gfsx_flip_step");
    });
}

void gfsx_flip_step_noalloc(gfsx_ctx *ctx, const gfsx_scene *sc,
                            const gfsx_step_params *pr, int *err) {
    guarded(ctx, err, [&] {
        const char *who = "gfsx_flip_step_noalloc";
        require(sc && pr, "gfsx_flip_step_noalloc: null scene or params");
        void *base = nullptr;
        if (ctx->device < kMaxDevices) {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            base = g_scratch_dev[ctx->device];
        }
        require(base != nullptr, "gfsx_flip_step_noalloc: no static scratch on this device");
        Carve carve;
        carve.base = static_cast<char *>(base);
        carve.cap = kNoallocScratchBytes;
        flip_step_impl(ctx, sc, pr, 1, who, false, &carve);
    });
}

void gfsx_flip_step_batched(gfsx_ctx *ctx, const gfsx_scene *scenes, int nscenes,
                            const gfsx_step_params *params, int *err) {
    guarded(ctx, err, [&] {
        require(scenes && params, "gfsx_flip_step_batched: null scenes or params");
        require(nscenes >= 1 && nscenes <= 65536, "gfsx_flip_step_batched: nscenes must be 1..65536");
        flip_step_impl(ctx, scenes, params, nscenes, "gfsx_flip_step_batched");
    });
}

// ---------------------------------------------------------------- adjoints

void gfsx_p2g_backward_device(gfsx_ctx *ctx, const float *pos, const float *val,
                              int64_t n, double radius, const float offset[3],
                              double dx, int ni, int nj, int nk, const float *gS,
                              const float *gW, float *gpos, float *gval, int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, pos);
        require(std::isfinite(radius) && radius > 0.0, "gfsx: radius must be positive and finite");
        require(radius / dx <= 64.0, "gfsx: radius larger than 64 cells");
        require(offset != nullptr, "gfsx: null offset");
        require(gS || gW, "gfsx_p2g_backward_device: no field adjoint given");
        require(n == 0 || gpos, "gfsx_p2g_backward_device: null position adjoint");
        gfsx::p2g_adjoint(pos, val, n, radius, offset, dx, ni, nj, nk, gS, gW, gpos, gval,
                          ctx->stream);
    });
}

void gfsx_flip_step_backward(gfsx_ctx *ctx, const gfsx_scene *sc,
                             const gfsx_step_params *pr, const gfsx_scene_grad *gr,
                             int *err) {
    guarded(ctx, err, [&] {
        const char *who = "gfsx_flip_step_backward";
        const AdjointIo io = check_adjoint(who, sc, pr, gr);
        // forward grids of this step, recomputed into the workspace
        const int64_t n = sc->n;
        int64_t nf[3], tot = 0;
        for (int c = 0; c < 3; c++) tot += (nf[c] = mac_count(c, sc->ni, sc->nj, sc->nk));
        float *grid = ctx->f32(S_B_GRID, tot), *old = ctx->f32(S_B_OLD, tot);
        float *wgt = ctx->f32(S_B_WGT, tot);
        float *xo = ctx->f32(S_B_XO, 3 * n), *vo = ctx->f32(S_B_VO, 3 * n);
        const int64_t at[3] = {0, nf[0], nf[0] + nf[1]};
        gfsx_scene f = *sc;
        f.u = grid;
        f.v = grid + at[1];
        f.w = grid + at[2];
        f.u_old = old;
        f.v_old = old + at[1];
        f.w_old = old + at[2];
        f.wu = wgt;
        f.wv = wgt + at[1];
        f.ww = wgt + at[2];
        f.x_out = xo;
        f.y_out = xo + n;
        f.z_out = xo + 2 * n;
        f.vx_out = vo;
        f.vy_out = vo + n;
        f.vz_out = vo + 2 * n;
        f.id_in = nullptr;
        f.id_out = nullptr;
        // (the brick sort, its sorted copies and the P2G grids; no particle
        // update), always by the scatter P2G, whose sort the adjoint reads
        const std::vector<gfsx::SceneDesc> laid = scatter_step(ctx, &f, pr, who, true);
        gfsx::AdjointSort so;
        so.pstart = (const uint32_t *)ctx->buf[S_K_PSTART];
        so.perm = (const uint32_t *)ctx->buf[S_PERM];
        for (int c = 0; c < 3; c++) so.xs[c] = (const float *)ctx->buf[S_XS0 + c];
        so.nbb = laid[0].nbb;
        so.gbx = laid[0].gbx;
        so.gby = laid[0].gby;
        run_adjoint(ctx, sc, pr, io, f, so);
    });
}

int64_t gfsx_step_tape_starts(int ni, int nj, int nk, int *err) {
    try {
        check_grid(ni, nj, nk, 1.0);
        ok(err);
        return tape_starts(ni, nj, nk);
    } catch (const std::exception &e) {
        fail(err, e.what());
        return 0;
    }
}

void gfsx_flip_step_record(gfsx_ctx *ctx, const gfsx_scene *sc, const gfsx_step_params *pr,
                           gfsx_step_tape *tape, int *err) {
    guarded(ctx, err, [&] {
        const char *who = "gfsx_flip_step_record";
        check_taped(who, sc, pr, tape);
        const int64_t n = sc->n;
        const std::vector<gfsx::SceneDesc> laid = scatter_step(ctx, sc, pr, who, false);
        if (n > 0) {
            const size_t b = sizeof(float) * (size_t)n;
            GFSX_HIP(hipMemcpyAsync(tape->perm, ctx->buf[S_PERM], sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice,
                                    ctx->stream));
            float *xs[3] = {tape->xs, tape->ys, tape->zs};
            for (int c = 0; c < 3; c++)
                GFSX_HIP(hipMemcpyAsync(xs[c], ctx->buf[S_XS0 + c], b, hipMemcpyDeviceToDevice,
                                        ctx->stream));
            GFSX_HIP(hipMemcpyAsync(tape->start, ctx->buf[S_K_PSTART],
                                    sizeof(uint32_t) * (size_t)tape->starts,
                                    hipMemcpyDeviceToDevice, ctx->stream));
        }
        tape->gbx = laid[0].gbx;
        tape->gby = laid[0].gby;
    });
}

void gfsx_flip_step_backward_tape(gfsx_ctx *ctx, const gfsx_scene *sc,
                                  const gfsx_step_params *pr, const gfsx_scene_grad *gr,
                                  const gfsx_step_tape *tape, int *err) {
    guarded(ctx, err, [&] {
        const char *who = "gfsx_flip_step_backward_tape";
        const AdjointIo io = check_adjoint(who, sc, pr, gr);
        check_taped(who, sc, pr, tape);
        const int gbx = (sc->ni + 7) / 8, gby = (sc->nj + 7) / 8;
        if (tape->gbx != gbx || tape->gby != gby)
            throw Error(std::string(who) + ": the tape was not recorded for this grid");
        gfsx::AdjointSort so;
        so.pstart = tape->start;
        so.perm = tape->perm;
        so.xs[0] = tape->xs;
        so.xs[1] = tape->ys;
        so.xs[2] = tape->zs;
        so.nbb = (uint32_t)(tape->starts - 2);
        so.gbx = gbx;
        so.gby = gby;
        run_adjoint(ctx, sc, pr, io, *sc, so);
    });
}

// ---------------------------------------------------------------- removal

int64_t gfsx_remove_particles(gfsx_ctx *ctx, const gfsx_scene *sc, int max_per_cell,
                              uint64_t seed, int *err) {
    int64_t kept = 0;
    guarded(ctx, err, [&] {
        const char *who = "gfsx_remove_particles";
        auto req = [&](bool c, const char *what) {
            if (!c) throw Error(std::string(who) + ": " + what);
        };
        req(sc != nullptr, "null scene");
        check_grid(sc->ni, sc->nj, sc->nk, sc->dx);
        check_particles(sc->n, sc->x_in);
        req(max_per_cell >= 1, "max_per_cell must be at least 1");
        const int64_t n = sc->n;
        if (n > 0) {
            req(sc->x_in && sc->y_in && sc->z_in && sc->vx_in && sc->vy_in && sc->vz_in,
                "null particle input");
            req(sc->x_out && sc->y_out && sc->z_out && sc->vx_out && sc->vy_out && sc->vz_out,
                "null particle output");
            const void *ins[6] = {sc->x_in, sc->y_in, sc->z_in, sc->vx_in, sc->vy_in, sc->vz_in};
            const void *outs[6] = {sc->x_out, sc->y_out, sc->z_out, sc->vx_out, sc->vy_out,
                                   sc->vz_out};
            for (int a = 0; a < 6; a++)
                for (int b = 0; b < 6; b++) req(ins[a] != outs[b], "outputs must not alias inputs");
            req(!sc->id_out || sc->id_out != sc->id_in, "id_out must not alias id_in");
        }
        // the cell sort over the grid's cells (Grid3d::positionToGridIndex)
        gfsx::BinGrid bg;
        bg.b[0] = sc->ni;
        bg.b[1] = sc->nj;
        bg.b[2] = sc->nk;
        bg.pad[0] = bg.pad[1] = bg.pad[2] = 0;
        bg.off[0] = bg.off[1] = bg.off[2] = 0.0f;
        bg.inv_dx = 1.0 / sc->dx;
        bg.nbin = (uint32_t)((int64_t)sc->ni * sc->nj * sc->nk);
        std::vector<gfsx::SceneDesc> t(1);
        std::memset(t.data(), 0, sizeof(gfsx::SceneDesc));
        t[0].xin[0] = sc->x_in;
        t[0].xin[1] = sc->y_in;
        t[0].xin[2] = sc->z_in;
        t[0].n = n;
        t[0].bg = bg;
        gfsx::SortBuffers sb = ctx->sort_buffers(t, n, bg.nbin + 1, (uint32_t)gfsx::sort_blocks(n));
        gfsx::cell_sort(sb, n, ctx->stream, false);
        gfsx::RemoveArgs r;
        std::memset(&r, 0, sizeof r);
        r.n = n;
        const float *xin[3] = {sc->x_in, sc->y_in, sc->z_in};
        const float *vin[3] = {sc->vx_in, sc->vy_in, sc->vz_in};
        float *xout[3] = {sc->x_out, sc->y_out, sc->z_out};
        float *vout[3] = {sc->vx_out, sc->vy_out, sc->vz_out};
        for (int c = 0; c < 3; c++) {
            r.x[c] = xin[c];
            r.v[c] = vin[c];
            r.xo[c] = xout[c];
            r.vo[c] = vout[c];
        }
        r.id_in = sc->id_in;
        r.id_out = sc->id_out;
        r.bg = bg;
        r.start = sb.start;
        r.perm = sb.perm;
        r.solid = sc->solid;
        r.max_per_cell = max_per_cell;
        r.seed = seed;
        r.keep = ctx->u32(S_R_KEEP, n + 1);
        r.dst = ctx->u32(S_R_DST, n + 1);
        // every over-full cell holds more than max_per_cell particles
        r.full = ctx->u32(S_R_FULL, std::min<int64_t>((int64_t)bg.nbin, n / (max_per_cell + 1) + 1));
        r.nfull = ctx->u32(S_R_NFULL, 1);
        uint32_t *scratch = ctx->u32(S_R_SCAN, gfsx::scan_scratch_words(n + 1));
        gfsx::fill_u32(r.keep, 0u, n + 1, ctx->stream);
        gfsx::remove_mark(r, ctx->stream);
        // keep[n] = 0, so dst[n] is the number kept
        gfsx::exclusive_scan(r.keep, r.dst, n + 1, scratch, ctx->stream);
        gfsx::remove_compact(r, ctx->stream);
        uint32_t total = 0;
        GFSX_HIP(hipMemcpyAsync(&total, r.dst + n, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                ctx->stream));
        GFSX_HIP(hipStreamSynchronize(ctx->stream));
        kept = (int64_t)total;
    });
    return kept;
}

void gfsx_set_option(gfsx_ctx *ctx, const char *name, int value, int *err) {
    guarded(ctx, err, [&] {
        require(name != nullptr, "gfsx_set_option: null name");
        if (std::strcmp(name, "p2g_variant") == 0) {
            require(value >= 0 && value <= 4, "gfsx_set_option: p2g_variant must be 0..4");
            ctx->p2g_variant = value;
        } else if (std::strcmp(name, "g2p_variant") == 0) {
            require(value >= 0 && value <= 1, "gfsx_set_option: g2p_variant must be 0..1");
            ctx->g2p_variant = value;
        } else if (std::strcmp(name, "g2p_brick") == 0) {
            require(value == 8 || value == 4, "gfsx_set_option: g2p_brick must be 8 or 4");
            ctx->g2p_brick = value;
        } else if (std::strcmp(name, "fp32_cells") == 0) {
            require(value == 0 || value == 1, "gfsx_set_option: fp32_cells must be 0 or 1");
            ctx->fp32_cells = value;
        } else if (std::strcmp(name, "particle_gather") == 0) {
            require(value >= 0 && value <= 2, "gfsx_set_option: particle_gather must be 0..2");
            ctx->particle_gather = value;
        } else if (std::strcmp(name, "adjoint_variant") == 0) {
            require(value >= 0 && value <= 1, "gfsx_set_option: adjoint_variant must be 0..1");
            ctx->adjoint_variant = value;
        } else {
            throw Error(std::string("gfsx_set_option: unknown option ") + name);
        }
    });
}

int gfsx_get_option(gfsx_ctx *ctx, const char *name, int *err) {
    int v = 0;
    guarded(ctx, err, [&] {
        require(name != nullptr, "gfsx_get_option: null name");
        if (std::strcmp(name, "p2g_variant") == 0) v = ctx->p2g_variant;
        else if (std::strcmp(name, "g2p_variant") == 0) v = ctx->g2p_variant;
        else if (std::strcmp(name, "g2p_brick") == 0) v = ctx->g2p_brick;
        else if (std::strcmp(name, "particle_gather") == 0) v = ctx->particle_gather;
        else if (std::strcmp(name, "fp32_cells") == 0) v = ctx->fp32_cells;
        else if (std::strcmp(name, "adjoint_variant") == 0) v = ctx->adjoint_variant;
        else throw Error(std::string("gfsx_get_option: unknown option ") + name);
    });
    return v;
}

void gfsx_timing_begin(gfsx_ctx *ctx, int *err) {
    guarded(ctx, err, [&] {
        ctx->timing = true;
        ctx->events_used = 0;
    });
}

int gfsx_timing_end(gfsx_ctx *ctx, double stage_ms[4], int *err) {
    int steps = 0;
    guarded(ctx, err, [&] {
        require(stage_ms != nullptr, "gfsx_timing_end: null output");
        GFSX_HIP(hipStreamSynchronize(ctx->stream));
        for (int a = 0; a < 4; a++) stage_ms[a] = 0.0;
        for (size_t e = 0; e < ctx->events_used; e++) {
            const auto &ev = ctx->events[e];
            for (int a = 0; a < 3; a++) {
                float ms = 0.0f;
                GFSX_HIP(hipEventElapsedTime(&ms, ev[a], ev[a + 1]));
                stage_ms[a] += ms;
            }
            float tot = 0.0f;
            GFSX_HIP(hipEventElapsedTime(&tot, ev[0], ev[3]));
            stage_ms[3] += tot;
        }
        steps = (int)ctx->events_used;
        ctx->events_used = 0;
        ctx->timing = false;
    });
    return steps;
}

void gfsx_cell_sort_device(gfsx_ctx *ctx, const float *x, const float *y,
                           const float *z, int64_t n, double dx, int ni,
                           int nj, int nk, uint32_t *perm,
                           uint32_t *cell_start, int *err) {
    guarded(ctx, err, [&] {
        check_grid(ni, nj, nk, dx);
        check_particles(n, x);
        require(n == 0 || (y && z && perm), "gfsx: null array");
        require(cell_start != nullptr, "gfsx: null cell_start");
        gfsx::BinGrid bg;
        bg.b[0] = ni;
        bg.b[1] = nj;
        bg.b[2] = nk;
        bg.pad[0] = bg.pad[1] = bg.pad[2] = 0;
        bg.off[0] = bg.off[1] = bg.off[2] = 0.0f;
        bg.inv_dx = 1.0 / dx;
        bg.nbin = (uint32_t)((int64_t)ni * nj * nk);
        std::vector<gfsx::SceneDesc> t(1);
        std::memset(t.data(), 0, sizeof(gfsx::SceneDesc));
        t[0].xin[0] = x;
        t[0].xin[1] = y;
        t[0].xin[2] = z;
        t[0].n = n;
        t[0].bg = bg;
        gfsx::SortBuffers sb = ctx->sort_buffers(t, n, bg.nbin + 1, (uint32_t)gfsx::sort_blocks(n));
        gfsx::cell_sort(sb, n, ctx->stream, false);
        if (n > 0)
            GFSX_HIP(hipMemcpyAsync(perm, sb.perm, sizeof(uint32_t) * n,
                                    hipMemcpyDeviceToDevice, ctx->stream));
        GFSX_HIP(hipMemcpyAsync(cell_start, sb.start, sizeof(uint32_t) * ((size_t)bg.nbin + 1),
                                hipMemcpyDeviceToDevice, ctx->stream));
    });
}

}  // extern "C"
