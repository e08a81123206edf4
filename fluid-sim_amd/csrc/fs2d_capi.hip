// fs2d_capi.hip -- C ABI (include/fs2d.h) over the gfx950 kernels.
//
// Host orchestration of FluidSim2D::runFrame (src/FluidSim2D.cpp:94-138): one
// HIP stream per handle, all state resident in HBM, no per-step allocation
// (the reference allocates temporaries in every stage, FluidSim2D.cpp:147,
// 186, 253, 334, 366, 390, 399, 471).  The PCG runs its iterations without
// host round trips; the host reads the convergence flag once per batch.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fs2d.h"
#include "fs2d_adjoint.hpp"
#include "fs2d_batch.hpp"

using namespace fs2d;

namespace {

enum KernelClass {
    K_SL_ADVECT, K_GRAVITY, K_ASSEMBLE, K_FACTOR, K_PCG_FWD, K_PCG_BWD, K_PCG_SPMV, K_PCG_AXPY,
    K_UPDATE_VEL, K_EXTRAPOLATE, K_PARTICLE_SORT, K_P2G, K_G2P, K_PARTICLE_ADVECT,
    K_LS_SWEEP, K_LS_OTHER, K_DIAGNOSTICS, K_MISC, K_ADJ_SL, K_ADJ_OTHER, K_COUNT
};
const char* kKernelNames[K_COUNT] = {
    "sl_advect", "gravity", "assemble", "mic0_factor", "pcg_forward_solve", "pcg_backward_solve",
    "pcg_spmv_dot", "pcg_axpy_norm", "update_velocity", "extrapolate", "particle_sort", "p2g", "g2p",
    "particle_advect", "levelset_sweep", "levelset_other", "diagnostics", "misc", "adjoint_sl",
    "adjoint_other"};

constexpr int kMaxBlocks = 2048;
// Solver layouts of at least this many entries (a 512^2 grid has 294912) run
// a PCG iteration as two launches by default: measured per iteration against
// four launches, 0.94-0.98x from 512^2 up, 1.03-1.08x at 64^2-384^2, where
// the separate SpMV and axpy are short and the fused launches' head and tail
// cost more than they hide.
constexpr long kFusedPcgMinEntries = 1L << 18;

int blocks_for(long n) {
    long b = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min(b, (long)kMaxBlocks));
}

struct TimedLaunch {
    int cls;
    hipEvent_t a, b;
};

// What a captured step depends on besides device memory: the kernels'
// arguments (sizes, parameters, buffer pointers, which the semi-Lagrangian
// stage swaps) and the host flags that choose launches.  A step is replayed
// only from the state it was captured in.
struct GraphKey {
    double dt, gx, gy, alpha, tau, sigma, tol;
    int max_iters, budget, pcg_grid, sorted_valid;
    long np;
    const void* ptr[12];  // u, v, u2, v2, dens, dens2, px, py, pvx, pvy, pkey, sorted
    bool operator==(const GraphKey& o) const { return memcmp(this, &o, sizeof o) == 0; }
};
// The handle's host state after the captured step.
struct GraphAfter {
    float *u, *v, *u2, *v2, *dens, *dens2;
    bool sorted_valid;
};
struct StepGraph {
    GraphKey key;
    GraphAfter after;
    hipGraphExec_t pre = nullptr, post = nullptr;
    int budget = 0;
};

}  // namespace

struct fs2d_sim {
    int nx = 0, ny = 0, k = 0, mode = 0, flags = 0;
    double dt = 0, dx = 0, dr = 0, rho = 0, gx = 0, gy = 0, alpha = 0;
    double tau = 0.999, sigma = 0.25, tol = 1e-12;
    int max_iters = 200;
    hipStream_t stream = nullptr;
    int device = 0;

    // grid state
    float *u = nullptr, *v = nullptr, *u2 = nullptr, *v2 = nullptr, *nu = nullptr, *nv = nullptr;
    float *dens = nullptr, *dens2 = nullptr;
    uint8_t* cell = nullptr;
    uint8_t* cell2 = nullptr;  // flags written by the bubble fill (FS2D_FLAG_PARTICLE_CELLS)
    double *phi = nullptr, *phi2 = nullptr;
    int *t0 = nullptr, *t1 = nullptr;
    int* ls_chg = nullptr;  // 16 closest-particle pass change flags (ls_first_unchanged)
    // projection: row-major pressure + SK4 solver arrays (fs2d_skew.hpp)
    float* p = nullptr;
    SkGeom geom{};
    float *r = nullptr, *z = nullptr, *q = nullptr, *w = nullptr, *s[2] = {nullptr, nullptr}, *pk = nullptr;
    float *Adiag = nullptr, *P = nullptr, *fx = nullptr, *fy = nullptr, *gx_ = nullptr, *gy_ = nullptr;
    uint8_t* cpl = nullptr;  // coupling bits (CPL_*)
    PcgScalars* sc = nullptr;
    PcgScalars* h_sc = nullptr;
    double *qq = nullptr, *ws_part = nullptr, *rmax_part = nullptr;
    unsigned long long* sk_ticket = nullptr;
    unsigned long long* sk_handoff = nullptr;
    // a PCG iteration as two launches: the p/r update with the forward solve
    // (fwd_axpy_kernel: its own ticket, hand-off rows and per-run flags) and
    // the backward solve with the next SpMV (bwd_spmv_kernel: ticket,
    // hand-off rows, per-strip completion flags, tagged sigma words)
    bool fused_pcg = true;
    int ncu = 0;  // compute units of the device
    unsigned long long* fa_ticket = nullptr;
    unsigned long long* fa_handoff = nullptr;
    unsigned* fa_flags = nullptr;
    unsigned long long* fb_ticket = nullptr;
    unsigned long long* fb_handoff = nullptr;
    unsigned* fb_flags = nullptr;
    unsigned long long* sig_words = nullptr;
    // extrapolation
    uint32_t *mu = nullptr, *mv = nullptr;
    int *fu0 = nullptr, *fv0 = nullptr;    // entries in layer order
    int *fu1 = nullptr, *fv1 = nullptr;    // per-layer counts, offsets, cursors
    int* fcount = nullptr;                 // layers of u, v
    int* fpos = nullptr;                   // u entries then v entries: position in layer order
    int4* fsrc = nullptr;                  // sources of each ordered entry (u then v)
    float4* fkv = nullptr;                 // known sources' values (u then v)
    // particles
    float *px = nullptr, *py = nullptr, *pvx = nullptr, *pvy = nullptr;
    long np = 0, np_cap = 0;
    int *pkey = nullptr, *counts = nullptr, *starts = nullptr, *cursor = nullptr, *sorted = nullptr;
    int* tile_sums = nullptr;
    bool sorted_valid = false;
    // scalars
    Diag* dg = nullptr;
    double* partials = nullptr;
    int* werr = nullptr;
    Diag* h_dg = nullptr;
    int* h_err = nullptr;

    std::vector<void*> allocs;
    int* render_cnt = nullptr;  // fs2d_render_buffers: block counts and totals, allocated on first use
    double current_time = 0, orig_volume = 0;
    long long steps = 0;
    int last_iters_hint = 8;
    int pcg_grid = 0;  // physical workgroups of the PCG SpMV / axpy (0: one per virtual workgroup)
    // Captured steps (fs2d_set_graph): each replays the stages up to a fixed
    // number of PCG iterations, then (after one host check of convergence)
    // the rest of the step.
    bool use_graph = false;
    int graph_budget = 0;
    std::vector<StepGraph> graphs;
    std::string err;

    unsigned long long timing = 0;  // bit k: launches of kernel class k are timed
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double ktime[K_COUNT] = {};
    long long kcount[K_COUNT] = {};
};

namespace {

fs2d_status fail(fs2d_sim* sim, fs2d_status st, const char* fmt, ...) {
    if (sim) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        sim->err = buf;
    }
    return st;
}

#define HIPCHK(sim, x)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return fail(sim, e_ == hipErrorOutOfMemory ? FS2D_ERR_OOM : FS2D_ERR_DEVICE, "%s: %s", #x, \
                        hipGetErrorString(e_));                                             \
    } while (0)

// Per-kernel timing state lives in the handle (fs2d_sim) or the batch
// (fs2d_batch): members timing (class mask), stream, pending, event_pool,
// ktime, kcount.
template <class H>
hipEvent_t get_event(H* s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Times one launch on the owner's stream when its class is in the timing mask.
template <class H>
struct Timed {
    H* s;
    int cls;
    hipEvent_t a = nullptr;
    Timed(H* s_, int c) : s(s_), cls(c) {
        if ((s->timing >> c) & 1ull) {
            a = get_event(s);
            if (a) (void)hipEventRecord(a, s->stream);
        }
    }
    ~Timed() {
        if (a) {
            hipEvent_t b = get_event(s);
            if (b) {
                (void)hipEventRecord(b, s->stream);
                s->pending.push_back({cls, a, b});
            } else {
                s->event_pool.push_back(a);
            }
        }
    }
};

template <class H>
void resolve_timing(H* s) {
    for (auto& t : s->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            s->ktime[t.cls] += ms;
            s->kcount[t.cls] += 1;
        }
        s->event_pool.push_back(t.a);
        s->event_pool.push_back(t.b);
    }
    s->pending.clear();
}

template <class T>
fs2d_status dalloc(fs2d_sim* s, T** p, size_t n) {
    void* ptr = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(s, FS2D_ERR_OOM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    s->allocs.push_back(ptr);
    *p = (T*)ptr;
    e = hipMemsetAsync(ptr, 0, bytes, s->stream);
    if (e != hipSuccess) return fail(s, FS2D_ERR_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
    return FS2D_OK;
}

#define TRY(x)                           \
    do {                                 \
        fs2d_status st_ = (x);           \
        if (st_ != FS2D_OK) return st_;  \
    } while (0)

template <class T>
void fill(fs2d_sim* s, T* a, T v, size_t n) {
    hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks_for((long)n)), dim3(kBlock), 0, s->stream, a, v, n);
}
template <class T>
void dcopy(fs2d_sim* s, T* dst, const T* src, size_t n) {
    hipLaunchKernelGGL(copy_kernel<T>, dim3(blocks_for((long)n)), dim3(kBlock), 0, s->stream, dst, src, n);
}

fs2d_status h2d(fs2d_sim* s, void* dst, const void* src, size_t bytes) {
    HIPCHK(s, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return FS2D_OK;
}
fs2d_status d2h(fs2d_sim* s, void* dst, const void* src, size_t bytes) {
    HIPCHK(s, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return FS2D_OK;
}

// USE_GHOST_PRESSURE of the reference (src/FluidSim2D.cpp:15), on by default.
int ghost_pressure(const fs2d_sim* s) { return (s->flags & FS2D_FLAG_NO_GHOST_PRESSURE) ? 0 : 1; }

long n_u(const fs2d_sim* s) { return (long)(s->nx + 1) * s->ny; }
long n_v(const fs2d_sim* s) { return (long)s->nx * (s->ny + 1); }
long n_c(const fs2d_sim* s) { return (long)s->nx * s->ny; }

MacView mac_view(const fs2d_sim* s) { return make_mac(s->u, s->v, s->nx, s->ny, s->dx); }

fs2d_status ensure_particle_capacity(fs2d_sim* s, long n) {
    if (n <= s->np_cap) return FS2D_OK;
    // the closest-particle sweeps carry index + 1 in a 29-bit field
    if (n >= (1L << kCandPosShift)) return fail(s, FS2D_ERR_CONFIG, "at most 2^29 - 1 particles");
    float* a[4];
    for (int k = 0; k < 4; k++) TRY(dalloc(s, &a[k], (size_t)n));
    int *key, *sorted;
    TRY(dalloc(s, &key, (size_t)n));
    TRY(dalloc(s, &sorted, (size_t)n));
    s->px = a[0]; s->py = a[1]; s->pvx = a[2]; s->pvy = a[3];
    s->pkey = key; s->sorted = sorted;
    s->np_cap = n;
    return FS2D_OK;
}

// ---------------------------------------------------------------------------
// Stages.
// ---------------------------------------------------------------------------
void sort_particles(fs2d_sim* s) {
    if (s->sorted_valid) return;
    Timed t(s, K_PARTICLE_SORT);
    const long nc = n_c(s);
    fill(s, s->counts, 0, (size_t)nc);
    if (s->np > 0)
        hipLaunchKernelGGL(particle_keys_kernel, dim3(blocks_for(s->np)), dim3(kBlock), 0, s->stream, s->px, s->py,
                           s->np, s->dx, s->nx, s->ny, s->pkey, s->counts);
    const int ntiles = (int)((nc + 1023) / 1024);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(kBlock), 0, s->stream, s->counts, s->starts,
                       s->tile_sums, nc);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(64), 0, s->stream, s->tile_sums, ntiles);
    hipLaunchKernelGGL(scan_add_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->starts, s->tile_sums, nc);
    if (s->np > 0) {
        dcopy(s, s->cursor, s->starts, (size_t)nc);
        hipLaunchKernelGGL(particle_scatter_kernel, dim3(blocks_for(s->np)), dim3(kBlock), 0, s->stream, s->pkey,
                           s->np, s->cursor, s->sorted);
        hipLaunchKernelGGL(segment_sort_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->starts,
                           s->counts, s->sorted, nc);
    }
    s->sorted_valid = true;
}

int extrap_row_blocks(int ny) { return std::max(1, std::min((2 * ny + 1 + 3) / 4, kMaxBlocks)); }
int extrap_col_blocks(int nx) { return (nx + 1 + 63) / 64 + (nx + 63) / 64; }

// Array2D::extrapolate of u and v (include/Array2D.h:552-591), from masks that
// mark the known faces 0 and the others ~0.
void extrapolate_uv(fs2d_sim* s, float* uval, float* vval) {
    Timed t(s, K_EXTRAPOLATE);
    const size_t nU = n_u(s);
    ExtraArr A{G(uval),     G(s->mu),  G(s->fu0), G(s->fpos), G((i4v*)s->fsrc), G((f4v*)s->fkv), G(s->fu1), G(s->fcount + 0),
               s->nx + 1, s->ny};
    ExtraArr B{G(vval),    G(s->mv),        G(s->fv0),      G(s->fpos + nU), G((i4v*)(s->fsrc + nU)),
               G((f4v*)(s->fkv + nU)), G(s->fv1), G(s->fcount + 1), s->nx,         s->ny + 1};
    const int nb = blocks_for(n_u(s) + n_v(s));
    hipLaunchKernelGGL(extrap_rows_kernel, dim3(extrap_row_blocks(s->ny)), dim3(kBlock), 0, s->stream, A, B);
    hipLaunchKernelGGL(extrap_cols_kernel, dim3(extrap_col_blocks(s->nx)), dim3(kExtraColThreads), 0, s->stream, A,
                       B);
    hipLaunchKernelGGL(extrap_hist_kernel, dim3(nb), dim3(kBlock), 0, s->stream, A, B);
    hipLaunchKernelGGL(extrap_scan_kernel, dim3(2), dim3(kBlock), 0, s->stream, A, B);
    hipLaunchKernelGGL(extrap_scatter_kernel, dim3(nb), dim3(kBlock), 0, s->stream, A, B);
    hipLaunchKernelGGL(extrap_link_kernel, dim3(nb), dim3(kBlock), 0, s->stream, A, B);
    hipLaunchKernelGGL(extrap_values_kernel, dim3(2), dim3(1024), 0, s->stream, A, B);
}

void diagnostics(fs2d_sim* s) {
    Timed t(s, K_DIAGNOSTICS);
    const long n = n_c(s) + s->np;
    const double pmass = s->k > 0 ? s->rho * s->dx * s->dx / (s->k * s->k) : 0.0;
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(diagnostics_kernel, dim3(nb), dim3(kBlock), 0, s->stream, s->cell, mac_view(s), s->px,
                       s->py, s->pvx, s->pvy, s->np, s->rho, s->gx, s->gy, pmass, s->partials);
    hipLaunchKernelGGL(diag_finalize_kernel, dim3(1), dim3(kBlock), 0, s->stream, s->partials, nb, s->dx, s->dg);
}

template <class Op, bool B>
void launch_sweep(fs2d_sim* s, const Op& op);
int sk_blocks(const fs2d_sim* s);
int pcg_blocks(const fs2d_sim* s);
template <int FX, int FY>
void launch_closest(fs2d_sim* s, int k) {
    launch_sweep<ClosestSweepOp<FX, FY>, false>(
        s, ClosestSweepOp<FX, FY>{{s->r, s->z, s->q}, nullptr, {s->w}, s->nx, s->ny, s->dx, s->dr, s->ls_chg, k});
}

// FluidSim2D::createWaterLevelSet (src/FluidSim2D.cpp:653-732).
void stage_level_set(fs2d_sim* s) {
    if (s->flags & FS2D_FLAG_PARTICLE_CELLS) {
        // createWaterLevelSet without USE_LEVEL_SET (src/FluidSim2D.cpp:665-706)
        Timed t(s, K_LS_OTHER);
        const long nc = n_c(s);
        hipLaunchKernelGGL(pc_clear_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->cell, s->cell2, nc);
        if (s->np > 0)
            hipLaunchKernelGGL(pc_mark_kernel, dim3(blocks_for(s->np)), dim3(kBlock), 0, s->stream, s->px, s->py,
                               s->np, s->cell, s->cell2, s->nx, s->ny);
        hipLaunchKernelGGL(pc_fill_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->cell, s->cell2,
                           s->nx, s->ny);
        dcopy(s, s->cell, (const uint8_t*)s->cell2, (size_t)nc);
    } else if (!(s->flags & FS2D_FLAG_STATIC_CELLS)) {
        const long nc = n_c(s);
        sort_particles(s);
        {
            Timed t(s, K_LS_OTHER);
            hipLaunchKernelGGL(ls_seed_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->px, s->py,
                               s->starts, s->counts, s->sorted, s->t0, s->nx, nc, s->dx, s->dr);
        }
        // 16 closest-particle passes (direction d = k mod 4: fx = d & 1, fy = d
        // >> 1) on the sweep engine: candidates in SK4 order of the pass's
        // flipped coordinates (the solver's scratch: r, z, q), read from t0
        // (row-major) for the first pass and from the previous pass's output
        // (w, SK4 order of its direction) after; the passes after the first
        // one that changes nothing return at once (ls_first_unchanged); the
        // last output back to t0
        fill(s, s->ls_chg, 0, 16);
        for (int k = 0; k < 16; k++) {
            const int d = k & 3, pd = (k + 3) & 3;
            {
                Timed t(s, K_LS_OTHER);
                hipLaunchKernelGGL(ls_cand_sk_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->geom,
                                   k == 0 ? s->t0 : nullptr, (const float*)s->w, pd & 1, pd >> 1, s->px, s->py, s->r,
                                   s->z, s->q, d & 1, d >> 1, s->dx, s->dr, s->ls_chg, k);
            }
            Timed t(s, K_LS_SWEEP);
            switch (d) {
                case 0: launch_closest<0, 0>(s, k); break;
                case 1: launch_closest<1, 0>(s, k); break;
                case 2: launch_closest<0, 1>(s, k); break;
                default: launch_closest<1, 1>(s, k); break;
            }
        }
        int* const tin = s->t0;
        {
            Timed t(s, K_LS_OTHER);
            hipLaunchKernelGGL(closest_t_to_rm_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->geom,
                               s->w, tin, (const int*)s->ls_chg);
        }
        double* a = s->phi;
        double* b = s->phi2;
        {
            Timed t(s, K_LS_OTHER);
            hipLaunchKernelGGL(ls_phi_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, tin, s->px, s->py,
                               b, s->nx, nc, s->dx, s->dr);
            hipLaunchKernelGGL(ls_surface_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, b, a, s->nx,
                               s->ny);
        }
        // 16 eikonal passes (directions d = k mod 4: fx = d & 1, fy = d >> 1) on
        // the sweep engine, in SK4 order of each direction's flipped
        // coordinates, in the solver's SK4 scratch (z, q, cpl in; w, s1 out).
        // Each pass's output moves straight to the next pass's order; phi is
        // row-major only before the first pass (in a) and after the last (in
        // a again, where the stages after redistance read it).
        {
            Timed t(s, K_LS_OTHER);
            hipLaunchKernelGGL(eik_to_sk_kernel, dim3(sk_blocks(s)), dim3(kBlock), 0, s->stream, s->geom, a, s->z,
                               s->q, s->cpl, 0, 0);
        }
        for (int k = 0; k < 16; k++) {
            const int d = k & 3, d2 = (k + 1) & 3;
            {
                Timed t(s, K_LS_SWEEP);
                launch_sweep<EikSweepOp, false>(s, EikSweepOp{{s->z, s->q}, s->cpl, {s->w, s->s[1]}, s->dx});
            }
            Timed t(s, K_LS_OTHER);
            if (k < 15)
                hipLaunchKernelGGL(eik_sk_to_sk_kernel, dim3(sk_blocks(s)), dim3(kBlock), 0, s->stream, s->geom,
                                   s->w, s->s[1], s->z, s->q, s->cpl, d & 1, d >> 1, d2 & 1, d2 >> 1);
            else
                hipLaunchKernelGGL(eik_from_sk_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->geom,
                                   s->w, s->s[1], a, d & 1, d >> 1);
        }
        {
            Timed t(s, K_LS_OTHER);
            for (int rep = 0; rep < 2; rep++) {
                hipLaunchKernelGGL(ls_smooth_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, a, b, s->nx,
                                   s->ny);
                std::swap(a, b);
            }
            // 16 sweeps and 2 smoothing passes: the result is back in a == s->phi
            hipLaunchKernelGGL(classify_kernel, dim3(blocks_for(nc)), dim3(kBlock), 0, s->stream, s->cell, a, nc);
        }
    }
    diagnostics(s);
}

// applySemiLagrangianAdvection (src/FluidSim2D.cpp:206-235), optionally fused
// with applyGravity.
void stage_sl(fs2d_sim* s, bool with_gravity) {
    Timed t(s, K_SL_ADVECT);
    const bool dens = (s->flags & FS2D_FLAG_DENSITY) != 0;
    const float gu = with_gravity ? (float)(s->dt * s->gx) : 0.0f;
    const float gv = with_gravity ? (float)(s->dt * s->gy) : 0.0f;
    hipLaunchKernelGGL(sl_advect_kernel, dim3(sl_tiles(s->nx, s->ny)), dim3(kBlock), 0, s->stream, mac_view(s), s->u2,
                       s->v2, dens ? s->dens : nullptr, s->dens2, s->dt, clamp_box(s->nx, s->ny, s->dx), gu, gv);
    std::swap(s->u, s->u2);
    std::swap(s->v, s->v2);
    if (dens) std::swap(s->dens, s->dens2);
}

void stage_gravity(fs2d_sim* s) {
    Timed t(s, K_GRAVITY);
    hipLaunchKernelGGL(gravity_kernel, dim3(blocks_for(n_u(s) + n_v(s))), dim3(kBlock), 0, s->stream, s->u, s->v,
                       n_u(s), n_v(s), (float)(s->dt * s->gx), (float)(s->dt * s->gy));
}

// transferVelocityToGrid (src/FluidSim2D.cpp:144-204).
void stage_p2g(fs2d_sim* s) {
    sort_particles(s);
    {
        Timed t(s, K_P2G);
        hipLaunchKernelGGL(p2g_kernel, dim3(blocks_for(n_u(s) + n_v(s))), dim3(kBlock), 0, s->stream, s->px, s->py,
                           s->pvx, s->pvy, s->starts, s->counts, s->sorted, s->u, s->v, s->mu, s->mv, s->nx, s->ny,
                           s->dx);
    }
    extrapolate_uv(s, s->u, s->v);
}

// applyProjection (src/FluidSim2D.cpp:252-467).  Per iteration: SpMV with the
// s update, the p/r update, forward and backward solves; scalars travel in
// device memory, and the host reads the stopping flag once per batch.
SkCtl sk_ctl(fs2d_sim* s) { return SkCtl{s->sk_ticket, s->sk_handoff, s->werr}; }

template <class Op, bool B>
void launch_sweep(fs2d_sim* s, const Op& op) {
    hipLaunchKernelGGL((sweep_kernel<Op, B>), dim3(s->geom.S), dim3(kSweepThreads), 0, s->stream, op, s->geom,
                       sk_ctl(s));
}

// Workgroups of the SK4 stencil kernels: a multiple of 8 (see xcd_loop).
int sk_blocks(const fs2d_sim* s) { return std::max(8, blocks_for(s->geom.size()) / 8 * 8); }
// Workgroups of the PCG SpMV and axpy, a multiple of 8 (xcd_loop): about 8
// entries per thread, at most kMaxBlocks, but no fewer than sk_blocks' one
// entry per thread would give up to 256.  Each workgroup pays partial
// reductions and barriers, so one entry per thread made the batched step
// reduction-bound; a lone small grid still wants its 256 workgroups for
// latency (a 256^2 solve took 17% longer at 40).  The w.s and |r|_inf
// partial counts follow this number, in the single and the batched step alike.
int pcg_blocks(const fs2d_sim* s) {
    const long per = (long)kBlock * 8;
    const int want = (int)std::min<long>(kMaxBlocks, (s->geom.size() + per - 1) / per) / 8 * 8;
    return std::max(std::max(8, want), std::min(sk_blocks(s), 256));
}

// The solver works on A / scaleA and rhs / scaleA, scaleA = dt / (rho dx^2):
// the same solution, the same iterates in exact arithmetic (every CG scalar
// and the relative stopping test are scale-free), but in FP32 the
// off-diagonals are exactly -1 and an interior cell's diagonal exactly its
// count of FLUID neighbours.  With scaleA folded into FP32 coefficients, the
// cells next to walls (3 * scaleA) round differently from interior ones
// (4 * scaleA, exact): a perturbation of one sign along every wall, which at
// 512^2 moved the pressure's smoothest mode by 2e-5 of the solution.
// (dt = 0 makes the reference's matrix zero; its velocity update then adds
// nothing whatever p is, and a unit scale keeps p finite.)
double solver_scale(const fs2d_sim* s) { return s->dt > 0 ? s->dt / (s->rho * s->dx * s->dx) : 1.0; }

// Ghost-fluid matrix and rhs of the current velocity (src/FluidSim2D.cpp:252-362),
// in solver units.
void assemble(fs2d_sim* s) {
    Timed t(s, K_ASSEMBLE);
    hipLaunchKernelGGL(assemble_sk_kernel, dim3(sk_blocks(s)), dim3(kBlock), 0, s->stream, s->geom, s->cell, s->phi,
                       s->u, s->v, s->Adiag, s->cpl, s->r, s->pk, s->s[0], 1.0 / (s->dx * solver_scale(s)),
                       s->rmax_part, ghost_pressure(s));
}

// Physical workgroups of the PCG SpMV and axpy launches: their loops run the
// pcg_blocks virtual workgroups (pcg_grid 0: one each), so this choice
// changes no result; a multiple of 8 (xcd_loop).
int pcg_launch(const fs2d_sim* s) {
    const int nvb = pcg_blocks(s);
    return s->pcg_grid > 0 ? std::min(nvb, s->pcg_grid) : nvb;
}

// The backward solve of iteration h (next_iter = h + 1; 0 for the initial
// solve) and the SpMV of iteration h + 1 in one launch (bwd_spmv_kernel):
// sigma_{h+1} goes to sig_out, beta = sigma_{h+1} / sigma_h with sigma_h in
// sig_prev (unused when first), s_new = z + beta s_old, w = A s_new.
// Workgroups: K for sigma (one (strip, loader) pair per wave), the S strips,
// and M for the SpMV, which together fill the device once (one workgroup per
// CU fits).  More SpMV workgroups than that measured far slower: a workgroup
// waits for a CU of its own XCD, and on XCDs whose CUs all hold strips or
// other SpMV workgroups its statically assigned runs start only at the end.
void launch_bwd_spmv(fs2d_sim* s, int next_iter, double* sig_out, const double* sig_prev, int first,
                     const float* s_old, float* s_new, int check) {
    const SkGeom& g = s->geom;
    const int waves = kSweepThreads / 64;
    const int K = (g.S * kLoaders + waves - 1) / waves;
    const int nbp = pcg_blocks(s);
    const int M = std::max(8, std::min(nbp, (s->ncu > 0 ? s->ncu : 256) - g.S - K));
    BwdFusedOp op{{s->q, s->P, s->gx_, s->gy_}, nullptr, {s->z}, s->sc, nullptr, next_iter, s->fb_flags, 0u,
                  check, s->rmax_part, nbp};
    FusedSpmvArgs f{s->Adiag, s->cpl,     s->z, s->q,         s_old, s_new, s->w, s->sc, sig_out, s->sig_words,
                    sig_prev, first,      s->ws_part, nbp, s->fb_flags, K, M, check, s->rmax_part};
    hipLaunchKernelGGL(bwd_spmv_kernel, dim3(K + g.S + M), dim3(kSweepThreads), 0, s->stream, op, g,
                       SkCtl{s->fb_ticket, s->fb_handoff, s->werr}, f);
}

// The p/r update of iteration h (s' in s_new, alpha = sigma_h / w.s' with
// sigma_h in sig_cur) and its forward solve in one launch (fwd_axpy_kernel).
// The update's workgroups sit beside the S strips: enough for two strips'
// runs at a time, which keeps it far ahead of the solve's wavefront; more
// (every free CU) measured slower at 2048^2 (158.6 against 154.1 us), the
// update's burst of traffic slowing the first strips.  The stopping test of
// iteration h runs in the next backward launch.
void launch_fwd_axpy(fs2d_sim* s, const float* s_new, const double* sig_cur) {
    const SkGeom& g = s->geom;
    const int nbp = pcg_blocks(s);
    const int per_strip = (nbp + g.S - 1) / g.S;
    const int M = std::max(8, std::min({nbp, 2 * per_strip, (s->ncu > 0 ? s->ncu : 256) - g.S}));
    FwdFusedOp op{{s->r, s->P, s->fx, s->fy}, nullptr, {s->q}, s->sc, s->rmax_part, nbp, 0, s->fa_flags, g, 0u};
    FusedAxpyArgs f{s->pk, s->r, s_new, s->w, s->sc, sig_cur, s->ws_part, s->rmax_part, s->fa_flags, nbp, M};
    hipLaunchKernelGGL(fwd_axpy_kernel, dim3(M + g.S), dim3(kSweepThreads), 0, s->stream, op, g,
                       SkCtl{s->fa_ticket, s->fa_handoff, s->werr}, f);
}

// MIC(0) factor and the first preconditioner application of the PCG
// (src/FluidSim2D.cpp:364-430) on the assembled matrix and the rhs in r (with
// p = s = 0 and the |rhs|_inf partials in rmax_part).  With fused_pcg the
// backward solve's launch also runs iteration 0's SpMV.
void pcg_begin(fs2d_sim* s) {
    const int nb = sk_blocks(s);
    hipLaunchKernelGGL(pcg_init_kernel, dim3(1), dim3(64), 0, s->stream, s->sc, s->tol);
    {
        Timed t(s, K_FACTOR);
        // the reference skips the solve when |rhs|_inf <= 1e-12 (:447-448); r holds rhs / scaleA
        FactorOp op{{s->Adiag}, s->cpl, {s->P, s->fx, s->fy, s->gx_, s->gy_}, s->tau, s->sigma, s->rmax_part, nb,
                    s->max_iters, s->sc, 1e-12 / solver_scale(s)};
        launch_sweep<FactorOp, false>(s, op);
    }
    {
        Timed t(s, K_PCG_FWD);
        FwdOp op{{s->r, s->P, s->fx, s->fy}, nullptr, {s->q}, s->sc, s->rmax_part, nb, 0};
        launch_sweep<FwdOp, false>(s, op);
    }
    {
        Timed t(s, K_PCG_BWD);
        if (s->fused_pcg) {
            launch_bwd_spmv(s, 0, s->qq, nullptr, 1, s->s[0], s->s[1], 0);
        } else {
            BwdOp op{{s->q, s->P, s->gx_, s->gy_}, nullptr, {s->z}, s->sc, s->qq, 0};
            launch_sweep<BwdOp, true>(s, op);
        }
    }
}

// Iterations h0 .. h1-1 of the PCG loop (:431-463): SpMV with the s update,
// the p/r update, forward and backward solves.  Every kernel returns at once
// once the stopping test has passed, so a fixed number of iterations can be
// queued (or captured) without host round trips.  With fused_pcg iteration
// h's SpMV ran in the previous backward solve's launch; this iteration is the
// p/r update with the forward solve, then the backward solve (after the
// stopping test) with iteration h + 1's SpMV.
void pcg_iterations(fs2d_sim* s, int h0, int h1) {
    const SkGeom& g = s->geom;
    double* qq0 = s->qq;
    double* qq1 = s->qq + (size_t)g.S * kLoaders;
    const int nbp = pcg_blocks(s), grid = pcg_launch(s);
    for (int h = h0; h < h1; h++) {
        float* s_old = s->s[h & 1];
        float* s_new = s->s[(h + 1) & 1];
        double* sig_cur = (h & 1) ? qq1 : qq0;   // sigma_h
        double* sig_prev = (h & 1) ? qq0 : qq1;  // sigma_{h-1}; receives sigma_{h+1}
        if (s->fused_pcg) {
            {
                Timed t(s, K_PCG_FWD);
                launch_fwd_axpy(s, s_new, sig_cur);
            }
            Timed t(s, K_PCG_BWD);
            launch_bwd_spmv(s, h + 1, sig_prev, sig_cur, 0, s_new, s_old, 1);
            continue;
        }
        {
            Timed t(s, K_PCG_SPMV);
            hipLaunchKernelGGL(pcg_spmv_sk_kernel, dim3(grid), dim3(kBlock), 0, s->stream, g, s->Adiag, s->cpl, s->z,
                               s_old, s_new, s->w, s->sc, sig_cur, sig_prev, h == 0 ? 1 : 0, s->ws_part,
                               nbp);
        }
        {
            Timed t(s, K_PCG_AXPY);
            hipLaunchKernelGGL(pcg_axpy_sk_kernel, dim3(grid), dim3(kBlock), 0, s->stream, g, s->pk, s->r, s_new,
                               s->w, s->sc, sig_cur, s->ws_part, nbp, s->rmax_part, nbp);
        }
        {
            Timed t(s, K_PCG_FWD);
            FwdOp op{{s->r, s->P, s->fx, s->fy}, nullptr, {s->q}, s->sc, s->rmax_part, nbp, 1};
            launch_sweep<FwdOp, false>(s, op);
        }
        {
            Timed t(s, K_PCG_BWD);
            BwdOp op{{s->q, s->P, s->gx_, s->gy_}, nullptr, {s->z}, s->sc, sig_prev, h + 1};
            launch_sweep<BwdOp, true>(s, op);
        }
    }
}

// The solution, row-major.
void pcg_end(fs2d_sim* s, float* out) {
    Timed t(s, K_MISC);
    hipLaunchKernelGGL(sk_to_rowmajor_kernel, dim3(blocks_for(n_c(s))), dim3(kBlock), 0, s->stream, s->geom, s->pk,
                       out);
}

// The whole PCG: the host reads the stopping flag once per chunk of queued
// iterations (the first chunk sized by the previous solve's count).
fs2d_status pcg_solve(fs2d_sim* s, float* out) {
    pcg_begin(s);
    int h = 0;
    int batch = std::max(2, std::min(s->last_iters_hint, s->max_iters));
    while (h < s->max_iters) {
        const int nbatch = std::min(batch, s->max_iters - h);
        pcg_iterations(s, h, h + nbatch);
        h += nbatch;
        HIPCHK(s, hipMemcpyAsync(s->h_sc, s->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        if (s->h_sc->done) break;
        batch = 4;
    }
    pcg_end(s, out);
    TRY(d2h(s, s->h_sc, s->sc, sizeof(PcgScalars)));
    s->last_iters_hint = std::max(2, s->h_sc->iter + 1);
    return s->h_sc->iter >= s->max_iters ? FS2D_WARN_MAXITER : FS2D_OK;
}

fs2d_status stage_projection(fs2d_sim* s) {
    assemble(s);
    return pcg_solve(s, s->p);
}

// updateVelocity (src/FluidSim2D.cpp:469-550).
void stage_update_velocity(fs2d_sim* s) {
    {
        Timed t(s, K_UPDATE_VEL);
        hipLaunchKernelGGL(update_velocity_kernel, dim3(blocks_for(n_u(s) + n_v(s))), dim3(kBlock), 0, s->stream,
                           s->cell, s->phi, s->p, s->u, s->v, s->nu, s->nv, s->mu, s->mv, s->nx, s->ny,
                           s->dt / (s->rho * s->dx), ghost_pressure(s));
    }
    extrapolate_uv(s, s->nu, s->nv);
    if (s->mode == FS2D_SEMILAGRANGIAN) {
        Timed t(s, K_MISC);
        dcopy(s, s->u, s->nu, (size_t)n_u(s));
        dcopy(s, s->v, s->nv, (size_t)n_v(s));
    }
}

// updateParticleVelocities (src/FluidSim2D.cpp:552-568).
void stage_g2p(fs2d_sim* s) {
    if (s->np > 0) {
        Timed t(s, K_G2P);
        hipLaunchKernelGGL(g2p_kernel, dim3(blocks_for(s->np)), dim3(kBlock), 0, s->stream, s->px, s->py, s->pvx,
                           s->pvy, s->np, s->nu, s->nv, s->u, s->v, s->nx, s->ny, s->dx, (float)s->alpha);
    }
    Timed t(s, K_MISC);
    dcopy(s, s->u, s->nu, (size_t)n_u(s));
    dcopy(s, s->v, s->nv, (size_t)n_v(s));
}

// applyAdvection (src/FluidSim2D.cpp:570-604).
void stage_particle_advect(fs2d_sim* s) {
    if (s->np == 0) return;
    Timed t(s, K_PARTICLE_ADVECT);
    const int nb = blocks_for(s->np);
    hipLaunchKernelGGL(particle_advect_kernel, dim3(nb), dim3(kBlock), 0, s->stream, s->px, s->py, s->pvx, s->pvy,
                       s->np, mac_view(s), s->dt, clamp_box(s->nx, s->ny, s->dx), s->dg, s->partials,
                       (s->flags & FS2D_FLAG_RK2) ? 1 : 0);
    hipLaunchKernelGGL(cfl_finalize_kernel, dim3(1), dim3(kBlock), 0, s->stream, s->partials, nb, s->dg);
    s->sorted_valid = false;
}

// Status after a sequence of stages: one host round trip.
fs2d_status finish(fs2d_sim* s, fs2d_status st) {
    HIPCHK(s, hipMemcpyAsync(s->h_dg, s->dg, sizeof(Diag), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->h_err, s->werr, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipGetLastError());
    if (s->timing) resolve_timing(s);
    if (*s->h_err) return fail(s, FS2D_ERR_HANG, "a wavefront hand-off wait expired");
    if (s->h_dg->nan_flag) return fail(s, FS2D_ERR_NAN, "advected particle position is NaN");
    if (st == FS2D_WARN_MAXITER) {
        fail(s, st, "PCG: maximum iteration limit (%d) exceeded", s->max_iters);
        return st;
    }
    if (s->np > 0 && s->h_dg->cfl_max > 5.0) {
        fail(s, FS2D_WARN_CFL, "particle velocity violates CFL condition (C=%f)", s->h_dg->cfl_max);
        return FS2D_WARN_CFL;
    }
    return st;
}

fs2d_status run_stage_impl(fs2d_sim* s, int stage, bool fuse_gravity, fs2d_status* worst) {
    fs2d_status st = FS2D_OK;
    switch (stage) {
        case FS2D_STAGE_INIT: break;
        case FS2D_STAGE_CREATE_WATER_LEVEL_SET: stage_level_set(s); break;
        case FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID: stage_p2g(s); break;
        case FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION: stage_sl(s, fuse_gravity); break;
        case FS2D_STAGE_APPLY_GRAVITY: stage_gravity(s); break;
        case FS2D_STAGE_CREATE_SOLID_LEVEL_SET: break;  // no-op (src/FluidSim2D.cpp:734-736)
        case FS2D_STAGE_APPLY_PROJECTION: st = stage_projection(s); break;
        case FS2D_STAGE_UPDATE_VELOCITY: stage_update_velocity(s); break;
        case FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES: stage_g2p(s); break;
        case FS2D_STAGE_APPLY_ADVECTION: stage_particle_advect(s); break;
        default: return fail(s, FS2D_ERR_ARG, "unknown stage %d", stage);
    }
    if (st < 0) return st;
    if (st > *worst) *worst = st;
    return FS2D_OK;
}

// ---------------------------------------------------------------------------
// Captured steps.
// ---------------------------------------------------------------------------
GraphKey graph_key(const fs2d_sim* s) {
    GraphKey k;
    memset(&k, 0, sizeof k);  // padding too: keys are compared bytewise
    k.dt = s->dt; k.gx = s->gx; k.gy = s->gy; k.alpha = s->alpha;
    k.tau = s->tau; k.sigma = s->sigma; k.tol = s->tol;
    k.max_iters = s->max_iters; k.budget = s->graph_budget; k.pcg_grid = s->pcg_grid;
    k.sorted_valid = s->sorted_valid ? 1 : 0;
    k.np = s->np;
    const void* p[12] = {s->u, s->v, s->u2, s->v2, s->dens, s->dens2, s->px, s->py, s->pvx, s->pvy, s->pkey, s->sorted};
    memcpy(k.ptr, p, sizeof p);
    return k;
}
GraphAfter graph_after(const fs2d_sim* s) {
    return GraphAfter{s->u, s->v, s->u2, s->v2, s->dens, s->dens2, s->sorted_valid};
}
void graph_restore(fs2d_sim* s, const GraphAfter& a) {
    s->u = a.u; s->v = a.v; s->u2 = a.u2; s->v2 = a.v2; s->dens = a.dens; s->dens2 = a.dens2;
    s->sorted_valid = a.sorted_valid;
}
void graph_free(StepGraph& G) {
    if (G.pre) (void)hipGraphExecDestroy(G.pre);
    if (G.post) (void)hipGraphExecDestroy(G.post);
    G.pre = G.post = nullptr;
}

// Stages of fs2d_step before and after applyProjection (runFrame,
// src/FluidSim2D.cpp:100-133; gravity fused into the SL advection).
void step_stages(const fs2d_sim* s, std::vector<int>& before, std::vector<int>& after) {
    if (s->mode == FS2D_SEMILAGRANGIAN) {
        before = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION,
                  FS2D_STAGE_CREATE_SOLID_LEVEL_SET};
        after = {FS2D_STAGE_UPDATE_VELOCITY, FS2D_STAGE_APPLY_ADVECTION};
    } else {
        before = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID, FS2D_STAGE_APPLY_GRAVITY,
                  FS2D_STAGE_CREATE_SOLID_LEVEL_SET};
        after = {FS2D_STAGE_UPDATE_VELOCITY, FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES, FS2D_STAGE_APPLY_ADVECTION};
    }
}

fs2d_status end_capture(fs2d_sim* s, hipGraphExec_t* exec) {
    hipGraph_t g = nullptr;
    HIPCHK(s, hipStreamEndCapture(s->stream, &g));
    hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(s, FS2D_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return FS2D_OK;
}

// Captures the step from the handle's current host state: `pre` runs the
// stages before the projection, the assembly, the MIC(0) factor and
// graph_budget PCG iterations; `post` the solution's copy and the stages
// after the projection.  Nothing runs; the host state is left as it was.
fs2d_status capture_step(fs2d_sim* s, StepGraph& G) {
    G.key = graph_key(s);
    G.budget = s->graph_budget;
    const GraphAfter before = graph_after(s);
    std::vector<int> st_before, st_after;
    step_stages(s, st_before, st_after);
    fs2d_status worst = FS2D_OK, st = FS2D_OK;
    HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const bool fuse = s->mode == FS2D_SEMILAGRANGIAN;  // gravity inside the SL advection, as fs2d_step
    for (int stg : st_before)
        if (st == FS2D_OK) st = run_stage_impl(s, stg, fuse, &worst);
    if (st == FS2D_OK) {
        assemble(s);
        pcg_begin(s);
        pcg_iterations(s, 0, G.budget);
    }
    fs2d_status ce = end_capture(s, &G.pre);
    if (st == FS2D_OK && ce == FS2D_OK) {
        HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        pcg_end(s, s->p);
        for (int stg : st_after)
            if (st == FS2D_OK) st = run_stage_impl(s, stg, fuse, &worst);
        ce = end_capture(s, &G.post);
    }
    G.after = graph_after(s);
    graph_restore(s, before);
    if (st != FS2D_OK || ce != FS2D_OK) {
        graph_free(G);
        return st != FS2D_OK ? st : ce;
    }
    return FS2D_OK;
}

// PCG iterations a captured step runs before its host check: the last
// solve's count plus a sixteenth and 4, in steps of 32, kept while it
// suffices and is at most twice what is needed.
void update_graph_budget(fs2d_sim* s, int needed) {
    const int want = std::min(s->max_iters, std::max(16, (needed + needed / 16 + 4 + 31) / 32 * 32));
    if (s->graph_budget <= 0 || needed > s->graph_budget || 2 * needed < s->graph_budget) s->graph_budget = want;
}

// fs2d_step by replaying a captured step (capturing it first when the host
// state is new).  Convergence within the budget leaves one host round trip
// in the step besides finish(); a solve that needs more iterations continues
// with queued chunks, as the uncaptured step does.
fs2d_status step_graph(fs2d_sim* s) {
    if (s->graph_budget <= 0) update_graph_budget(s, s->last_iters_hint);
    s->graph_budget = std::min(s->graph_budget, s->max_iters);
    const GraphKey key = graph_key(s);
    StepGraph* G = nullptr;
    for (auto& g : s->graphs)
        if (g.key == key) G = &g;
    if (!G) {
        if (s->graphs.size() >= 4) {  // two buffer parities x two budgets
            graph_free(s->graphs.front());
            s->graphs.erase(s->graphs.begin());
        }
        StepGraph ng;
        TRY(capture_step(s, ng));
        s->graphs.push_back(ng);
        G = &s->graphs.back();
    }
    HIPCHK(s, hipGraphLaunch(G->pre, s->stream));
    TRY(d2h(s, s->h_sc, s->sc, sizeof(PcgScalars)));
    int h = G->budget;
    while (!s->h_sc->done && h < s->max_iters) {
        const int n = std::min(8, s->max_iters - h);
        pcg_iterations(s, h, h + n);
        h += n;
        TRY(d2h(s, s->h_sc, s->sc, sizeof(PcgScalars)));
    }
    HIPCHK(s, hipGraphLaunch(G->post, s->stream));
    graph_restore(s, G->after);
    const int iters = s->h_sc->iter;
    s->last_iters_hint = std::max(2, iters + 1);
    update_graph_budget(s, iters);
    s->current_time += s->dt;
    s->steps++;
    return finish(s, iters >= s->max_iters ? FS2D_WARN_MAXITER : FS2D_OK);
}

bool valid_dtype(int field, int dtype) {
    if (field == FS2D_FIELD_CELL) return dtype == FS2D_U8;
    if (field == FS2D_FIELD_PARTICLE_CELL || field == FS2D_FIELD_PARTICLE_ORDER) return dtype == FS2D_I32;
    return dtype == FS2D_F32 || dtype == FS2D_F64;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

fs2d_status fs2d_config_init(fs2d_config* c) {
    if (!c) return FS2D_ERR_ARG;
    memset(c, 0, sizeof *c);
    c->particles_per_cell_sqrt = 2;
    c->dt = 0.0001;             // FluidSim2D.h:89
    c->dx = 0.001;              // FluidSim2D.h:90
    c->rho = 997.0;             // FluidSim2D.h:82
    c->gravity_x = 0.0;         // FluidSim2D.h:80
    c->gravity_y = -9.81;
    c->mode = FS2D_PICFLIP;
    c->pic_flip_alpha = 1.0;
    c->tau = 0.999;             // FluidSim2D.cpp:369
    c->sigma = 0.25;            // FluidSim2D.cpp:370
    c->pcg_tol = 1e-12;         // FluidSim2D.cpp:453
    c->pcg_max_iters = 200;     // FluidSim2D.cpp:429
    c->seed = 1;                // glibc's default rand() seed
    c->flags = 0;
    return FS2D_OK;
}

static fs2d_status validate(const fs2d_config* c, char* why, size_t n) {
    if (c->size_x < 3 || c->size_y < 3 || c->size_x > 16384 || c->size_y > 16384) {
        snprintf(why, n, "size %dx%d outside [3, 16384]", c->size_x, c->size_y);
        return FS2D_ERR_CONFIG;
    }
    if (!(c->dx > 0) || !(c->dt >= 0) || !(c->rho > 0) || !std::isfinite(c->dt)) {
        snprintf(why, n, "dx, rho must be > 0 and dt >= 0");
        return FS2D_ERR_CONFIG;
    }
    if (c->particles_per_cell_sqrt < 0 || c->particles_per_cell_sqrt > 16) {
        snprintf(why, n, "particles_per_cell_sqrt %d outside [0, 16]", c->particles_per_cell_sqrt);
        return FS2D_ERR_CONFIG;
    }
    if (c->mode != FS2D_SEMILAGRANGIAN && c->mode != FS2D_PICFLIP) {
        snprintf(why, n, "mode %d", c->mode);
        return FS2D_ERR_CONFIG;
    }
    if (c->pcg_max_iters < 0 || !(c->tau >= 0) || !(c->sigma >= 0) || !(c->pcg_tol >= 0)) {
        snprintf(why, n, "solver parameters");
        return FS2D_ERR_CONFIG;
    }
    if (c->flags & ~(FS2D_FLAG_STATIC_CELLS | FS2D_FLAG_DENSITY | FS2D_FLAG_PARTICLE_CELLS |
                     FS2D_FLAG_NO_GHOST_PRESSURE | FS2D_FLAG_UNIFORM_SEEDING | FS2D_FLAG_RK2)) {
        snprintf(why, n, "unknown flags 0x%x", c->flags);
        return FS2D_ERR_CONFIG;
    }
    if ((c->flags & FS2D_FLAG_PARTICLE_CELLS) &&
        ((c->flags & FS2D_FLAG_STATIC_CELLS) || !(c->flags & FS2D_FLAG_NO_GHOST_PRESSURE))) {
        // without a level set the ghost-fluid terms would divide phi = 0 by phi = 0
        snprintf(why, n, "FS2D_FLAG_PARTICLE_CELLS needs FS2D_FLAG_NO_GHOST_PRESSURE and no FS2D_FLAG_STATIC_CELLS");
        return FS2D_ERR_CONFIG;
    }
    if (!c->initial_values) {
        snprintf(why, n, "initial_values is NULL");
        return FS2D_ERR_CONFIG;
    }
    const int nx = c->size_x, ny = c->size_y;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const uint8_t f = c->initial_values[(size_t)j * nx + i];
            if (f > FS2D_SOLID) {
                snprintf(why, n, "cell (%d,%d) has flag %u", i, j, f);
                return FS2D_ERR_CONFIG;
            }
            // The reference reads cell(i-1,j) etc. of FLUID cells without bounds
            // checks (FluidSim2D.cpp:267-299); a non-solid border is out of bounds there.
            if ((i == 0 || j == 0 || i == nx - 1 || j == ny - 1) && f != FS2D_SOLID) {
                snprintf(why, n, "border cell (%d,%d) is not SOLID", i, j);
                return FS2D_ERR_CONFIG;
            }
        }
    return FS2D_OK;
}

static thread_local std::string g_create_error;

fs2d_status fs2d_create(const fs2d_config* cfg, fs2d_sim** out) {
    if (!cfg || !out) return FS2D_ERR_ARG;
    *out = nullptr;
    char why[256] = {0};
    fs2d_status vs = validate(cfg, why, sizeof why);
    if (vs != FS2D_OK) {
        g_create_error = why;
        return vs;
    }
    fs2d_sim* s = new fs2d_sim();
    s->nx = cfg->size_x;
    s->ny = cfg->size_y;
    s->k = cfg->particles_per_cell_sqrt;
    s->dt = cfg->dt;
    s->dx = cfg->dx;
    s->dr = 0.9 * cfg->dx;  // FluidSim2D.cpp:32
    s->rho = cfg->rho;
    s->gx = cfg->gravity_x;
    s->gy = cfg->gravity_y;
    s->mode = cfg->mode;
    s->alpha = cfg->pic_flip_alpha;
    s->tau = cfg->tau;
    s->sigma = cfg->sigma;
    s->tol = cfg->pcg_tol;
    s->max_iters = cfg->pcg_max_iters;
    s->flags = cfg->flags;

    auto bail = [&](fs2d_status st) {
        g_create_error = s->err;
        fs2d_destroy(s);
        return st;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        s->err = std::string("HIP device unavailable: ") + hipGetErrorString(e);
        return bail(FS2D_ERR_DEVICE);
    }
    const size_t nU = n_u(s), nV = n_v(s), nC = n_c(s);
    fs2d_status st = FS2D_OK;
#define A(p, n) \
    if ((st = dalloc(s, &s->p, (n))) != FS2D_OK) return bail(st)
    A(u, nU); A(v, nV); A(u2, nU); A(v2, nV); A(nu, nU); A(nv, nV);
    A(dens, nC); A(dens2, nC); A(cell, nC); A(phi, nC); A(phi2, nC); A(t0, nC); A(t1, nC); A(ls_chg, 16);
    if (s->flags & FS2D_FLAG_PARTICLE_CELLS) A(cell2, nC);
    s->geom = make_geom(s->nx, s->ny);
    const size_t nS = (size_t)s->geom.size();
    s->fused_pcg = (long)nS >= kFusedPcgMinEntries;
    A(p, nC);
    A(r, nS); A(z, nS); A(q, nS); A(w, nS); A(s[0], nS); A(s[1], nS); A(pk, nS);
    A(Adiag, nS); A(P, nS); A(fx, nS); A(fy, nS); A(gx_, nS); A(gy_, nS); A(cpl, nS);
    A(sc, 1); A(qq, 2 * (size_t)s->geom.S * kLoaders); A(ws_part, kMaxBlocks); A(rmax_part, kMaxBlocks); A(sk_ticket, 1);
    A(mu, nU); A(mv, nV); A(fu0, nU); A(fv0, nV); A(fcount, 2);
    A(fu1, extra_aux_len(s->nx + 1, s->ny)); A(fv1, extra_aux_len(s->nx, s->ny + 1));
    A(fpos, nU + nV); A(fsrc, nU + nV); A(fkv, nU + nV);
    A(counts, nC); A(starts, nC); A(cursor, nC); A(tile_sums, (nC + 1023) / 1024 + 1);
    A(dg, 1); A(partials, 3 * kMaxBlocks);
    A(werr, 1);
    const int nstrips = (s->ny + 63) / 64;
    A(sk_handoff, (size_t)nstrips * sk_handoff_stride(s->nx, 3));
    A(fb_ticket, 1); A(fb_flags, nstrips); A(sig_words, 2 * (size_t)nstrips * kLoaders);
    A(fb_handoff, (size_t)nstrips * sk_handoff_stride(s->nx, BwdFusedOp::NW));
    A(fa_ticket, 1); A(fa_flags, kMaxBlocks);
    A(fa_handoff, (size_t)nstrips * sk_handoff_stride(s->nx, FwdFusedOp::NW));
#undef A
    if (hipDeviceGetAttribute(&s->ncu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess) s->ncu = 0;
    e = hipHostMalloc((void**)&s->h_sc, sizeof(PcgScalars), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_dg, sizeof(Diag), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_err, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
        s->err = std::string("hipHostMalloc: ") + hipGetErrorString(e);
        return bail(FS2D_ERR_OOM);
    }
    memset(s->h_sc, 0, sizeof(PcgScalars));
    memset(s->h_dg, 0, sizeof(Diag));
    *s->h_err = 0;
    if (hipMemcpyAsync(s->cell, cfg->initial_values, nC, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
        s->err = "initial upload failed";
        return bail(FS2D_ERR_DEVICE);
    }

    // Particle seeding of FluidSim2D::create (src/FluidSim2D.cpp:49-69): k^2
    // jittered samples per FLUID cell in row-major cell order, x jitter drawn
    // before y jitter, with glibc's rand() stream (randf, :19-21).  random_r
    // with a private 128-byte state is that stream without touching the
    // process-wide generator.  FS2D_FLAG_UNIFORM_SEEDING draws each sample
    // uniformly over the whole cell instead: ((i + U) dx, (j + U) dx),
    // U = rand()/RAND_MAX, x before y.
    std::vector<float> hx, hy;
    size_t fluid = 0;
    {
        struct random_data rd;
        char rstate[128];
        memset(&rd, 0, sizeof rd);
        memset(rstate, 0, sizeof rstate);
        initstate_r(cfg->seed, rstate, sizeof rstate, &rd);
        auto randf = [&](float mx) {
            int32_t val;
            random_r(&rd, &val);
            return ((double)val / (double)RAND_MAX) * mx;
        };
        const int k = s->k, ppc = k * k;
        const float dist = k > 0 ? 1.0f / k : 0.0f;
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                if (cfg->initial_values[(size_t)j * s->nx + i] == FS2D_FLUID) {
                    fluid++;
                    for (int e2 = 0; e2 < ppc; e2++) {
                        if (cfg->flags & FS2D_FLAG_UNIFORM_SEEDING) {
                            const double ux = randf(1.0f);
                            const double uy = randf(1.0f);
                            hx.push_back((float)(((double)i + ux) * s->dx));
                            hy.push_back((float)(((double)j + uy) * s->dx));
                            continue;
                        }
                        const int x = e2 % k, y = e2 / k;
                        const double jx = randf(dist);
                        const double jy = randf(dist);
                        const float fx = dist * x, fy = dist * y;
                        hx.push_back((float)(((double)i + fx + jx) * s->dx));
                        hy.push_back((float)(((double)j + fy + jy) * s->dx));
                    }
                }
    }
    s->orig_volume = fluid * s->dx * s->dx;
    s->np = (long)hx.size();
    if ((st = ensure_particle_capacity(s, std::max(s->np, 1L))) != FS2D_OK) return bail(st);
    if (s->np > 0) {
        if ((st = h2d(s, s->px, hx.data(), s->np * sizeof(float))) != FS2D_OK ||
            (st = h2d(s, s->py, hy.data(), s->np * sizeof(float))) != FS2D_OK)
            return bail(st);
    }
    if (s->flags & FS2D_FLAG_STATIC_CELLS) {
        // Static cells: ghost-fluid faces (if any EMPTY cell exists) sit half-way
        // between the centres.
        std::vector<double> ph(nC);
        for (size_t c = 0; c < nC; c++) ph[c] = cfg->initial_values[c] == FS2D_EMPTY ? 0.5 * s->dx : -0.5 * s->dx;
        if ((st = h2d(s, s->phi, ph.data(), nC * sizeof(double))) != FS2D_OK) return bail(st);
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) {
        s->err = "device synchronisation failed";
        return bail(FS2D_ERR_DEVICE);
    }
    *out = s;
    return FS2D_OK;
}

void fs2d_destroy(fs2d_sim* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& g : s->graphs) graph_free(g);
    for (void* p : s->allocs) (void)hipFree(p);
    for (auto& t : s->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : s->event_pool) (void)hipEventDestroy(e);
    if (s->h_sc) (void)hipHostFree(s->h_sc);
    if (s->h_dg) (void)hipHostFree(s->h_dg);
    if (s->h_err) (void)hipHostFree(s->h_err);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

fs2d_status fs2d_step(fs2d_sim* s) {
    if (!s) return FS2D_ERR_ARG;
    if (s->use_graph && !s->timing) return step_graph(s);
    fs2d_status worst = FS2D_OK, st;
    if (s->mode == FS2D_SEMILAGRANGIAN) {
        // createWaterLevelSet, applySemiLagrangianAdvection + applyGravity (fused),
        // createSolidLevelSet, applyProjection, updateVelocity, applyAdvection
        const int stages[] = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION,
                              FS2D_STAGE_CREATE_SOLID_LEVEL_SET, FS2D_STAGE_APPLY_PROJECTION,
                              FS2D_STAGE_UPDATE_VELOCITY, FS2D_STAGE_APPLY_ADVECTION};
        for (int stg : stages)
            if ((st = run_stage_impl(s, stg, true, &worst)) != FS2D_OK) return st;
    } else {
        const int stages[] = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID,
                              FS2D_STAGE_APPLY_GRAVITY, FS2D_STAGE_CREATE_SOLID_LEVEL_SET,
                              FS2D_STAGE_APPLY_PROJECTION, FS2D_STAGE_UPDATE_VELOCITY,
                              FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES, FS2D_STAGE_APPLY_ADVECTION};
        for (int stg : stages)
            if ((st = run_stage_impl(s, stg, false, &worst)) != FS2D_OK) return st;
    }
    s->current_time += s->dt;
    s->steps++;
    return finish(s, worst);
}

fs2d_status fs2d_run_stage(fs2d_sim* s, int stage) {
    if (!s) return FS2D_ERR_ARG;
    fs2d_status worst = FS2D_OK;
    fs2d_status st = run_stage_impl(s, stage, false, &worst);
    if (st != FS2D_OK) return st;
    return finish(s, worst);
}

fs2d_status fs2d_synchronize(fs2d_sim* s) {
    if (!s) return FS2D_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (s->timing) resolve_timing(s);
    return FS2D_OK;
}

size_t fs2d_field_count(const fs2d_sim* s, int field) {
    if (!s) return 0;
    switch (field) {
        case FS2D_FIELD_U: case FS2D_FIELD_NEWU: return (size_t)n_u(s);
        case FS2D_FIELD_V: case FS2D_FIELD_NEWV: return (size_t)n_v(s);
        case FS2D_FIELD_P: case FS2D_FIELD_PHI: case FS2D_FIELD_DENSITY: case FS2D_FIELD_CELL: return (size_t)n_c(s);
        case FS2D_FIELD_PARTICLE_POS: case FS2D_FIELD_PARTICLE_VEL: return (size_t)s->np * 2;
        case FS2D_FIELD_PARTICLE_CELL: case FS2D_FIELD_PARTICLE_ORDER: return (size_t)s->np;
        default: return 0;
    }
}

static float* fp32_field(fs2d_sim* s, int field) {
    switch (field) {
        case FS2D_FIELD_U: return s->u;
        case FS2D_FIELD_V: return s->v;
        case FS2D_FIELD_NEWU: return s->nu;
        case FS2D_FIELD_NEWV: return s->nv;
        case FS2D_FIELD_P: return s->p;
        case FS2D_FIELD_DENSITY: return s->dens;
        default: return nullptr;
    }
}

static bool is_particle_field(int field) {
    return field == FS2D_FIELD_PARTICLE_POS || field == FS2D_FIELD_PARTICLE_VEL ||
           field == FS2D_FIELD_PARTICLE_CELL || field == FS2D_FIELD_PARTICLE_ORDER;
}

fs2d_status fs2d_get_field(fs2d_sim* s, int field, int dtype, void* host, size_t count) {
    if (!s || !host) return FS2D_ERR_ARG;
    const size_t n = fs2d_field_count(s, field);
    if (n == 0 && !is_particle_field(field)) return fail(s, FS2D_ERR_ARG, "unknown field %d", field);
    if (count != n) return fail(s, FS2D_ERR_ARG, "field %d has %zu elements, buffer has %zu", field, n, count);
    if (!valid_dtype(field, dtype)) return fail(s, FS2D_ERR_ARG, "dtype %d invalid for field %d", dtype, field);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (n == 0) return FS2D_OK;
    if (field == FS2D_FIELD_PARTICLE_CELL || field == FS2D_FIELD_PARTICLE_ORDER) {
        sort_particles(s);
        TRY(d2h(s, host, field == FS2D_FIELD_PARTICLE_CELL ? s->pkey : s->sorted, n * sizeof(int)));
        return FS2D_OK;
    }
    if (field == FS2D_FIELD_CELL) {
        TRY(d2h(s, host, s->cell, n));
        return FS2D_OK;
    }
    if (field == FS2D_FIELD_PHI) {
        std::vector<double> tmp(n);
        TRY(d2h(s, tmp.data(), s->phi, n * sizeof(double)));
        if (dtype == FS2D_F64) memcpy(host, tmp.data(), n * 8);
        else for (size_t i = 0; i < n; i++) ((float*)host)[i] = (float)tmp[i];
        return FS2D_OK;
    }
    std::vector<float> tmp(n);
    if (field == FS2D_FIELD_PARTICLE_POS || field == FS2D_FIELD_PARTICLE_VEL) {
        const float* a = field == FS2D_FIELD_PARTICLE_POS ? s->px : s->pvx;
        const float* b = field == FS2D_FIELD_PARTICLE_POS ? s->py : s->pvy;
        std::vector<float> xs(s->np), ys(s->np);
        TRY(d2h(s, xs.data(), a, s->np * sizeof(float)));
        TRY(d2h(s, ys.data(), b, s->np * sizeof(float)));
        for (long e = 0; e < s->np; e++) {
            tmp[2 * e] = xs[e];
            tmp[2 * e + 1] = ys[e];
        }
    } else {
        TRY(d2h(s, tmp.data(), fp32_field(s, field), n * sizeof(float)));
    }
    if (dtype == FS2D_F32) memcpy(host, tmp.data(), n * 4);
    else for (size_t i = 0; i < n; i++) ((double*)host)[i] = tmp[i];
    return FS2D_OK;
}

static bool particle_in_box(const fs2d_sim* s, double x, double y) {
    // Positions inside [dx, (n-1)dx) keep every transfer stencil inside the grid
    // (the reference's particles never leave it: seeding in interior cells and
    // clampPos after every move, FluidSim2D.cpp:52-64, 645-651).
    return std::isfinite(x) && std::isfinite(y) && (float)x >= (float)s->dx && (float)y >= (float)s->dx &&
           (double)(float)x < (s->nx - 1) * s->dx && (double)(float)y < (s->ny - 1) * s->dx;
}

fs2d_status fs2d_set_field(fs2d_sim* s, int field, int dtype, const void* host, size_t count) {
    if (!s || !host) return FS2D_ERR_ARG;
    const size_t n = fs2d_field_count(s, field);
    if (field == FS2D_FIELD_PARTICLE_CELL || field == FS2D_FIELD_PARTICLE_ORDER)
        return fail(s, FS2D_ERR_ARG, "field %d is read-only", field);
    if (n == 0 && !is_particle_field(field)) return fail(s, FS2D_ERR_ARG, "unknown field %d", field);
    if (count != n) return fail(s, FS2D_ERR_ARG, "field %d has %zu elements, buffer has %zu", field, n, count);
    if (!valid_dtype(field, dtype)) return fail(s, FS2D_ERR_ARG, "dtype %d invalid for field %d", dtype, field);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (n == 0) return FS2D_OK;
    if (field == FS2D_FIELD_CELL) {
        const uint8_t* c = (const uint8_t*)host;
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                uint8_t f = c[(size_t)j * s->nx + i];
                if (f > FS2D_SOLID || ((i == 0 || j == 0 || i == s->nx - 1 || j == s->ny - 1) && f != FS2D_SOLID))
                    return fail(s, FS2D_ERR_ARG, "invalid cell flag %u at (%d,%d)", f, i, j);
            }
        TRY(h2d(s, s->cell, host, n));
        return FS2D_OK;
    }
    if (field == FS2D_FIELD_PHI) {
        std::vector<double> tmp(n);
        for (size_t i = 0; i < n; i++)
            tmp[i] = dtype == FS2D_F64 ? ((const double*)host)[i] : (double)((const float*)host)[i];
        TRY(h2d(s, s->phi, tmp.data(), n * sizeof(double)));
        return FS2D_OK;
    }
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; i++)
        tmp[i] = dtype == FS2D_F32 ? ((const float*)host)[i] : (float)((const double*)host)[i];
    if (field == FS2D_FIELD_PARTICLE_POS || field == FS2D_FIELD_PARTICLE_VEL) {
        std::vector<float> xs(s->np), ys(s->np);
        for (long e = 0; e < s->np; e++) {
            xs[e] = tmp[2 * e];
            ys[e] = tmp[2 * e + 1];
            if (field == FS2D_FIELD_PARTICLE_POS && !particle_in_box(s, xs[e], ys[e]))
                return fail(s, FS2D_ERR_ARG, "particle %ld at (%g, %g) outside [dx, (n-1)dx)", e, xs[e], ys[e]);
        }
        float* a = field == FS2D_FIELD_PARTICLE_POS ? s->px : s->pvx;
        float* b = field == FS2D_FIELD_PARTICLE_POS ? s->py : s->pvy;
        TRY(h2d(s, a, xs.data(), s->np * sizeof(float)));
        TRY(h2d(s, b, ys.data(), s->np * sizeof(float)));
        if (field == FS2D_FIELD_PARTICLE_POS) s->sorted_valid = false;
        return FS2D_OK;
    }
    TRY(h2d(s, fp32_field(s, field), tmp.data(), n * sizeof(float)));
    return FS2D_OK;
}

fs2d_status fs2d_set_particles(fs2d_sim* s, int dtype, const void* pos_xy, size_t n) {
    if (!s || (n > 0 && !pos_xy)) return FS2D_ERR_ARG;
    if (dtype != FS2D_F32 && dtype != FS2D_F64) return fail(s, FS2D_ERR_ARG, "dtype %d", dtype);
    if (n > (size_t)INT32_MAX / 2) return fail(s, FS2D_ERR_ARG, "too many particles");
    std::vector<float> xs(n), ys(n);
    for (size_t e = 0; e < n; e++) {
        double x = dtype == FS2D_F32 ? ((const float*)pos_xy)[2 * e] : ((const double*)pos_xy)[2 * e];
        double y = dtype == FS2D_F32 ? ((const float*)pos_xy)[2 * e + 1] : ((const double*)pos_xy)[2 * e + 1];
        if (!particle_in_box(s, x, y))
            return fail(s, FS2D_ERR_ARG, "particle %zu at (%g, %g) outside [dx, (n-1)dx)", e, x, y);
        xs[e] = (float)x;
        ys[e] = (float)y;
    }
    HIPCHK(s, hipStreamSynchronize(s->stream));
    TRY(ensure_particle_capacity(s, std::max<long>((long)n, 1)));
    s->np = (long)n;
    if (n > 0) {
        TRY(h2d(s, s->px, xs.data(), n * sizeof(float)));
        TRY(h2d(s, s->py, ys.data(), n * sizeof(float)));
        fill(s, s->pvx, 0.0f, n);
        fill(s, s->pvy, 0.0f, n);
        HIPCHK(s, hipStreamSynchronize(s->stream));
    }
    s->sorted_valid = false;
    return FS2D_OK;
}

fs2d_status fs2d_device_field(fs2d_sim* s, int field, void** dptr) {
    if (!s || !dptr) return FS2D_ERR_ARG;
    switch (field) {
        case FS2D_FIELD_CELL: *dptr = s->cell; return FS2D_OK;
        case FS2D_FIELD_PHI: *dptr = s->phi; return FS2D_OK;
        case FS2D_FIELD_PARTICLE_POS: *dptr = s->px; return FS2D_OK;
        case FS2D_FIELD_PARTICLE_VEL: *dptr = s->pvx; return FS2D_OK;
        default: break;
    }
    float* f = fp32_field(s, field);
    if (!f) return fail(s, FS2D_ERR_ARG, "unknown field %d", field);
    *dptr = f;
    return FS2D_OK;
}

void* fs2d_stream(fs2d_sim* s) { return s ? (void*)s->stream : nullptr; }

fs2d_status fs2d_set_gravity(fs2d_sim* s, double gx, double gy) {
    if (!s) return FS2D_ERR_ARG;
    s->gx = gx;
    s->gy = gy;
    return FS2D_OK;
}
fs2d_status fs2d_set_pic_flip_alpha(fs2d_sim* s, double a) {
    if (!s) return FS2D_ERR_ARG;
    s->alpha = a;
    return FS2D_OK;
}
fs2d_status fs2d_set_dt(fs2d_sim* s, double dt) {
    if (!s || !(dt >= 0) || !std::isfinite(dt)) return FS2D_ERR_ARG;
    s->dt = dt;
    return FS2D_OK;
}
fs2d_status fs2d_set_solver(fs2d_sim* s, double tau, double sigma, double tol, int max_iters) {
    if (!s || !(tau >= 0) || !(sigma >= 0) || !(tol >= 0) || max_iters < 0) return FS2D_ERR_ARG;
    s->tau = tau;
    s->sigma = sigma;
    s->tol = tol;
    s->max_iters = max_iters;
    return FS2D_OK;
}

fs2d_status fs2d_get_stats(fs2d_sim* s, fs2d_stats* out) {
    if (!s || !out) return FS2D_ERR_ARG;
    HIPCHK(s, hipMemcpyAsync(s->h_dg, s->dg, sizeof(Diag), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->h_sc, s->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    memset(out, 0, sizeof *out);
    out->water_volume = s->h_dg->volume;
    out->total_energy = s->h_dg->grid_energy;
    out->particle_total_energy = s->h_dg->particle_energy;
    out->orig_water_volume = s->orig_volume;
    out->current_time = s->current_time;
    out->pcg_residual = s->h_sc->rhs_norm > 0 ? s->h_sc->r_norm / s->h_sc->rhs_norm : 0.0;
    out->cfl_max = s->h_dg->cfl_max;
    out->pcg_iters = s->h_sc->iter;
    out->pcg_hit_max = s->h_sc->iter >= s->max_iters;
    out->n_particles = s->np;
    out->steps = s->steps;
    return FS2D_OK;
}

fs2d_status fs2d_avg_pressure(fs2d_sim* s, int fluid_only, double* out) {
    if (!s || !out) return FS2D_ERR_ARG;
    const size_t n = (size_t)n_c(s);
    std::vector<float> p(n);
    std::vector<uint8_t> c(n);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    TRY(d2h(s, p.data(), s->p, n * 4));
    TRY(d2h(s, c.data(), s->cell, n));
    double acc = 0;
    size_t cnt = 0;
    for (size_t i = 0; i < n; i++)
        if (!fluid_only || c[i] == FS2D_FLUID) {
            acc += p[i];
            cnt++;
        }
    // avgPressure divides by all cells (:612), avgPressureInFluid by the fluid count (:625)
    *out = acc / (double)(fluid_only ? cnt : n);
    return FS2D_OK;
}

fs2d_status fs2d_max_velocity(fs2d_sim* s, double* out) {
    if (!s || !out) return FS2D_ERR_ARG;
    // FluidSim2D::maxVelocity (:629-638): |velInterp| at the corners of FLUID cells.
    const size_t nU = (size_t)n_u(s), nV = (size_t)n_v(s), nC = (size_t)n_c(s);
    std::vector<float> u(nU), v(nV);
    std::vector<uint8_t> c(nC);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    TRY(d2h(s, u.data(), s->u, nU * 4));
    TRY(d2h(s, v.data(), s->v, nV * 4));
    TRY(d2h(s, c.data(), s->cell, nC));
    auto cr = [](const std::vector<float>& a, int w, int h, double gx, double gy) -> double {
        int x = (int)gx, y = (int)gy;
        if (x < 0 || x >= w || y < 0 || y >= h) return 0;
        double fx = gx - x, fy = gy - y;
        double uu[4] = {-0.5 * fx * fx * fx + fx * fx - 0.5 * fx, 1.5 * fx * fx * fx - 2.5 * fx * fx + 1,
                        -1.5 * fx * fx * fx + 2 * fx * fx + 0.5 * fx, 0.5 * fx * fx * fx - 0.5 * fx * fx};
        double vv[4] = {-0.5 * fy * fy * fy + fy * fy - 0.5 * fy, 1.5 * fy * fy * fy - 2.5 * fy * fy + 1,
                        -1.5 * fy * fy * fy + 2 * fy * fy + 0.5 * fy, 0.5 * fy * fy * fy - 0.5 * fy * fy};
        double acc = 0;
        for (int r = 0; r < 4; r++) {
            int yy = std::min(std::max(y - 1 + r, 0), h - 1);
            double row = 0;
            for (int q = 0; q < 4; q++) row += uu[q] * a[(size_t)yy * w + std::min(std::max(x - 1 + q, 0), w - 1)];
            acc += row * vv[r];
        }
        return acc;
    };
    double m = 0;
    for (int j = 0; j < s->ny; j++)
        for (int i = 0; i < s->nx; i++) {
            if (c[(size_t)j * s->nx + i] != FS2D_FLUID) continue;
            double X = i, Y = j;
            double ux = std::min(std::max(X, 1e-6), s->nx - 1 - 1e-6);
            double uy = std::min(std::max(Y - 0.5, 1e-6), s->ny - 1 - 1e-6);
            double vx = std::min(std::max(X - 0.5, 1e-6), s->nx - 1 - 1e-6);
            double vy = std::min(std::max(Y, 1e-6), s->ny - 1 - 1e-6);
            double a = cr(u, s->nx + 1, s->ny, ux, uy), b = cr(v, s->nx, s->ny + 1, vx, vy);
            m = std::max(m, std::sqrt(a * a + b * b));
        }
    *out = m;
    return FS2D_OK;
}

fs2d_status fs2d_set_graph(fs2d_sim* s, int enable) {
    if (!s) return FS2D_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->use_graph = enable != 0;
    if (!s->use_graph) {
        for (auto& g : s->graphs) graph_free(g);
        s->graphs.clear();
    }
    return FS2D_OK;
}

fs2d_status fs2d_set_timing(fs2d_sim* s, int enable) {
    return fs2d_set_timing_classes(s, enable ? ~0ull : 0ull);
}
fs2d_status fs2d_set_timing_classes(fs2d_sim* s, unsigned long long class_mask) {
    if (!s) return FS2D_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    resolve_timing(s);
    s->timing = class_mask & ((1ull << K_COUNT) - 1ull);
    return FS2D_OK;
}
int fs2d_kernel_count(void) { return K_COUNT; }
const char* fs2d_kernel_name(int k) { return (k >= 0 && k < K_COUNT) ? kKernelNames[k] : nullptr; }
fs2d_status fs2d_get_kernel_time(fs2d_sim* s, int k, double* total_ms, long long* launches) {
    if (!s || k < 0 || k >= K_COUNT || !total_ms || !launches) return FS2D_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    resolve_timing(s);
    *total_ms = s->ktime[k];
    *launches = s->kcount[k];
    return FS2D_OK;
}
fs2d_status fs2d_reset_kernel_times(fs2d_sim* s) {
    if (!s) return FS2D_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    resolve_timing(s);
    for (int k = 0; k < K_COUNT; k++) {
        s->ktime[k] = 0;
        s->kcount[k] = 0;
    }
    return FS2D_OK;
}

const char* fs2d_last_error(const fs2d_sim* s) { return s ? s->err.c_str() : g_create_error.c_str(); }

const char* fs2d_status_string(int st) {
    switch (st) {
        case FS2D_OK: return "FS2D_OK";
        case FS2D_WARN_MAXITER: return "FS2D_WARN_MAXITER";
        case FS2D_WARN_CFL: return "FS2D_WARN_CFL";
        case FS2D_ERR_NAN: return "FS2D_ERR_NAN";
        case FS2D_ERR_OOM: return "FS2D_ERR_OOM";
        case FS2D_ERR_CONFIG: return "FS2D_ERR_CONFIG";
        case FS2D_ERR_ARG: return "FS2D_ERR_ARG";
        case FS2D_ERR_DEVICE: return "FS2D_ERR_DEVICE";
        case FS2D_ERR_HANG: return "FS2D_ERR_HANG";
        default: return "unknown";
    }
}

}  // extern "C"

// ===========================================================================
// applyProjection without allocation (fs2d_project_noalloc).
//
// Every array the projection needs beyond its inputs and output lives in
// `scratch`, a file-scope device array: one copy per GPU, held from the
// library's load until the process ends.  Host code never names `scratch`
// itself (in host code the name is a host-side stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address reaches kernels.
// Calls share their GPU's copy, so calls that overlap on one GPU race; each
// call clears what must start at zero (the strip ticket, the hand-off words,
// the error flag) with fill kernels before its first sweep.
// ===========================================================================
namespace {

constexpr long kNoallocEntries = 1L << 19;  // SK4 entries (512 x 512 cells need 294,912)
constexpr int kNoallocMaxStrips = 128;     // S * T4 <= kNoallocEntries / 256 and T4 >= 24 give S <= 85
// hand-off words of S strips: S * (3 * (nx + 192) + 64) with nx <= 4 T4 - 64
constexpr long kNoallocHandWords = 12L * (kNoallocEntries / 256) + 448L * kNoallocMaxStrips;
constexpr int kNoallocFloatArrays = 13;  // r z q w s0 s1 pk Adiag P fx fy gx gy

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaFloats = na_align(sizeof(float) * kNoallocEntries);
constexpr size_t kNaCpl = na_align(kNoallocEntries);
constexpr size_t kNaSc = na_align(sizeof(PcgScalars));
constexpr size_t kNaQq = na_align(sizeof(double) * 2 * kNoallocMaxStrips * kLoaders);
constexpr size_t kNaPart = na_align(sizeof(double) * kMaxBlocks);
constexpr size_t kNaWord = 256;
constexpr size_t kNaHand = na_align(sizeof(unsigned long long) * kNoallocHandWords);
constexpr size_t kNoallocBytes =
    kNoallocFloatArrays * kNaFloats + kNaCpl + kNaSc + kNaQq + 2 * kNaPart + 2 * kNaWord + kNaHand;

__device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kNoallocMaxDevices] = {};

// Failures are explained by fs2d_last_error(NULL), like those of fs2d_create.
fs2d_status noalloc_fail(fs2d_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_create_error = buf;
    return st;
}

// The device address of `scratch` on the current device (looked up once).
fs2d_status noalloc_scratch(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kNoallocMaxDevices) return noalloc_fail(FS2D_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return noalloc_fail(FS2D_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_noalloc_ptr[dev] = (unsigned char*)p;
    }
    *out = g_noalloc_ptr[dev];
    return FS2D_OK;
}

}  // namespace

extern "C" {

fs2d_status fs2d_project_noalloc(const fs2d_projection_args* a, int* iters, double* residual) {
    if (!a || !a->cell || !a->phi || !a->u || !a->v || !a->p)
        return noalloc_fail(FS2D_ERR_ARG, "fs2d_project_noalloc: NULL argument");
    if (a->size_x < 3 || a->size_y < 3 || a->size_x > 16384 || a->size_y > 16384)
        return noalloc_fail(FS2D_ERR_ARG, "fs2d_project_noalloc: size %dx%d outside [3, 16384]", a->size_x,
                            a->size_y);
    const SkGeom geom = make_geom(a->size_x, a->size_y);
    const long long entries = geom.size();
    const long hand = (long)geom.S * (long)sk_handoff_stride(a->size_x, 3);
    if (entries > kNoallocEntries || geom.S > kNoallocMaxStrips || hand > kNoallocHandWords)
        return noalloc_fail(FS2D_ERR_ARG,
                            "fs2d_project_noalloc: a %d x %d grid needs %lld solver entries; the static scratch "
                            "holds %ld (512 x 512 cells fit)",
                            a->size_x, a->size_y, entries, kNoallocEntries);
    if (!(a->dx > 0) || !(a->rho > 0) || !(a->dt >= 0) || !(a->tau >= 0) || !(a->sigma >= 0) ||
        !(a->pcg_tol >= 0) || a->pcg_max_iters < 0 || (a->flags & ~FS2D_FLAG_NO_GHOST_PRESSURE))
        return noalloc_fail(FS2D_ERR_ARG, "fs2d_project_noalloc: invalid solver parameters or flags");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(FS2D_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    unsigned char* base = nullptr;
    fs2d_status st = noalloc_scratch(dev, &base);
    if (st != FS2D_OK) return st;

    // A handle on the host stack over the caller's arrays and the scratch.
    fs2d_sim t;
    t.nx = a->size_x;
    t.ny = a->size_y;
    t.dt = a->dt;
    t.dx = a->dx;
    t.rho = a->rho;
    t.tau = a->tau;
    t.sigma = a->sigma;
    t.tol = a->pcg_tol;
    t.max_iters = a->pcg_max_iters;
    t.flags = a->flags;
    t.stream = (hipStream_t)a->stream;
    t.device = dev;
    t.geom = geom;
    t.cell = const_cast<uint8_t*>(a->cell);
    t.phi = const_cast<double*>(a->phi);
    t.u = const_cast<float*>(a->u);
    t.v = const_cast<float*>(a->v);
    unsigned char* q = base;
    float** fl[kNoallocFloatArrays] = {&t.r, &t.z, &t.q, &t.w, &t.s[0], &t.s[1], &t.pk,
                                       &t.Adiag, &t.P, &t.fx, &t.fy, &t.gx_, &t.gy_};
    for (float** f : fl) {
        *f = (float*)q;
        q += kNaFloats;
    }
    t.cpl = q;                                   q += kNaCpl;
    t.sc = (PcgScalars*)q;                       q += kNaSc;
    t.qq = (double*)q;                           q += kNaQq;
    t.ws_part = (double*)q;                      q += kNaPart;
    t.rmax_part = (double*)q;                    q += kNaPart;
    t.sk_ticket = (unsigned long long*)q;        q += kNaWord;
    t.werr = (int*)q;                            q += kNaWord;
    t.sk_handoff = (unsigned long long*)q;
    t.fused_pcg = false;  // the separate SpMV launch: no completion flags or tagged sigma in the scratch
    PcgScalars h_sc{};
    t.h_sc = &h_sc;  // pageable: read back with a synchronising copy

    // A fresh start, as fs2d_create's zeroed buffers give a handle.
    fill(&t, t.sk_ticket, 0ull, 1);
    fill(&t, t.sk_handoff, 0ull, (size_t)hand);
    fill(&t, t.werr, 0, 1);
    assemble(&t);
    st = pcg_solve(&t, a->p);
    int h_err = 0;
    e = hipMemcpyAsync(&h_err, t.werr, sizeof(int), hipMemcpyDeviceToHost, t.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return noalloc_fail(FS2D_ERR_DEVICE, "fs2d_project_noalloc: %s", hipGetErrorString(e));
    if (st < 0) return noalloc_fail(st, "fs2d_project_noalloc: %s", t.err.c_str());
    if (h_err) return noalloc_fail(FS2D_ERR_HANG, "fs2d_project_noalloc: a wavefront hand-off wait expired");
    if (iters) *iters = h_sc.iter;
    if (residual) *residual = h_sc.rhs_norm > 0 ? h_sc.r_norm / h_sc.rhs_norm : 0.0;
    if (st == FS2D_WARN_MAXITER)
        noalloc_fail(st, "PCG: maximum iteration limit (%d) exceeded", t.max_iters);
    return st;
}

}  // extern "C"

#include "fs2d_adjoint_api.hpp"
#include "fs2d_batch_api.hpp"
#include "fs2d_tune_api.hpp"
#include "fs2d_render.hpp"
#include "fs3d_api.hpp"
