// fs2d_batch.hpp -- batched launches: one launch runs a kernel of the step
// over every simulation of an ensemble, each with its own grid size.
//
// Each kernel's body takes the workgroup's index and count within its
// simulation (vb, nvb).  A batched launch concatenates the simulations'
// workgroup ranges, each as large as the simulation's own single launch and
// starting at a multiple of 8 (so vb % 8 is the XCD, as xcd_loop assumes); a
// map gives every workgroup its simulation and vb.  The arithmetic and the
// reduction orders are therefore those of the single-simulation launches, and
// a member of a batch steps bit-identically to the same simulation alone.
//
// The wavefront passes (level-set sweeps, MIC(0) factor, triangular solves)
// take strips from one ticket over the concatenated strip lists of all
// simulations: a strip's upstream neighbour belongs to the same simulation and
// has a smaller position in the list, so it was taken earlier and the launch
// cannot deadlock.  Per-simulation solver state (PcgScalars) masks converged
// simulations: their workgroups return at once.
#pragma once

#include <climits>

#include "fs2d_kernels.hpp"

namespace fs2d {

// Device pointers and scalars of one simulation, copied from its handle.
struct InstBufs {
    int nx, ny, mode, flags;
    long np;
    double dt, dx, dr, rho, gx, gy, alpha, tau, sigma;
    double pmass;  // rho dx^2 / k^2 (0 without particles)
    // the solver's rhs factor 1 / (dx scaleA) and |rhs|_inf zero threshold in
    // solver units, computed on the host as the single handle computes them
    double rhs_scale, rhs_zero;
    int max_iters;
    SkGeom geom;
    float *u, *v, *u2, *v2, *nu, *nv, *dens, *dens2;
    uint8_t* cell;
    uint8_t* cell2;
    double *phi, *phi2;
    int *t0, *t1;
    int* ls_chg;  // closest-particle pass change flags (16)
    float *p, *r, *z, *q, *w, *s0, *s1, *pk, *Adiag, *P, *fx, *fy, *gx_, *gy_;
    uint8_t* cpl;
    PcgScalars* sc;
    double *qq, *ws_part, *rmax_part;
    uint32_t *mu, *mv;
    int *fu0, *fu1, *fv0, *fv1, *fcount, *fpos;
    int4* fsrc;
    float4* fkv;
    float *px, *py, *pvx, *pvy;
    int *pkey, *counts, *starts, *cursor, *sorted, *tile_sums;
    Diag* dg;
    double* partials;
    unsigned long long* sk_handoff;  // hand-off rows of the batched solver passes
    unsigned* fa_flags;              // the batch's p/r update run flags (b_fwd_axpy_kernel)
    unsigned long long* fa_handoff;  // the batch's hand-off rows of that launch's forward solve
};

// Workgroup -> (simulation, workgroup within it) of one batched launch.
struct BlockMap {
    const int* inst;   // per workgroup; -1 for padding
    const int* first;  // per simulation
    const int* count;  // per simulation
};
__device__ __forceinline__ bool bmap(const BlockMap& m, int& b, int& vb, int& nvb) {
    b = m.inst[blockIdx.x];
    if (b < 0) return false;
    vb = (int)blockIdx.x - m.first[b];
    nvb = m.count[b];
    return true;
}

// Global strip -> (simulation, strip order within it).
struct StripMap {
    const int* inst;
    const int* first;
    int total;
};

// ---------------------------------------------------------------------------
// Element-wise and stencil kernels.
// ---------------------------------------------------------------------------
enum BatchFill { BF_COUNTS, BF_FCOUNT };
__global__ __launch_bounds__(kBlock) void b_fill_kernel(const InstBufs* __restrict__ ins, BlockMap bm, int what) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    if (what == BF_COUNTS) fill_body(I.counts, 0, (size_t)I.nx * I.ny, vb, nvb);
    else fill_body(I.fcount, 0, 2, vb, nvb);
}

// Copies over the face domain (u then v) or the cells.
enum BatchCopy { BC_UV_FROM_NEW, BC_UV_FROM_U2, BC_DENS_FROM_DENS2, BC_CURSOR_FROM_STARTS };
__global__ __launch_bounds__(kBlock) void b_copy_kernel(const InstBufs* __restrict__ ins, BlockMap bm, int what) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    const long nU = (long)(I.nx + 1) * I.ny, nV = (long)I.nx * (I.ny + 1), nC = (long)I.nx * I.ny;
    if (what == BC_UV_FROM_NEW || what == BC_UV_FROM_U2) {
        const float* su = what == BC_UV_FROM_NEW ? I.nu : I.u2;
        const float* sv = what == BC_UV_FROM_NEW ? I.nv : I.v2;
        for (long e = vb * (long)blockDim.x + threadIdx.x; e < nU + nV; e += (long)nvb * blockDim.x) {
            if (e < nU) I.u[e] = su[e];
            else I.v[e - nU] = sv[e - nU];
        }
    } else if (what == BC_DENS_FROM_DENS2) {
        copy_body(I.dens, (const float*)I.dens2, (size_t)nC, vb, nvb);
    } else {
        copy_body(I.cursor, (const int*)I.starts, (size_t)nC, vb, nvb);
    }
}

__global__ __launch_bounds__(kBlock) void b_sl_advect_kernel(const InstBufs* __restrict__ ins, BlockMap bm,
                                                             int with_gravity) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    const bool dens = (I.flags & 2) != 0;  // FS2D_FLAG_DENSITY
    const float gu = with_gravity ? (float)(I.dt * I.gx) : 0.0f;
    const float gv = with_gravity ? (float)(I.dt * I.gy) : 0.0f;
    sl_advect_body(make_mac(I.u, I.v, I.nx, I.ny, I.dx), I.u2, I.v2, dens ? I.dens : nullptr, I.dens2, I.dt,
                   clamp_box(I.nx, I.ny, I.dx), gu, gv, vb, nvb);
}

__global__ __launch_bounds__(kBlock) void b_gravity_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    gravity_body(I.u, I.v, (long)(I.nx + 1) * I.ny, (long)I.nx * (I.ny + 1), (float)(I.dt * I.gx),
                 (float)(I.dt * I.gy), vb, nvb);
}

// PcgScalars of every simulation: zero, tolerance set (one workgroup each);
// no member is known to be done.
__global__ void b_pcg_init_kernel(const InstBufs* __restrict__ ins, int n, double tol, int* all_done_since) {
    const int b = blockIdx.x;
    if (b == 0 && threadIdx.x == 0) *all_done_since = INT_MAX;
    if (b >= n || threadIdx.x != 0) return;
    PcgScalars z{};
    z.tol = tol;
    *ins[b].sc = z;
}

__global__ __launch_bounds__(kBlock) void b_assemble_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    // solver units, A / scaleA and rhs / scaleA (as the single path's assemble)
    assemble_sk_body(I.geom, I.cell, I.phi, I.u, I.v, I.Adiag, I.cpl, I.r, I.pk, I.s0, I.rhs_scale,
                     I.rmax_part, (I.flags & 8) ? 0 : 1 /* FS2D_FLAG_NO_GHOST_PRESSURE */, vb, nvb);
}

// The members still iterating and their virtual workgroups, packed: member b
// owns virtual workgroups first[b] .. first[b] + count[b] - 1 of the
// iteration's SpMV and axpy loops, each range starting at a multiple of 8 (so
// vb % 8 stays the XCD), and map[k] names the member of virtual workgroup k
// (-1 for padding).  Built once per iteration, from the members' done flags,
// which only the forward solve's stopping test and the factor write, so the
// SpMV and the axpy of an iteration share it.  Members are packed in `order`,
// heaviest virtual workgroups first: workgroup g of a loop runs k = g, g + G,
// g + 2 G, ..., and each round of G consecutive entries then spans members of
// similar weight.
struct PcgPlanBufs {
    int* map;     // [sum of aligned counts]
    int* first;   // [n]
    int* total;   // [1]
    int* active;  // [n] 1 while the member iterates (its done flag as the plan saw it)
};
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
    __shared__ int wsum[kBlock / 64];
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if ((threadIdx.x & 63) >= o) x += y;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; k++) {
        if (k < (int)(threadIdx.x >> 6)) before += wsum[k];
        total += wsum[k];
    }
    return before + x - v;
}
// n workgroups: every one forms the prefix sums of the members' aligned counts
// (dynamic LDS: n ints), workgroup i writes the map range of member order[i],
// and workgroup 0 the total.  When no member iterates it records h in
// all_done_since: the solves of iteration h and later then skip their strips
// at once.
__global__ __launch_bounds__(kBlock) void b_pcg_plan_kernel(const InstBufs* __restrict__ ins, int n,
                                                            const int* __restrict__ count,
                                                            const int* __restrict__ order, PcgPlanBufs P, int h,
                                                            int* all_done_since) {
    extern __shared__ int vals[];  // aligned count of order[i], 0 if done
    __shared__ int my_first;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const int b = order[i];
        vals[i] = ins[b].sc->done ? 0 : (count[b] + 7) & ~7;
    }
    __syncthreads();
    const int c = (n + kBlock - 1) / kBlock;
    const int i0 = min((int)threadIdx.x * c, n), i1 = min(i0 + c, n);
    int sum = 0;
    for (int i = i0; i < i1; i++) sum += vals[i];
    int total;
    int ps = block_excl_scan(sum, total);
    for (int i = i0; i < i1; i++) {
        if (i == (int)blockIdx.x) my_first = ps;
        ps += vals[i];
    }
    __syncthreads();
    const int i = blockIdx.x;
    const int b = order[i];
    const int v = vals[i];
    if (threadIdx.x == 0) P.active[b] = v != 0;
    if (v) {
        const int f = my_first, cnt = count[b];
        for (int k = threadIdx.x; k < v; k += kBlock) P.map[f + k] = k < cnt ? b : -1;
        if (threadIdx.x == 0) P.first[b] = f;
        // sigma_h and beta of the member, once, in the order and arithmetic
        // of spmv_scalars (one wave's fixed-order sums), for b_spmv_kernel
        if (threadIdx.x < 64) {
            const InstBufs& I = ins[b];
            const int ns = I.geom.S * kLoaders;
            const double* qq0 = I.qq;
            const double* qq1 = I.qq + (size_t)ns;
            const double sig = wave_sum_of((h & 1) ? qq1 : qq0, ns);
            double beta_d = 0;
            if (h != 0) beta_d = sig / wave_sum_of((h & 1) ? qq0 : qq1, ns);
            if (threadIdx.x == 0) {
                I.sc->sigma = sig;
                I.sc->beta = (float)beta_d;
            }
        }
    }
    if (i == 0 && threadIdx.x == 0) {
        *P.total = total;
        if (total == 0 && *all_done_since > h) *all_done_since = h;
    }
}

// Iteration h of the PCG loop over the members still iterating (one
// persistent launch of any size: workgroup g runs virtual workgroups g, g +
// gridDim.x, ... of the iteration's plan; s double-buffered by parity; beta
// formed once per member by the plan kernel from sigma_h and sigma_{h-1}, the
// two halves of qq).  Each member's virtual workgroups compute what its single
// launch computes, so members step bit-identically to single simulations.
__global__ __launch_bounds__(kBlock) void b_spmv_kernel(const InstBufs* __restrict__ ins,
                                                        const int* __restrict__ count, PcgPlanBufs P, int h) {
    const int total = *P.total;
    int b_next = (int)blockIdx.x < total ? P.map[blockIdx.x] : -1;
    for (int k = blockIdx.x; k < total; k += gridDim.x) {
        const int b = b_next;
        if (k + (int)gridDim.x < total) b_next = P.map[k + gridDim.x];
        if (b < 0) continue;  // padding to a multiple of 8
        const int vb = k - P.first[b], nvb = count[b];
        const InstBufs& I = ins[b];
        const float* s_old = (h & 1) ? I.s1 : I.s0;
        float* s_new = (h & 1) ? I.s0 : I.s1;
        const SpmvPre pre = spmv_pre(I.geom, I.Adiag, I.cpl, I.z, s_old, vb, nvb);
        const float beta = I.sc->beta;  // b_pcg_plan_kernel's, as spmv_scalars forms it
        spmv_core(pre, I.geom, I.Adiag, I.cpl, I.z, s_old, s_new, I.w, beta, I.ws_part, vb);
    }
}
// w.s of every member still iterating, summed once per member (one wave per
// member) in the order the single launch's axpy sums it, so that the axpy's
// workgroups need not each re-read all of the member's partials.  `nparts`:
// the member's SpMV virtual workgroup count.
__global__ __launch_bounds__(64) void b_ws_reduce_kernel(const InstBufs* __restrict__ ins, int n,
                                                         const int* nparts) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const InstBufs& I = ins[b];
    if (I.sc->done) return;
    const double t = wave_sum_of(I.ws_part, nparts[b]);
    if (threadIdx.x == 0) I.sc->ws = t;
}
__global__ __launch_bounds__(kBlock) void b_axpy_kernel(const InstBufs* __restrict__ ins,
                                                        const int* __restrict__ count, PcgPlanBufs P, int h) {
    const int total = *P.total;
    int b_next = (int)blockIdx.x < total ? P.map[blockIdx.x] : -1;
    for (int k = blockIdx.x; k < total; k += gridDim.x) {
        const int b = b_next;
        if (k + (int)gridDim.x < total) b_next = P.map[k + gridDim.x];
        if (b < 0) continue;
        const int vb = k - P.first[b], nvb = count[b];
        const InstBufs& I = ins[b];
        float4* p4 = (float4*)I.pk;
        float4* r4 = (float4*)I.r;
        const float4* s4 = (const float4*)((h & 1) ? I.s0 : I.s1);
        const float4* w4 = (const float4*)I.w;
        const XcdLoop it = xcd_loop(I.geom.size() / 4, vb, nvb);
        AxpyBuf Bf;
        Bf.load(it, it.e, p4, r4, s4, w4);
        // sigma (written by the SpMV) and w.s come reduced in the member's scalars
        const float a = (float)(I.sc->sigma / I.sc->ws);
        axpy_core(Bf, it, p4, r4, s4, w4, a, I.rmax_part, vb);
    }
}
__global__ __launch_bounds__(kBlock) void b_sk_to_rowmajor_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    sk_to_rowmajor_body(I.geom, I.pk, I.p, vb, nvb);
}

__global__ __launch_bounds__(kBlock) void b_update_velocity_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    update_velocity_body(I.cell, I.phi, I.p, I.u, I.v, I.nu, I.nv, I.mu, I.mv, I.nx, I.ny, I.dt / (I.rho * I.dx),
                         (I.flags & 8) ? 0 : 1 /* FS2D_FLAG_NO_GHOST_PRESSURE */, vb, nvb);
}

// Extrapolation of (u, v) (which = 0, after P2G) or (nu, nv) (which = 1).
__device__ __forceinline__ void b_extra_arrays(const InstBufs& I, int which, ExtraArr& A, ExtraArr& B) {
    const long nU = (long)(I.nx + 1) * I.ny;
    A = ExtraArr{G(which ? I.nu : I.u), G(I.mu),  G(I.fu0), G(I.fpos), G((i4v*)I.fsrc), G((f4v*)I.fkv), G(I.fu1),
                 G(I.fcount + 0),       I.nx + 1, I.ny};
    B = ExtraArr{G(which ? I.nv : I.v), G(I.mv),        G(I.fv0),         G(I.fpos + nU), G((i4v*)(I.fsrc + nU)),
                 G((f4v*)(I.fkv + nU)),         G(I.fv1),       G(I.fcount + 1),  I.nx,           I.ny + 1};
}
enum BatchExtra { BX_ROWS, BX_HIST, BX_SCATTER, BX_LINK };
template <int WHAT>
__global__ __launch_bounds__(kBlock) void b_extrap_kernel(const InstBufs* __restrict__ ins, BlockMap bm, int which) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    ExtraArr A, B;
    b_extra_arrays(ins[b], which, A, B);
    if (WHAT == BX_ROWS) extrap_rows_body(A, B, vb, nvb);
    else if (WHAT == BX_HIST) extrap_hist_body(A, B, vb, nvb);
    else if (WHAT == BX_SCATTER) extrap_scatter_body(A, B, vb, nvb);
    else extrap_link_body(A, B, vb, nvb);
}
__global__ __launch_bounds__(kExtraColThreads) void b_extrap_cols_kernel(const InstBufs* __restrict__ ins, BlockMap bm,
                                                                         int which) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    ExtraArr A, B;
    b_extra_arrays(ins[b], which, A, B);
    extrap_cols_body(A, B, vb, nvb);
}
// Two workgroups per simulation: u and v.
__global__ __launch_bounds__(kBlock) void b_extrap_scan_kernel(const InstBufs* __restrict__ ins, int n, int which) {
    const int b = blockIdx.x >> 1;
    if (b >= n) return;
    ExtraArr A, B;
    b_extra_arrays(ins[b], which, A, B);
    extrap_scan_body(A, B, blockIdx.x & 1);
}
__global__ __launch_bounds__(1024) void b_extrap_values_kernel(const InstBufs* __restrict__ ins, int n, int which) {
    const int b = blockIdx.x >> 1;
    if (b >= n) return;
    ExtraArr A, B;
    b_extra_arrays(ins[b], which, A, B);
    extrap_values_body(A, B, blockIdx.x & 1);
}

// ---------------------------------------------------------------------------
// Particles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void b_particle_keys_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    particle_keys_body(I.px, I.py, I.np, I.dx, I.nx, I.ny, I.pkey, I.counts, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_scan_tiles_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    scan_tiles_body(I.counts, I.starts, I.tile_sums, (long)I.nx * I.ny, vb, nvb);
}
__global__ void b_scan_sums_kernel(const InstBufs* __restrict__ ins, int n) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const InstBufs& I = ins[b];
    scan_sums_body(I.tile_sums, (int)(((long)I.nx * I.ny + 1023) / 1024), 0, 1);
}
__global__ __launch_bounds__(kBlock) void b_scan_add_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    scan_add_body(I.starts, I.tile_sums, (long)I.nx * I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_particle_scatter_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    particle_scatter_body(I.pkey, I.np, I.cursor, I.sorted, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_segment_sort_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    segment_sort_body(I.starts, I.counts, I.sorted, (long)I.nx * I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_p2g_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    p2g_body(I.px, I.py, I.pvx, I.pvy, I.starts, I.counts, I.sorted, I.u, I.v, I.mu, I.mv, I.nx, I.ny, I.dx, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_g2p_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    g2p_body(I.px, I.py, I.pvx, I.pvy, I.np, I.nu, I.nv, I.u, I.v, I.nx, I.ny, I.dx, (float)I.alpha, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_particle_advect_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    particle_advect_body(I.px, I.py, I.pvx, I.pvy, I.np, make_mac(I.u, I.v, I.nx, I.ny, I.dx), I.dt,
                         clamp_box(I.nx, I.ny, I.dx), I.dg, I.partials, (I.flags & 32) ? 1 : 0 /* FS2D_FLAG_RK2 */,
                         vb, nvb);
}
// One workgroup per simulation; `bm` is the map of the launch that wrote the partials.
__global__ __launch_bounds__(kBlock) void b_cfl_finalize_kernel(const InstBufs* __restrict__ ins, int n, BlockMap bm) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const InstBufs& I = ins[b];
    if (I.np == 0) return;
    cfl_finalize_body(I.partials, bm.count[b], I.dg, 0, 1);
}

// ---------------------------------------------------------------------------
// Level set.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void b_ls_seed_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    ls_seed_body(I.px, I.py, I.starts, I.counts, I.sorted, I.t0, I.nx, (long)I.nx * I.ny, I.dx, I.dr, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_ls_phi_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    ls_phi_body(I.t0, I.px, I.py, I.phi2, I.nx, (long)I.nx * I.ny, I.dx, I.dr, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_ls_surface_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    ls_surface_body(I.phi2, I.phi, I.nx, I.ny, vb, nvb);
}
// Smoothing pass `parity`: phi -> phi2 (0) or phi2 -> phi (1).
__global__ __launch_bounds__(kBlock) void b_ls_smooth_kernel(const InstBufs* __restrict__ ins, BlockMap bm,
                                                             int parity) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    ls_smooth_body(parity ? I.phi2 : I.phi, parity ? I.phi : I.phi2, I.nx, I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_classify_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    classify_body(I.cell, I.phi, (long)I.nx * I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_diagnostics_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    diagnostics_body(I.cell, make_mac(I.u, I.v, I.nx, I.ny, I.dx), I.px, I.py, I.pvx, I.pvy, I.np, I.rho, I.gx, I.gy,
                     I.pmass, I.partials, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_diag_finalize_kernel(const InstBufs* __restrict__ ins, int n,
                                                                 BlockMap bm) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const InstBufs& I = ins[b];
    diag_finalize_body(I.partials, bm.count[b], I.dx, I.dg, 0, 1);
}

// Flags of every simulation's solver and diagnostics, gathered for one host
// read: out[4b] done, out[4b+1] iterations, out[4b+2] NaN flag, out[4b+3]
// CFL warning (particle CFL > 5, src/FluidSim2D.cpp:582-585).
__global__ void b_gather_status_kernel(const InstBufs* __restrict__ ins, int n, int* out) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        out[4 * b] = ins[b].sc->done;
        out[4 * b + 1] = ins[b].sc->iter;
        out[4 * b + 2] = ins[b].dg->nan_flag;
        out[4 * b + 3] = (ins[b].np > 0 && ins[b].dg->cfl_max > 5.0) ? 1 : 0;
    }
}

// Cells from particles (createWaterLevelSet without USE_LEVEL_SET).
__global__ __launch_bounds__(kBlock) void b_pc_clear_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    pc_clear_body(I.cell, I.cell2, (long)I.nx * I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_pc_mark_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    pc_mark_body(I.px, I.py, I.np, I.cell, I.cell2, I.nx, I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_pc_copy_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    copy_body(I.cell, (const uint8_t*)I.cell2, (size_t)I.nx * I.ny, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_pc_fill_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    pc_fill_body(I.cell, I.cell2, I.nx, I.ny, vb, nvb);
}

// Solver passes.  check: the forward solve applies the stopping test first;
// h: the backward solve of iteration h (-1: the initial solve).
struct FactorMaker {
    const int* nparts;  // workgroups of the SK launches per simulation (the rhs partials)
    __device__ FactorOp operator()(const InstBufs& I, int b) const {
        return FactorOp{{I.Adiag}, I.cpl, {I.P, I.fx, I.fy, I.gx_, I.gy_}, I.tau, I.sigma, I.rmax_part,
                        nparts[b], I.max_iters, I.sc, I.rhs_zero};
    }
};
// The eikonal passes on the sweep engine (EikSweepOp): conversions between a
// member's phi (row-major, ping-pong by parity) and the SK4 order of a
// direction's flipped coordinates, in the member's solver scratch.
__global__ __launch_bounds__(kBlock) void b_eik_to_sk_kernel(const InstBufs* __restrict__ ins, BlockMap bm,
                                                             int parity, int fx, int fy) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    eik_to_sk_body(I.geom, parity ? I.phi2 : I.phi, I.z, I.q, I.cpl, fx, fy, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_eik_from_sk_kernel(const InstBufs* __restrict__ ins, BlockMap bm,
                                                               int parity, int fx, int fy) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    eik_from_sk_body(I.geom, I.w, I.s1, parity ? I.phi : I.phi2, fx, fy, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_eik_sk_to_sk_kernel(const InstBufs* __restrict__ ins, BlockMap bm,
                                                                int fx, int fy, int fx2, int fy2) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    eik_sk_to_sk_body(I.geom, I.w, I.s1, I.z, I.q, I.cpl, fx, fy, fx2, fy2, vb, nvb);
}
// The closest-particle passes on the sweep engine (ClosestSweepOp), as in the
// single path: candidates from t0 (first pass) or from the previous pass's
// output (w, SK4 order of direction pdir); the last output back to t0.
__global__ __launch_bounds__(kBlock) void b_ls_cand_sk_kernel(const InstBufs* __restrict__ ins, BlockMap bm, int dir,
                                                              int first, int pdir, int k) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    ls_cand_sk_body(I.geom, first ? I.t0 : nullptr, I.w, pdir & 1, pdir >> 1, I.px, I.py, I.r, I.z, I.q, dir & 1,
                    dir >> 1, I.dx, I.dr, I.ls_chg, k, vb, nvb);
}
__global__ __launch_bounds__(kBlock) void b_closest_t_kernel(const InstBufs* __restrict__ ins, BlockMap bm) {
    int b, vb, nvb;
    if (!bmap(bm, b, vb, nvb)) return;
    const InstBufs& I = ins[b];
    closest_t_to_rm_body(I.geom, I.w, I.t0, I.ls_chg, vb, nvb);
}
// Every member's 16 pass change flags cleared (one thread per flag).
__global__ __launch_bounds__(kBlock) void b_ls_chg_clear_kernel(const InstBufs* __restrict__ ins, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 16 * n) ins[e >> 4].ls_chg[e & 15] = 0;
}
template <int FX, int FY>
struct ClosestSweepMaker {
    int k;
    __device__ ClosestSweepOp<FX, FY> operator()(const InstBufs& I, int) const {
        return ClosestSweepOp<FX, FY>{{I.r, I.z, I.q}, nullptr, {I.w}, I.nx, I.ny, I.dx, I.dr, I.ls_chg, k};
    }
};
struct EikSweepMaker {
    __device__ EikSweepOp operator()(const InstBufs& I, int) const {
        return EikSweepOp{{I.z, I.q}, I.cpl, {I.w, I.s1}, I.dx};
    }
};
struct FwdMaker {
    const int* nparts;
    int check;
    __device__ FwdOp operator()(const InstBufs& I, int b) const {
        return FwdOp{{I.r, I.P, I.fx, I.fy}, nullptr, {I.q}, I.sc, I.rmax_part, nparts[b], check};
    }
};
struct BwdMaker {
    int h;
    const int* nparts;  // with check: the members' p/r update partial counts
    int check;          // the stopping test of iteration h (after b_fwd_axpy_kernel)
    __device__ BwdOp operator()(const InstBufs& I, int b) const {
        // sigma_{h+1} goes where sigma_{h-1} was: the half of parity h + 1
        double* qq = I.qq + (size_t)(((h + 1) & 1) ? I.geom.S * kLoaders : 0);
        return BwdOp{{I.q, I.P, I.gx_, I.gy_}, nullptr, {I.z}, I.sc, qq, h + 1, nullptr, 0u, check, I.rmax_part,
                     check ? nparts[b] : 0};
    }
};
// The batched passes use the single launch's ring depth (sweep_ring).  A
// 4-supergroup ring would halve the LDS, but the sweep workgroup's 134 VGPRs
// already keep a second one off the CU, and the shorter lookahead made the
// ensemble's solves slower.  The ring depth changes no arithmetic.
// PREFETCH: loaders issue their first loads before the skip test (which a
// skipped strip then wastes).  all_done_since (PCG solves): when every member
// had converged before iteration h, the first workgroup advances the ticket
// by the launch's strip count and every workgroup returns, so the epochs of
// later launches stay as if every strip had taken its ticket.
template <class Op, bool BACKWARD, class Maker, bool PREFETCH = true>
// With `active` (the PCG plan's flags of this iteration), a strip of a member
// that had converged before it returns once it holds its ticket, without the
// loads its skip test would discard: its done flag makes skip() true anyway.
// The flag is read beside the member's descriptor, so an iterating member's
// strips wait no longer.
__global__ __launch_bounds__(kSweepThreads) void b_sweep_kernel(const InstBufs* __restrict__ ins, StripMap sm,
                                                                unsigned long long* ticket, int* err, Maker mk,
                                                                const int* all_done_since = nullptr, int h = 0,
                                                                const int* active = nullptr) {
    __shared__ __align__(16) SweepLds<Op, sweep_ring<Op>()> L;
    if (all_done_since && *all_done_since <= h) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(ticket, (unsigned long long)sm.total);
        return;
    }
    const unsigned long long t = sweep_ticket(L, ticket);
    const int gs = (int)(t % (unsigned long long)sm.total);
    const unsigned epoch = (unsigned)(t / (unsigned long long)sm.total) + 1u;
    const int b = sm.inst[gs];
    const bool idle = active && active[b] == 0;
    const InstBufs& I = ins[b];
    Op op = mk(I, b);
    const SkCtl c{ticket, I.sk_handoff, err};
    if (idle) return;
    sweep_run<Op, BACKWARD, PREFETCH>(op, I.geom, c, gs - sm.first[b], epoch, L);
}


// Iteration h's p/r update and forward solves of every member in one launch,
// as fwd_axpy_kernel does for one simulation: the waves of M workgroups run
// the plan's virtual workgroups, one at a time each, as strip-ordered runs of
// their member (pcg_run, axpy_run_wave; alpha as b_axpy_kernel forms it; r
// stored with device scope, each run flagged in its member's fa_flags after
// the wave's stores complete), beside the
// strips of every member, whose loaders wait for their member's runs over
// their strip.  The stopping test moves to the backward launch (BwdMaker
// check).  Each entry's update and the maximum of |r| are the separate
// launches', so members stay bit-identical to single simulations.
template <bool PREFETCH>
__global__ __launch_bounds__(kSweepThreads) void b_fwd_axpy_kernel(const InstBufs* __restrict__ ins,
                                                                   const int* __restrict__ count, PcgPlanBufs P,
                                                                   StripMap sm, unsigned long long* ticket, int* err,
                                                                   int M, const int* all_done_since, int h) {
    __shared__ __align__(16) SweepLds<FwdFusedOp, sweep_ring<FwdFusedOp>()> L;
    const unsigned long long per = (unsigned long long)(M + sm.total);
    if (*all_done_since <= h) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(ticket, per);
        return;
    }
    const unsigned long long t = sweep_ticket(L, ticket);
    const int idx = (int)(t % per);
    const unsigned epoch = (unsigned)(t / per) + 1u;
    if (idx < M) {
        // runner w M + idx (wave w) takes the plan's entries k = runner,
        // runner + 6 M, ... one at a time (axpy_run_wave)
        constexpr int kWaves = kSweepThreads / 64;
        const int total = *P.total, runners = kWaves * M;
        RunBuf Rb;
        for (int k = (int)(threadIdx.x >> 6) * M + idx; k < total; k += runners) {
            const int b = P.map[k];
            if (b < 0) continue;  // padding to a multiple of 8
            const int vb = k - P.first[b];
            const InstBufs& I = ins[b];
            const float a = (float)(I.sc->sigma / I.sc->ws);
            const __amdgpu_buffer_rsrc_t r_rs =
                __builtin_amdgcn_make_buffer_rsrc(I.r, 0, (int)(I.geom.size() * 4), kBufferCfg);
            long q0, q1;
            pcg_run(I.geom.size(), vb, count[b], q0, q1);
            axpy_run_wave(Rb, (const float4*)((h & 1) ? I.s0 : I.s1), (const float4*)I.w, (float4*)I.pk,
                          (const float4*)I.r, r_rs, a, q0, q1, I.rmax_part + vb, I.fa_flags + vb, epoch);
        }
        return;
    }
    const int gs = idx - M;
    const int b = sm.inst[gs];
    const InstBufs& I = ins[b];
    FwdFusedOp op{{I.r, I.P, I.fx, I.fy}, nullptr, {I.q}, I.sc, I.rmax_part, count[b], 0, I.fa_flags, I.geom, epoch};
    const SkCtl c{ticket, I.fa_handoff, err};
    sweep_run<FwdFusedOp, false, PREFETCH>(op, I.geom, c, gs - sm.first[b], epoch, L);
}

}  // namespace fs2d
