// fs2d_kernels.hpp -- gfx950 kernels of the 2D step.
//
// Layout in HBM (all row-major, element (i, j) at [j * width + i]):
//   u, u2, newu   FP32 (nx+1) x ny        v, v2, newv   FP32 nx x (ny+1)
//   p             FP32 nx x ny (row-major copy of the solver's pressure)
//   PCG vectors and coefficients (r, z, q, w, s[2], p, Adiag, Ax, Ay, P, fx,
//   fy, gx, gy) FP32 in the strip-skewed SK4 layout of fs2d_skew.hpp
//   cell          uint8 nx x ny           phi[2]        FP64 nx x ny
//   tidx[2]       int32 nx x ny (closest particle of the level-set sweeps)
//   px, py, pvx, pvy  FP32 SoA, particle order as created
// Reductions are deterministic and fence-free: each block writes a partial and
// a later kernel (kernel boundaries order the writes) combines the partials in
// index order.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fs2d_device.hpp"
#include "fs2d_gmem.hpp"
#include "fs2d_sweep.hpp"

namespace fs2d {

constexpr int kBlock = 256;

struct Diag {  // createWaterLevelSet diagnostics (src/FluidSim2D.cpp:709-731)
    double volume, grid_energy, particle_energy, cfl_max;
    int nan_flag, pad;
};

// ---------------------------------------------------------------------------
// Block reductions and the last-block protocol.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T x = __shfl_xor(v, o, 64);
        v = x > v ? x : v;
    }
    return v;
}
// Sum over a 256-thread block in a fixed order; result valid in thread 0.  In
// a larger block only the first 256 threads contribute (the others still take
// part in the barriers).
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kBlock / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && threadIdx.x < kBlock) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; k++) t += sh[k];
    return t;
}
__device__ __forceinline__ double block_max(double v) {
    __shared__ double sh[kBlock / 64];
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && threadIdx.x < kBlock) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; k++) t = sh[k] > t ? sh[k] : t;
    return t;
}
// Sum / max of partials[0..n) by a whole block in a fixed order; result in thread 0.
__device__ __forceinline__ double sum_partials(const double* partials, int n) {
    double acc = 0;
    for (int k = threadIdx.x; k < n; k += kBlock) acc += partials[k];
    return block_sum(acc);
}
__device__ __forceinline__ double max_partials(const double* partials, int n) {
    double acc = 0;
    for (int k = threadIdx.x; k < n; k += kBlock) acc = partials[k] > acc ? partials[k] : acc;
    return block_max(acc);
}

template <class T>
__device__ __forceinline__ void fill_body(T* a, T v, size_t n, int vb, int nvb) {
    for (size_t e = vb * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)nvb * blockDim.x)
        a[e] = v;
}
template <class T>
__global__ void fill_kernel(T* a, T v, size_t n) {
    fill_body(a, v, n, blockIdx.x, gridDim.x);
}
template <class T>
__device__ __forceinline__ void copy_body(T* dst, const T* src, size_t n, int vb, int nvb) {
    for (size_t e = vb * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)nvb * blockDim.x)
        dst[e] = src[e];
}
template <class T>
__global__ void copy_kernel(T* dst, const T* src, size_t n) {
    copy_body(dst, src, n, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Semi-Lagrangian advection (src/FluidSim2D.cpp:206-235), out of place, with
// applyGravity (:237-250) optionally fused into the store, and the passive
// density (no reference counterpart) traced back from cell centres.
// ---------------------------------------------------------------------------
// The stage runs over tiles of kSlTX x kSlTY cells, one at a time per
// workgroup: a tile's outputs are the u faces, v faces and density cells whose
// index (i, j) falls in it.  The velocities (and density) of the tile and a
// halo of kSlHalo cells are copied to LDS first, coalesced; a sample whose 16
// taps all lie in that window reads LDS, any other (a departure point more
// than about kSlHalo - 2.5 cells away, or the window's clipped border) reads
// the arrays: the same bits either way.  Departure points within 1.5 cells
// (CFL <= 1.5) read LDS only.
constexpr int kSlTX = 64, kSlTY = 16, kSlHalo = 4;
constexpr int kSlUW = kSlTX + 1 + 2 * kSlHalo, kSlUH = kSlTY + 2 * kSlHalo;
constexpr int kSlVW = kSlTX + 2 * kSlHalo, kSlVH = kSlTY + 1 + 2 * kSlHalo;
constexpr int kSlDW = kSlTX + 2 * kSlHalo, kSlDH = kSlTY + 2 * kSlHalo;
__host__ __device__ inline int sl_tiles(int nx, int ny) {
    return ((nx + 1 + kSlTX - 1) / kSlTX) * ((ny + 1 + kSlTY - 1) / kSlTY);
}
struct SlLds {
    float u[kSlUW * kSlUH];
    float v[kSlVW * kSlVH];
    float d[kSlDW * kSlDH];
};
// Copies columns x0 .. x0 + ww - 1 and rows y0 .. y0 + wh - 1 of the w x h
// array a, clipped to the array, into lds (row stride = the clipped width).
__device__ __forceinline__ LdsWin load_win(float* lds, const float* __restrict__ a, int w, int h, int x0, int y0,
                                           int ww, int wh) {
    const int xa = max(x0, 0), ya = max(y0, 0);
    const int cw = min(x0 + ww, w) - xa, ch = min(y0 + wh, h) - ya;
    for (int k = threadIdx.x; k < cw * ch; k += blockDim.x) {
        const int r = k / cw, c = k - r * cw;
        lds[k] = a[(size_t)(ya + r) * w + xa + c];
    }
    return LdsWin{lds, xa, ya, cw, ch};
}

__device__ __forceinline__ void sl_advect_body(MacView m, float* __restrict__ u_out,
                                                           float* __restrict__ v_out,
                                                           const float* __restrict__ d_in,
                                                           float* __restrict__ d_out, double dt,
                                                           ClampBox cb, float gu, float gv, int vb, int nvb) {
    __shared__ SlLds L;
    const int nx = m.nx, ny = m.ny;
    const int ntx = (nx + 1 + kSlTX - 1) / kSlTX, nt = sl_tiles(nx, ny);
    for (int t = vb; t < nt; t += nvb) {
        const int X0 = (t % ntx) * kSlTX, Y0 = (t / ntx) * kSlTY;
        __syncthreads();  // the previous tile's samples are done with L
        const LdsWin U = load_win(L.u, m.u, nx + 1, ny, X0 - kSlHalo, Y0 - kSlHalo, kSlUW, kSlUH);
        const LdsWin V = load_win(L.v, m.v, nx, ny + 1, X0 - kSlHalo, Y0 - kSlHalo, kSlVW, kSlVH);
        LdsWin D{};
        if (d_in) D = load_win(L.d, d_in, nx, ny, X0 - kSlHalo, Y0 - kSlHalo, kSlDW, kSlDH);
        __syncthreads();
        const WinSampler S{m, U, V};
        for (int k = threadIdx.x; k < kSlTX * kSlTY; k += blockDim.x) {
            const int i = X0 + k % kSlTX, j = Y0 + k / kSlTX;
            double ox, oy;
            if (i <= nx && j < ny) {  // u face (i, j + 1/2)
                backtrace_with(S, dt, cb, (double)i * m.dx, ((double)j + 0.5) * m.dx, ox, oy);
                u_out[(size_t)j * (nx + 1) + i] = S.u(ox, oy) + gu;
            }
            if (i < nx && j <= ny) {  // v face (i + 1/2, j)
                backtrace_with(S, dt, cb, ((double)i + 0.5) * m.dx, (double)j * m.dx, ox, oy);
                v_out[(size_t)j * nx + i] = S.v(ox, oy) + gv;
            }
            if (d_in && i < nx && j < ny) {  // density at the cell centre
                backtrace_with(S, dt, cb, ((double)i + 0.5) * m.dx, ((double)j + 0.5) * m.dx, ox, oy);
                const double gx = clampd(ox * m.inv_dx - 0.5, 1e-6, (double)(nx - 1) - 1e-6);
                const double gy = clampd(oy * m.inv_dx - 0.5, 1e-6, (double)(ny - 1) - 1e-6);
                d_out[(size_t)j * nx + i] = catmull_rom_win(D, d_in, nx, ny, gx, gy);
            }
        }
    }
}
__global__ __launch_bounds__(kBlock) void sl_advect_kernel(MacView m, float* __restrict__ u_out,
                                                           float* __restrict__ v_out,
                                                           const float* __restrict__ d_in,
                                                           float* __restrict__ d_out, double dt,
                                                           ClampBox cb, float gu, float gv) {
    sl_advect_body(m, u_out, v_out, d_in, d_out, dt, cb, gu, gv, blockIdx.x, gridDim.x);
}

// FluidSim2D::applyGravity (src/FluidSim2D.cpp:237-250): every face.
__device__ __forceinline__ void gravity_body(float* __restrict__ u, float* __restrict__ v, long nu, long nv, float gu,
                               float gv, int vb, int nvb) {
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < nu + nv; e += (long)nvb * blockDim.x) {
        if (e < nu) u[e] += gu;
        else v[e - nu] += gv;
    }
}
__global__ void gravity_kernel(float* __restrict__ u, float* __restrict__ v, long nu, long nv, float gu,
                               float gv) {
    gravity_body(u, v, nu, nv, gu, gv, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Projection (src/FluidSim2D.cpp:252-467) on the SK4 layout (fs2d_skew.hpp).
// ---------------------------------------------------------------------------
// Element loop of a stencil kernel over an SK4 array.  Workgroups go to the 8
// XCDs round-robin (workgroup b on XCD b % 8) and each XCD has its own L2, so a
// grid-stride loop would fetch a cell's neighbours (one SK4 group, 256
// elements, away) through another XCD's L2.  With a workgroup count that is a
// multiple of 8, XCD x's workgroups share the contiguous range [x n/8,
// (x+1) n/8) instead; otherwise this is a plain grid-stride loop.  vb / nvb
// are the workgroup's index and count within its simulation (a batched launch
// starts every simulation's workgroups at a multiple of 8, so vb % 8 is still
// the XCD).
struct XcdLoop {
    long e, end, stride;
};
__device__ __forceinline__ XcdLoop xcd_loop(long n, int vb, int nvb) {
    const int nb = nvb;
    if (nb & 7) return {vb * (long)blockDim.x + threadIdx.x, n, (long)nb * blockDim.x};
    const int x = vb & 7, k = vb >> 3;
    const long chunk = (n + 7) / 8;
    const long b0 = x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
    return {b0 + k * (long)blockDim.x + threadIdx.x, b1, (long)(nb >> 3) * blockDim.x};
}

// Decode an SK4 element index into its cell; false for padding entries.
__device__ __forceinline__ bool sk_cell(const SkGeom& g, long e, int& i, int& j) {
    const int e4 = (int)(e & 3), l = (int)((e >> 2) & 63);
    // 32-bit quotient: e >> 8 < 2^32 for any array that fits in memory
    const unsigned r = (unsigned)(e >> 8), t4 = (unsigned)g.T4;
    const unsigned sq = r / t4;
    const int grp = (int)(r - sq * t4), s = (int)sq;
    const int t = grp * 4 + e4;
    i = t - l;
    j = s * 64 + l;
    return i >= 0 && i < g.nx && j < g.ny;
}

// Ghost-fluid matrix (:262-303) and rhs (:338-362) written in SK4 order, in
// solver units (the reference's matrix and rhs divided by scaleA = dt / (rho
// dx^2)): the diagonal, and the coupling bits of each FLUID cell (its
// off-diagonals are -1 towards FLUID neighbours).  `scale` is the rhs factor
// 1 / (dx scaleA).  The PCG vectors p and s start at zero (:424).  Partial
// |rhs|_inf per block.
// Tiles of 64 rows (a strip) x 32 steps, a tile per workgroup and round
// (vb, vb + nvb, ...): each wave takes two rows of 32 consecutive steps at a
// time, so its reads of the row-major cell, phi, u and v rows are coalesced,
// and a tile's SK4 stores fill whole cache lines (a line holds 8 rows x 4
// steps) before they leave the L2.  Padding entries (no cell) get zeros like
// the cells without an equation.  partials[vb]: the largest |rhs| of the
// workgroup's tiles.
__device__ __forceinline__ void assemble_sk_body(
    SkGeom g, const uint8_t* __restrict__ cell, const double* __restrict__ phi, const float* __restrict__ u,
    const float* __restrict__ v, float* __restrict__ Adiag, uint8_t* __restrict__ flags, float* __restrict__ r,
    float* __restrict__ p, float* __restrict__ s0, double scale, double* partials, int ghost, int vb, int nvb) {
    constexpr double scaleA = 1.0;
    const int nx = g.nx;
    const int tiles_per_strip = g.T4 / 8;
    const long ntiles = (long)g.S * tiles_per_strip;
    const int rr = threadIdx.x >> 5, tt = threadIdx.x & 31;
    double m = 0;
    for (long tile = vb; tile < ntiles; tile += nvb) {
        const int s = (int)(tile / tiles_per_strip), t = (int)(tile % tiles_per_strip) * 32 + tt;
        for (int l = rr; l < 64; l += 8) {
            const int i = t - l, j = s * 64 + l;
            const long e = (((long)s * g.T4 + (t >> 2)) * 64 + l) * 4 + (t & 3);
            double ad = 0, rhs = 0;
            uint32_t f = 0;
            if (i >= 0 && i < nx && j < g.ny) {
                const long c = (long)j * nx + i;
                if (cell[c] == FLUID) {
                    const uint8_t cl = cell[c - 1], cr = cell[c + 1], cb = cell[c - nx], ct = cell[c + nx];
                    if (ghost) {
                        const double ph = phi[c];
                        if (cl == FLUID) { ad += scaleA; f |= CPL_L; }
                        else if (cl == EMPTY) ad -= scaleA * fmax(phi[c - 1] / ph, -1e3);
                        if (cr == FLUID) { ad += scaleA; f |= CPL_R; }
                        else if (cr == EMPTY) ad += scaleA * (1 - fmax(phi[c + 1] / ph, -1e3));
                        if (cb == FLUID) { ad += scaleA; f |= CPL_D; }
                        else if (cb == EMPTY) ad -= scaleA * fmax(phi[c - nx] / ph, -1e3);
                        if (ct == FLUID) { ad += scaleA; f |= CPL_U; }
                        else if (ct == EMPTY) ad += scaleA * (1 - fmax(phi[c + nx] / ph, -1e3));
                    } else {
                        // without USE_GHOST_PRESSURE (:304-331): only the +x and +y EMPTY
                        // neighbours add to the diagonal
                        if (cl == FLUID) { ad += scaleA; f |= CPL_L; }
                        if (cr == FLUID) { ad += scaleA; f |= CPL_R; }
                        else if (cr == EMPTY) ad += scaleA;
                        if (cb == FLUID) { ad += scaleA; f |= CPL_D; }
                        if (ct == FLUID) { ad += scaleA; f |= CPL_U; }
                        else if (ct == EMPTY) ad += scaleA;
                    }
                    const long iu = (long)j * (nx + 1) + i;
                    const double u0 = u[iu], u1 = u[iu + 1], v0 = v[c], v1 = v[c + nx];
                    rhs = -scale * (u1 - u0 + v1 - v0);
                    if (cl == SOLID) rhs -= scale * u0;
                    if (cr == SOLID) rhs += scale * u1;
                    if (cb == SOLID) rhs -= scale * v0;
                    if (ct == SOLID) rhs += scale * v1;
                }
            }
            Adiag[e] = (float)ad;
            flags[e] = (uint8_t)f;
            r[e] = (float)rhs;
            p[e] = 0.0f;
            s0[e] = 0.0f;
            const double a = fabs((double)(float)rhs);
            m = a > m ? a : m;
        }
    }
    double bm = block_max(m);
    if (threadIdx.x == 0) partials[vb] = bm;
}
__global__ __launch_bounds__(kBlock) void assemble_sk_kernel(
    SkGeom g, const uint8_t* __restrict__ cell, const double* __restrict__ phi, const float* __restrict__ u,
    const float* __restrict__ v, float* __restrict__ Adiag, uint8_t* __restrict__ flags, float* __restrict__ r,
    float* __restrict__ p, float* __restrict__ s0, double scale, double* partials, int ghost) {
    assemble_sk_body(g, cell, phi, u, v, Adiag, flags, r, p, s0, scale, partials, ghost, blockIdx.x, gridDim.x);
}

// Sum of a small partial array by one wave, broadcast to every thread of the
// block (fixed order, identical in every block).
__device__ __forceinline__ double block_sum_bcast(const double* a, int n) {
    __shared__ double sh;
    __syncthreads();  // earlier readers of sh are done
    if (threadIdx.x < 64) {
        double t = wave_sum_of(a, n);
        if (threadIdx.x == 0) sh = t;
    }
    __syncthreads();
    return sh;
}

// s' = z + beta s (:459, FMA as Array2D.h:185), w = A s' (:433-444), partial
// w.s' (:450).  sigma_k = z.r comes from the forward solves' per-strip q.q
// partials; beta = sigma_k / sigma_{k-1} (:457-458).  s is double-buffered
// because every cell also needs s' of its four neighbours.
//
// Neighbour offsets in SK4 (element e: step q = e & 3, lane l = (e >> 2) & 63):
// (i-1, j) is step t-1 of the lane, (i+1, j) step t+1; (i, j-1) is lane l-1 at
// step t-1 and (i, j+1) lane l+1 at step t+1, both inside the strip except on
// its edge lanes, which take the general index.  FLUID cells are interior, so
// every neighbour exists.
__device__ __forceinline__ long sk_step_back(long e) { return (e & 3) ? e - 1 : e - 253; }    // t-1, same lane
__device__ __forceinline__ long sk_step_fwd(long e) { return (e & 3) != 3 ? e + 1 : e + 253; }  // t+1, same lane
// (i, j-1) of an entry in lane 0: lane 63 of the strip below at step t+63;
// (i, j+1) of an entry in lane 63: lane 0 of the strip above at step t-63.
// Closed forms of SkGeom::idx from the entry's group r = e >> 8 and step
// residue e4 = e & 3 (no decode of the cell, no division).
__device__ __forceinline__ long sk_below_edge(const SkGeom& g, long e) {
    const long e4 = e & 3;
    return (((e >> 8) - g.T4 + 15 + ((e4 + 3) >> 2)) * 64 + 63) * 4 + ((e4 + 3) & 3);
}
__device__ __forceinline__ long sk_above_edge(const SkGeom& g, long e) {
    const long e4 = e & 3;
    return ((e >> 8) + g.T4 + ((e4 - 63) >> 2)) * 256 + ((e4 + 1) & 3);
}
// Entries of SpMV virtual workgroup vb of nvb (256 threads): a contiguous run
// of whole 256-entry groups, the n / 256 groups split as evenly as possible;
// thread tid takes entry tid of each group (threads past 256 of a larger block
// take none).  A strip's entries are contiguous, so a run touches one or two
// strips (and reads one more on either side): the fused backward launch starts
// a run as soon as those strips have finished.  The neighbours (i +- 1, j) at
// +-253 lie in the run's own groups except at its two ends.
__device__ __forceinline__ XcdLoop spmv_range(long n, int vb, int nvb, int tid) {
    if (tid >= kBlock) return {n, n, 1};
    const long ng = n >> 8;
    const long g0 = ng * vb / nvb, g1 = ng * (vb + 1) / nvb;
    return {g0 * 256 + tid, g1 * 256, 256};
}
// One virtual workgroup (vb of nvb) of the SpMV: the partial w.s of its
// entries (spmv_range) goes to partials[vb].  The entries' own values do not
// depend on beta, so the first entry's are loaded (spmv_pre) before the
// reductions that give beta, and each later entry's at the end of the
// previous round.
struct SpmvPre {
    XcdLoop it;
    float so, zc, ad;
    uint32_t f;
};
__device__ __forceinline__ SpmvPre spmv_pre(const SkGeom& g, const float* __restrict__ Adiag,
                                            const uint8_t* __restrict__ flags, const float* __restrict__ z,
                                            const float* __restrict__ s_old, int vb, int nvb) {
    SpmvPre r;
    r.it = spmv_range(g.size(), vb, nvb, (int)threadIdx.x);
    r.so = r.zc = r.ad = 0.0f;
    r.f = 0;
    const long e = r.it.e;
    if (e < r.it.end) {
        r.so = s_old[e];
        r.zc = z[e];
        r.ad = Adiag[e];
        r.f = flags[e];
    }
    return r;
}
// sigma_k = z.r from the backward solves' per-strip q.q partials and beta =
// sigma_k / sigma_{k-1} (:457-458), reduced by the whole block in a fixed
// order (the same in every workgroup).
struct SpmvScalars {
    double sig;
    float beta;
};
__device__ __forceinline__ SpmvScalars spmv_scalars(const SkGeom& g, const double* sig_cur, const double* sig_prev,
                                                    int first) {
    SpmvScalars r;
    r.sig = block_sum_bcast(sig_cur, g.S * kLoaders);
    double beta_d = 0;
    if (!first) beta_d = r.sig / block_sum_bcast(sig_prev, g.S * kLoaders);
    r.beta = (float)beta_d;
    return r;
}
// The matrix is in solver units (off-diagonals -1): w = ad s' - sum of the
// coupled neighbours' s', with one rounding (an explicit fma, so that no
// launch path can contract it differently).
__device__ __forceinline__ void spmv_core(const SpmvPre& pre, const SkGeom& g, const float* __restrict__ Adiag,
                                          const uint8_t* __restrict__ flags, const float* __restrict__ z,
                                          const float* __restrict__ s_old, float* __restrict__ s_new,
                                          float* __restrict__ w, float beta, double* partials, int vb) {
    const XcdLoop it = pre.it;
    long e = it.e;
    float so = pre.so, zc = pre.zc, ad = pre.ad;
    uint32_t f = pre.f;
    double acc = 0;
    while (e < it.end) {
        const float sc_ = fmaf(beta, so, zc);
        s_new[e] = sc_;
        float wv = 0.0f;
        if (ad != 0.0f) {  // FLUID: interior, all four neighbours exist
            const int l = (int)((e >> 2) & 63);
            const long nd = l > 0 ? sk_step_back(e) - 4 : sk_below_edge(g, e);
            const long nu = l < 63 ? sk_step_fwd(e) + 4 : sk_above_edge(g, e);
            float nsum = 0.0f;
            if (f & CPL_L) { const long n = sk_step_back(e); nsum += fmaf(beta, s_old[n], z[n]); }
            if (f & CPL_R) { const long n = sk_step_fwd(e); nsum += fmaf(beta, s_old[n], z[n]); }
            if (f & CPL_D) nsum += fmaf(beta, s_old[nd], z[nd]);
            if (f & CPL_U) nsum += fmaf(beta, s_old[nu], z[nu]);
            wv = fmaf(ad, sc_, -nsum);
        }
        w[e] = wv;
        acc += (double)wv * (double)sc_;
        e += it.stride;
        if (e < it.end) {
            so = s_old[e];
            zc = z[e];
            ad = Adiag[e];
            f = flags[e];
        }
    }
    double bs = block_sum(acc);
    if (threadIdx.x == 0) partials[vb] = bs;
}
// spmv_core for a launch of few workgroups per CU: U entries of the thread
// at a time, every load of them (the entry, and both operands of all four
// neighbours, indices clamped into the array) issued together, so a virtual
// workgroup pays two memory round trips per U entries instead of two per
// entry.  The same entries, arithmetic and summation order as spmv_core.
template <int U>
__device__ __forceinline__ void spmv_core_wide(const XcdLoop it, const SkGeom& g, const float* __restrict__ Adiag,
                                               const uint8_t* __restrict__ flags, const float* __restrict__ z,
                                               const float* __restrict__ s_old, float* __restrict__ s_new,
                                               float* __restrict__ w, float beta, double* partials, int vb) {
    const long last = g.size() - 1;
    double acc = 0;
    for (long e0 = it.e; e0 < it.end; e0 += U * it.stride) {
        float so[U], zc[U], ad[U], ns[U][4], nz[U][4];
        uint32_t f[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long e = e0 + u * it.stride;
            const long ec = e < it.end ? e : e0;
            const int l = (int)((ec >> 2) & 63);
            long nb[4] = {sk_step_back(ec), sk_step_fwd(ec), l > 0 ? sk_step_back(ec) - 4 : sk_below_edge(g, ec),
                          l < 63 ? sk_step_fwd(ec) + 4 : sk_above_edge(g, ec)};
            so[u] = s_old[ec];
            zc[u] = z[ec];
            ad[u] = Adiag[ec];
            f[u] = flags[ec];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const long m = nb[k] < 0 ? 0 : (nb[k] > last ? last : nb[k]);
                ns[u][k] = s_old[m];
                nz[u][k] = z[m];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long e = e0 + u * it.stride;
            if (e < it.end) {
                const float sc_ = fmaf(beta, so[u], zc[u]);
                s_new[e] = sc_;
                float wv = 0.0f;
                if (ad[u] != 0.0f) {
                    float nsum = 0.0f;
                    if (f[u] & CPL_L) nsum += fmaf(beta, ns[u][0], nz[u][0]);
                    if (f[u] & CPL_R) nsum += fmaf(beta, ns[u][1], nz[u][1]);
                    if (f[u] & CPL_D) nsum += fmaf(beta, ns[u][2], nz[u][2]);
                    if (f[u] & CPL_U) nsum += fmaf(beta, ns[u][3], nz[u][3]);
                    wv = fmaf(ad[u], sc_, -nsum);
                }
                w[e] = wv;
                acc += (double)wv * (double)sc_;
            }
        }
    }
    double bs = block_sum(acc);
    if (threadIdx.x == 0) partials[vb] = bs;
}
// s' = z + beta s (:459, FMA as Array2D.h:185), w = A s' (:433-444), partial
// w.s' (:450), over nvb virtual workgroups: workgroup k of the launch runs
// vb = k, k + gridDim.x, ...  Each vb computes what a launch of nvb
// workgroups computes in its workgroup vb, so the partials, and every result,
// do not depend on the launch's size.  sigma goes to the scalars (vb 0).
__global__ __launch_bounds__(kBlock) void pcg_spmv_sk_kernel(
    SkGeom g, const float* __restrict__ Adiag, const uint8_t* __restrict__ flags, const float* __restrict__ z,
    const float* __restrict__ s_old, float* __restrict__ s_new, float* __restrict__ w,
    PcgScalars* sc, const double* sig_cur, const double* sig_prev, int first, double* partials, int nvb) {
    if (sc->done) return;
    int vb = blockIdx.x;
    if (vb >= nvb) return;
    SpmvPre pre = spmv_pre(g, Adiag, flags, z, s_old, vb, nvb);
    const SpmvScalars S = spmv_scalars(g, sig_cur, sig_prev, first);
    if (vb == 0 && threadIdx.x == 0) sc->sigma = S.sig;
    for (;;) {
        spmv_core(pre, g, Adiag, flags, z, s_old, s_new, w, S.beta, partials, vb);
        vb += gridDim.x;
        if (vb >= nvb) break;
        pre = spmv_pre(g, Adiag, flags, z, s_old, vb, nvb);
    }
}

// alpha = sigma / (w.s') (:450); p += alpha s', r -= alpha w (:451-452);
// partial |r|_inf for the stopping test (:453), finished by the forward solve.
// One float4 of each array per thread and round (the SK4 arrays hold a
// multiple of 256 entries).  Rounds go in chunks of kAxpyU; a virtual
// workgroup's first chunk is loaded (AxpyBuf::load) before alpha is known
// when it is the launch workgroup's first.
constexpr int kAxpyU = 2;
struct AxpyBuf {
    float4 sv[kAxpyU], wv[kAxpyU], pv[kAxpyU], rv[kAxpyU];
    __device__ __forceinline__ void load(const XcdLoop& it, long e0, const float4* p4, const float4* r4,
                                         const float4* s4, const float4* w4) {
#pragma unroll
        for (int u = 0; u < kAxpyU; u++) {
            const long e = e0 + u * it.stride;
            if (e < it.end) {
                sv[u] = s4[e];
                wv[u] = w4[e];
                pv[u] = p4[e];
                rv[u] = r4[e];
            }
        }
    }
};
__device__ __forceinline__ void axpy_core(AxpyBuf& B, const XcdLoop& it, float4* p4, float4* r4, const float4* s4,
                                          const float4* w4, float a, double* partials, int vb) {
    float m = 0.0f;
    for (long e0 = it.e; e0 < it.end;) {
#pragma unroll
        for (int u = 0; u < kAxpyU; u++) {
            const long e = e0 + u * it.stride;
            if (e < it.end) {
                float4 pn = B.pv[u], rn = B.rv[u];
                pn.x = fmaf(a, B.sv[u].x, pn.x);
                pn.y = fmaf(a, B.sv[u].y, pn.y);
                pn.z = fmaf(a, B.sv[u].z, pn.z);
                pn.w = fmaf(a, B.sv[u].w, pn.w);
                rn.x = fmaf(-a, B.wv[u].x, rn.x);
                rn.y = fmaf(-a, B.wv[u].y, rn.y);
                rn.z = fmaf(-a, B.wv[u].z, rn.z);
                rn.w = fmaf(-a, B.wv[u].w, rn.w);
                p4[e] = pn;
                r4[e] = rn;
                m = fmaxf(m, fmaxf(fmaxf(fabsf(rn.x), fabsf(rn.y)), fmaxf(fabsf(rn.z), fabsf(rn.w))));
            }
        }
        e0 += kAxpyU * it.stride;
        if (e0 < it.end) B.load(it, e0, p4, r4, s4, w4);
    }
    double bm = block_max((double)m);
    if (threadIdx.x == 0) partials[vb] = bm;
}
// The axpy over nvb virtual workgroups (as the SpMV).  n_ws >= 0: sigma and
// w.s are reduced from the partials here; n_ws < 0: they were reduced into sc
// already, in the same order.
__global__ __launch_bounds__(kBlock) void pcg_axpy_sk_kernel(
    SkGeom g, float* __restrict__ p, float* __restrict__ r, const float* __restrict__ s,
    const float* __restrict__ w, const PcgScalars* sc, const double* sig_cur, const double* ws_partials,
    int n_ws, double* partials, int nvb) {
    if (sc->done) return;
    int vb = blockIdx.x;
    if (vb >= nvb) return;
    float4* p4 = (float4*)p;
    float4* r4 = (float4*)r;
    const float4* s4 = (const float4*)s;
    const float4* w4 = (const float4*)w;
    AxpyBuf B;
    XcdLoop it = xcd_loop(g.size() / 4, vb, nvb);
    B.load(it, it.e, p4, r4, s4, w4);
    const double sig = n_ws < 0 ? sc->sigma : block_sum_bcast(sig_cur, g.S * kLoaders);
    const double ws = n_ws < 0 ? sc->ws : block_sum_bcast(ws_partials, n_ws);
    const float a = (float)(sig / ws);
    for (;;) {
        axpy_core(B, it, p4, r4, s4, w4, a, partials, vb);
        vb += gridDim.x;
        if (vb >= nvb) break;
        it = xcd_loop(g.size() / 4, vb, nvb);
        B.load(it, it.e, p4, r4, s4, w4);
    }
}

// ---------------------------------------------------------------------------
// The backward solve and the next iteration's SpMV in one launch
// (src/FluidSim2D.cpp:410-421 then :457-459, :433-444, :450).  The backward
// strips finish top-down about 3 us apart and the last one runs long after
// most of the chip has gone idle; the SpMV of a region needs only z of its
// own and the neighbouring strips, so it runs beneath the solve's tail.
// With `check`, the launch applies the stopping test of the iteration whose
// p/r update ran in the previous launch (:453): the sigma and SpMV workgroups
// evaluate it from the same partials, so they agree whether there is work;
// the strips run regardless (see bwd_spmv_kernel).
// Workgroups take tickets of the launch (tickets / launch = K + S + M):
//   0 .. K-1        sigma: q.q per (strip, loader) in the order the solve's
//                   loaders would sum it (bwd_sigma), published as tagged
//                   words {half, epoch} (no fences) and as plain doubles for
//                   later launches
//   K .. K+S-1      the backward solve's strips (BwdFusedOp): z is stored
//                   with device scope (written through to memory), and after
//                   those stores have completed each strip flags itself
//   K+S .. K+S+M-1  SpMV: sigma from the tagged words, beta as spmv_scalars
//                   forms it, then the virtual workgroups of pcg_spmv_sk_kernel
//                   (the same entries, arithmetic and partials) in the order
//                   their strips finish, each after its strips' flags
// Every wait is on a workgroup with a smaller ticket, which started earlier,
// so the launch completes at any residency; every wait is bounded and raises
// the error flag on expiry.  Results are bit-identical to the separate
// launches.  Kernel starts invalidate the caches, and no workgroup reads a
// strip's z before its flag (a strip's entries are whole 1 KB groups, so no
// cache line holds two strips), so none can hold a stale copy.
// ---------------------------------------------------------------------------
struct FusedSpmvArgs {
    const float* Adiag;
    const uint8_t* flags;
    const float* z;
    const float* q;
    const float* s_old;
    float* s_new;
    float* w;
    PcgScalars* sc;
    double* sig_out;                // sigma_{k+1} partials [S][kLoaders], plain
    unsigned long long* sig_words;  // the same as {low half, epoch}, {high half, epoch}
    const double* sig_prev;         // sigma_k partials (an earlier launch's), plain
    int first;                      // no beta (the initial solve)
    double* ws_partials;            // [nvb]
    int nvb;                        // SpMV virtual workgroups (pcg_blocks)
    const unsigned* strip_flags;
    int K, M;
    int check;                      // the stopping test of the iteration (its p/r update's partials)
    const double* rmax_partials;     // [nvb]
};

// sigma workgroup k of K: one (strip, loader) pair per wave; the lane sums its
// q.q over the loader's supergroups (sg = me, me + kLoaders, ... in the
// backward solve's order), then the wave's tree, as BwdOp's loaders do.
__device__ __forceinline__ void bwd_sigma(const SkGeom& g, const FusedSpmvArgs& f, int k, unsigned epoch) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int kWaves = kSweepThreads / 64;
    const int NSG = g.T4 / kSub, T = g.T4 * 4;
    for (int pr = k * kWaves + wave; pr < g.S * kLoaders; pr += f.K * kWaves) {
        const int s = pr / kLoaders, me = pr % kLoaders;
        const float* qs = f.q + (long)s * g.T4 * 256 + lane * 4;
        double acc = 0;
        constexpr int U = 4;  // supergroups with loads in flight together
        for (int sg0 = me; sg0 < NSG; sg0 += U * kLoaders) {
            float4 v[U][kSub];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int sg = sg0 + u * kLoaders;
                const int t0 = T - 16 * ((sg < NSG ? sg : NSG - 1) + 1);
#pragma unroll
                for (int j = 0; j < kSub; j++) v[u][j] = *(const float4*)(qs + (long)(t0 + 4 * j) * 64);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (sg0 + u * kLoaders < NSG)
#pragma unroll
                    for (int j = 0; j < kSub; j++) acc += qq4(v[u][j]);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) {
            f.sig_out[pr] = acc;
            const unsigned long long b = (unsigned long long)__double_as_longlong(acc);
            const unsigned long long tag = (unsigned long long)epoch << 32;
            atomic_st64(f.sig_words + 2 * pr, (b & 0xffffffffull) | tag);
            atomic_st64(f.sig_words + 2 * pr + 1, (b >> 32) | tag);
        }
    }
}

// SpMV workgroup m of M.
__device__ __forceinline__ void bwd_spmv_worker(const SkGeom& g, const FusedSpmvArgs& f, int m, unsigned epoch,
                                                int* err) {
    __shared__ int sh_bad;
    __shared__ double sh_sig;
    const int lane = threadIdx.x & 63;
    const int n_sig = g.S * kLoaders;
    // sigma_{k+1}: wave 0 waits for this launch's tagged words and sums them in
    // wave_sum_of's order
    if (threadIdx.x < 64) {
        double t = 0;
        bool bad = false;
        for (int spin = 0;; spin++) {
            t = 0;
            bool ok = true;
            for (int i = lane; i < n_sig; i += 64) {
                const unsigned long long lo = atomic_ld64(f.sig_words + 2 * i);
                const unsigned long long hi = atomic_ld64(f.sig_words + 2 * i + 1);
                ok = ok && (unsigned)(lo >> 32) == epoch && (unsigned)(hi >> 32) == epoch;
                t += __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
            }
            if (__all(ok)) break;
            if (spin >= kSpinLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                bad = true;
                if (lane == 0) atomicExch(err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) {
            sh_bad = bad;
            sh_sig = t;
        }
    }
    __syncthreads();
    if (sh_bad) return;
    const double sig = sh_sig;
    double beta_d = 0;
    if (!f.first) beta_d = sig / block_sum_bcast(f.sig_prev, n_sig);
    const float beta = (float)beta_d;
    if (m == 0 && threadIdx.x == 0) f.sc->sigma = sig;

    // Virtual workgroups in the order their strips finish: the top of the
    // array (the last runs) first.  A run reads z of its strips and of one
    // strip on either side; the strips above it finish earlier.
    const long n = g.size(), strip_sz = (long)g.T4 * 256;
    int lo_ready = g.S;  // strips lo_ready .. S-1 are known to be done
    for (int it = m; it < f.nvb; it += f.M) {
        const int vb = f.nvb - 1 - it;
        const XcdLoop xl = spmv_range(n, vb, f.nvb, (int)threadIdx.x);
        const long b0 = (n >> 8) * vb / f.nvb * 256, b1 = (n >> 8) * (vb + 1) / f.nvb * 256;
        const int need = b0 < b1 ? max(0, (int)(b0 / strip_sz) - 1) : g.S;
        if (need < lo_ready) {
            if (threadIdx.x < 64) {
                bool bad = false;
                // strips finish top-down: one lane watches the lowest one
                // needed (few polls through the fabric while the strips hand
                // off), then the wave checks them all
                for (int spin = 0;; spin++) {
                    unsigned v = 0;
                    if (lane == 0)
                        v = __hip_atomic_load(f.strip_flags + need, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((unsigned)__shfl((int)v, 0, 64) == epoch) break;
                    if (spin >= kSpinLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        bad = true;
                        if (lane == 0) atomicExch(err, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(16);
                }
                for (int s0 = need; s0 < lo_ready && !bad; s0 += 64) {
                    const int s = s0 + lane;
                    for (int spin = 0;; spin++) {
                        const bool ok = s >= lo_ready ||
                                        __hip_atomic_load(f.strip_flags + s, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT) == epoch;
                        if (__all(ok)) break;
                        if (spin >= kSpinLimit ||
                            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                            bad = true;
                            if (lane == 0) atomicExch(err, 1);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(8);
                    }
                }
                if (lane == 0) sh_bad = bad;
            }
            __syncthreads();
            if (sh_bad) return;
            lo_ready = need;
        }
        // 9 entries at once: a run is n / 256 / nvb groups, about 8.4 at 2048^2
        // (pcg_blocks aims at 8 entries per thread), so one batch takes it whole
        // (8 at once: 1.2 us more per launch; 12: slower, register pressure)
        spmv_core_wide<9>(xl, g, f.Adiag, f.flags, f.z, f.s_old, f.s_new, f.w, beta, f.ws_partials, vb);
    }
}

__global__ __launch_bounds__(kSweepThreads) void bwd_spmv_kernel(BwdFusedOp op, SkGeom g, SkCtl c,
                                                                 FusedSpmvArgs f) {
    __shared__ __align__(16) SweepLds<BwdFusedOp, sweep_ring<BwdFusedOp>()> L;
    const unsigned long long ticket = sweep_ticket(L, c.ticket);
    const unsigned long long per = (unsigned long long)(f.K + g.S + f.M);
    const int idx = (int)(ticket % per);
    const unsigned epoch = (unsigned)(ticket / per) + 1u;
    // sc->done is set before the launch, or by this launch's stopping test
    // (:453), which the sigma and SpMV workgroups evaluate for themselves from
    // the same partials (so they agree whether there is work) and SpMV
    // workgroup 0 records (|r|_inf; done, or the iteration count).  The
    // strips do not wait for it: when it passes, their solve is wasted but
    // harmless (nothing reads z or sigma once done is set).
    const bool strip = idx >= f.K && idx < f.K + g.S;
    __shared__ int sh_done;
    if (threadIdx.x < 64) {
        int done = f.sc->done;
        if (!done && f.check && !strip) {
            const double rn = f.nvb <= 64 * 32 ? wave_max_of_wide<32>(f.rmax_partials, f.nvb)
                                               : wave_max_of(f.rmax_partials, f.nvb);
            done = rn <= f.sc->tol * f.sc->rhs_norm;
            if (idx == f.K + g.S && threadIdx.x == 0) {  // SpMV workgroup 0
                f.sc->r_norm = rn;
                if (done) f.sc->done = 1;
                else f.sc->iter = op.next_iter;
            }
        }
        if (threadIdx.x == 0) sh_done = done;
    }
    __syncthreads();
    if (sh_done) return;
    if (idx < f.K) {
        bwd_sigma(g, f, idx, epoch);
    } else if (idx < f.K + g.S) {
        op.epoch = epoch;
        sweep_run<BwdFusedOp, true, true>(op, g, c, idx - f.K, epoch, L);
    } else {
        bwd_spmv_worker(g, f, idx - f.K - g.S, epoch, c.err);
    }
}

// ---------------------------------------------------------------------------
// The p/r update and the forward solve in one launch (src/FluidSim2D.cpp:450-452
// then :400-409).  Forward strip s needs only r of strip s, so the update runs
// over strip-ordered runs of whole 256-entry groups (pcg_run) and flags each
// one; the solve's loaders start on a strip as soon as its runs are flagged.
// Workgroups take tickets of the launch (tickets / launch = M + S):
//   0 .. M-1      the update: alpha as pcg_axpy_sk_kernel forms it, then runs
//                 m, m + M, ...; r is stored with device scope, so after the
//                 stores' completion (vmcnt 0, every wave, then a barrier) the
//                 run's flag may follow; the |r|_inf partial of each run
//   M .. M+S-1    the forward solve's strips (FwdFusedOp)
// The stopping test, which needs every run's partial, moves to the backward
// solve's launch.  Each entry's update and the maximum over all partials are
// those of the separate launches, so every result is bit-identical.
// ---------------------------------------------------------------------------
struct FusedAxpyArgs {
    float* p;
    float* r;
    const float* s;                // s' of the iteration
    const float* w;
    const PcgScalars* sc;
    const double* sig_cur;         // sigma partials [S][kLoaders]
    const double* ws_partials;     // [nvb]
    double* rmax_partials;         // [nvb], one per run
    unsigned* run_flags;           // [nvb]
    int nvb, M;
};
// Float4 range of p/r update run vb of nvb: whole 256-entry groups (64
// float4s each), the n / 256 groups split evenly.
__device__ __forceinline__ void pcg_run(long n, int vb, int nvb, long& q0, long& q1) {
    const long ng = n >> 8;
    q0 = ng * vb / nvb * 64;
    q1 = ng * (vb + 1) / nvb * 64;
}
// One run of the p/r update by one wave: p += a s, r -= a w over float4s
// q0 .. q1-1 (r stored with device scope), its |r|_inf partial, then, once
// the wave's stores have completed, the run's flag.  The partial is a maximum,
// so it is the same whichever threads cover the run.  For runs of a few
// float4s per lane (the ensemble's members), where one wave per run keeps
// six runs per workgroup in flight.
constexpr int kRunU = 4;
struct RunBuf {
    float4 sv[kRunU], wv[kRunU], pv[kRunU], rv[kRunU];
    __device__ __forceinline__ void load(const float4* s4, const float4* w4, const float4* p4, const float4* r4,
                                         long e0, long q1) {
#pragma unroll
        for (int u = 0; u < kRunU; u++) {
            const long e = e0 + u * 64;
            if (e < q1) {
                sv[u] = s4[e];
                wv[u] = w4[e];
                pv[u] = p4[e];
                rv[u] = r4[e];
            }
        }
    }
};
__device__ __forceinline__ void axpy_run_wave(RunBuf& B, const float4* s4, const float4* w4, float4* p4,
                                              const float4* r4, __amdgpu_buffer_rsrc_t r_rs, float a, long q0,
                                              long q1, double* partial, unsigned* flag, unsigned epoch) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    float mx = 0.0f;
    for (long e0 = q0 + lane; e0 < q1 + lane; e0 += kRunU * 64) {
        B.load(s4, w4, p4, r4, e0, q1);
#pragma unroll
        for (int u = 0; u < kRunU; u++) {
            const long e = e0 + u * 64;
            if (e < q1) {
                float4 pn = B.pv[u], rn = B.rv[u];
                pn.x = fmaf(a, B.sv[u].x, pn.x);
                pn.y = fmaf(a, B.sv[u].y, pn.y);
                pn.z = fmaf(a, B.sv[u].z, pn.z);
                pn.w = fmaf(a, B.sv[u].w, pn.w);
                rn.x = fmaf(-a, B.wv[u].x, rn.x);
                rn.y = fmaf(-a, B.wv[u].y, rn.y);
                rn.z = fmaf(-a, B.wv[u].z, rn.z);
                rn.w = fmaf(-a, B.wv[u].w, rn.w);
                p4[e] = pn;
                const f4 rv4 = {rn.x, rn.y, rn.z, rn.w};
                __builtin_amdgcn_raw_buffer_store_b128(rv4, r_rs, (int)(e * 16), 0, kCpolSc1);
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf(rn.x), fabsf(rn.y)), fmaxf(fabsf(rn.z), fabsf(rn.w))));
            }
        }
    }
    const double m = wave_max((double)mx);
    if (lane == 0) *partial = m;
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(kWaitVm0);  // the wave's r stores are in memory
    asm volatile("" ::: "memory");
    if (lane == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Update workgroup m of M: runs m, m + M, ... in order (strip 0's first), all
// 256 threads of its first four waves on one run at a time: a run of a large
// grid is about 540 float4s, which one wave would take in three round trips
// (per-wave runs measured 157.2 against 150.9 us per launch at 2048^2).
__device__ __forceinline__ void fwd_axpy_worker(const SkGeom& g, const FusedAxpyArgs& f, int m, unsigned epoch) {
    float4* p4 = (float4*)f.p;
    const float4* s4 = (const float4*)f.s;
    const float4* w4 = (const float4*)f.w;
    const float4* r4 = (const float4*)f.r;
    constexpr int U = 4;
    float4 sv[U], wv[U], pv[U], rv[U];
    auto load = [&](long e0, long q1) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long e = e0 + u * kBlock;
            if (e < q1) {
                sv[u] = s4[e];
                wv[u] = w4[e];
                pv[u] = p4[e];
                rv[u] = r4[e];
            }
        }
    };
    // the first run's first entries do not depend on alpha: their loads go
    // out before its reduction
    long q0, q1;
    pcg_run(g.size(), m, f.nvb, q0, q1);
    if (threadIdx.x < kBlock) load(q0 + threadIdx.x, q1);
    // alpha = sigma / w.s as pcg_axpy_sk_kernel's block_sum_bcast forms it,
    // with every load of both sums in flight together (one memory round
    // trip: the update's first runs, and so the first strip, wait on it)
    __shared__ float sh_a;
    if (threadIdx.x < 64) {
        double sig, ws;
        if (g.S * kLoaders <= 64 * 12 && f.nvb <= 64 * 32) {
            wave_sum_of_wide2<12, 32>(f.sig_cur, g.S * kLoaders, f.ws_partials, f.nvb, sig, ws);
        } else {
            sig = wave_sum_of(f.sig_cur, g.S * kLoaders);
            ws = wave_sum_of(f.ws_partials, f.nvb);
        }
        if (threadIdx.x == 0) sh_a = (float)(sig / ws);
    }
    __syncthreads();
    const float a = sh_a;
    const __amdgpu_buffer_rsrc_t r_rs = __builtin_amdgcn_make_buffer_rsrc(f.r, 0, (int)(g.size() * 4), kBufferCfg);
    typedef float f4 __attribute__((ext_vector_type(4)));
    for (int vb = m; vb < f.nvb; vb += f.M) {
        pcg_run(g.size(), vb, f.nvb, q0, q1);
        float mx = 0.0f;
        if (threadIdx.x < kBlock) {
            for (long e0 = q0 + threadIdx.x; e0 < q1; e0 += U * kBlock) {
                if (vb != m || e0 != q0 + threadIdx.x) load(e0, q1);
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const long e = e0 + u * kBlock;
                    if (e < q1) {
                        float4 pn = pv[u], rn = rv[u];
                        pn.x = fmaf(a, sv[u].x, pn.x);
                        pn.y = fmaf(a, sv[u].y, pn.y);
                        pn.z = fmaf(a, sv[u].z, pn.z);
                        pn.w = fmaf(a, sv[u].w, pn.w);
                        rn.x = fmaf(-a, wv[u].x, rn.x);
                        rn.y = fmaf(-a, wv[u].y, rn.y);
                        rn.z = fmaf(-a, wv[u].z, rn.z);
                        rn.w = fmaf(-a, wv[u].w, rn.w);
                        p4[e] = pn;
                        const f4 rv4 = {rn.x, rn.y, rn.z, rn.w};
                        __builtin_amdgcn_raw_buffer_store_b128(rv4, r_rs, (int)(e * 16), 0, kCpolSc1);
                        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(rn.x), fabsf(rn.y)), fmaxf(fabsf(rn.z), fabsf(rn.w))));
                    }
                }
            }
        }
        const double bm = block_max((double)mx);
        if (threadIdx.x == 0) f.rmax_partials[vb] = bm;
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(kWaitVm0);  // this wave's r stores are in memory
        asm volatile("" ::: "memory");
        __syncthreads();                        // ... and every wave's
        if (threadIdx.x == 0) __hip_atomic_store(f.run_flags + vb, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(kSweepThreads) void fwd_axpy_kernel(FwdFusedOp op, SkGeom g, SkCtl c,
                                                                 FusedAxpyArgs f) {
    __shared__ __align__(16) SweepLds<FwdFusedOp, sweep_ring<FwdFusedOp>()> L;
    const unsigned long long ticket = sweep_ticket(L, c.ticket);
    const unsigned long long per = (unsigned long long)(f.M + g.S);
    const int idx = (int)(ticket % per);
    const unsigned epoch = (unsigned)(ticket / per) + 1u;
    // sc->done is set only between launches of this kind
    __shared__ int sh_done;
    if (threadIdx.x == 0) sh_done = f.sc->done;
    __syncthreads();
    if (sh_done) return;
    if (idx < f.M) {
        fwd_axpy_worker(g, f, idx, epoch);
    } else {
        op.epoch = epoch;
        sweep_run<FwdFusedOp, false, true>(op, g, c, idx - f.M, epoch, L);
    }
}

// PcgScalars of one projection: zero, tolerance set (:424-430).
__global__ void pcg_init_kernel(PcgScalars* sc, double tol) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PcgScalars z{};
        z.tol = tol;
        *sc = z;
    }
}

// SK4 -> row-major copy of the pressure.
__device__ __forceinline__ void sk_to_rowmajor_body(SkGeom g, const float* __restrict__ sk,
                                                               float* __restrict__ rm, int vb, int nvb) {
    const long n = (long)g.nx * g.ny;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % g.nx), j = (int)(c / g.nx);
        rm[c] = sk[g.idx(i, j)];
    }
}
__global__ __launch_bounds__(kBlock) void sk_to_rowmajor_kernel(SkGeom g, const float* __restrict__ sk,
                                                               float* __restrict__ rm) {
    sk_to_rowmajor_body(g, sk, rm, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// FluidSim2D::updateVelocity (src/FluidSim2D.cpp:469-550): newMac from mac and
// the pressure gradient with ghost-fluid terms; faces not next to FLUID are
// marked unknown (mask ~0) for extrapolation.  Faces u(nx, j) and v(i, ny)
// lie outside the reference's loops and keep mac's value (mask 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void update_velocity_body(
    const uint8_t* __restrict__ cell, const double* __restrict__ phi, const float* __restrict__ p,
    const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ nu, float* __restrict__ nv,
    uint32_t* __restrict__ mu, uint32_t* __restrict__ mv, int nx, int ny, double scale, int ghost, int vb, int nvb) {
    const long nU = (long)(nx + 1) * ny, nV = (long)nx * (ny + 1);
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < nU + nV; e += (long)nvb * blockDim.x) {
        if (e < nU) {
            const int i = (int)(e % (nx + 1)), j = (int)(e / (nx + 1));
            double val = u[e];
            uint32_t mk = 0;
            if (i < nx) {
                const long c = (long)j * nx + i;
                const uint8_t cc = cell[c];
                const uint8_t cl = i > 0 ? cell[c - 1] : SOLID;
                if ((i > 0 && cl == FLUID) || cc == FLUID) {
                    if ((i == 0 || cl == SOLID) || cc == SOLID) val = 0;
                    else if (!ghost) val -= scale * ((double)p[c] - (double)p[c - 1]);  // :486-505 without ghosts
                    else if (i > 0 && cl == EMPTY)
                        val -= scale * (1 - fmax(phi[c - 1] / phi[c], -1e3)) * (double)p[c];
                    else if (cc == EMPTY)
                        val -= scale * (fmax(phi[c] / phi[c - 1], -1e3) - 1) * (double)p[c - 1];
                    else if (i > 0) val -= scale * ((double)p[c] - (double)p[c - 1]);
                    else val = 0;
                } else {
                    mk = 0xFFFFFFFFu;
                }
            }
            nu[e] = (float)val;
            mu[e] = mk;
        } else {
            const long f = e - nU;
            const int i = (int)(f % nx), j = (int)(f / nx);
            double val = v[f];
            uint32_t mk = 0;
            if (j < ny) {
                const long c = (long)j * nx + i;
                const uint8_t cc = cell[c];
                const uint8_t cb = j > 0 ? cell[c - nx] : SOLID;
                if ((j > 0 && cb == FLUID) || cc == FLUID) {
                    if ((j == 0 || cb == SOLID) || cc == SOLID) val = 0;
                    else if (!ghost) val -= scale * ((double)p[c] - (double)p[c - nx]);
                    else if (j > 0 && cb == EMPTY)
                        val -= scale * (1 - fmax(phi[c - nx] / phi[c], -1e3)) * (double)p[c];
                    else if (cc == EMPTY)
                        val -= scale * (fmax(phi[c] / phi[c - nx], -1e3) - 1) * (double)p[c - nx];
                    else if (j > 0) val -= scale * ((double)p[c] - (double)p[c - nx]);
                    else val = 0;
                } else {
                    mk = 0xFFFFFFFFu;
                }
            }
            nv[f] = (float)val;
            mv[f] = mk;
        }
    }
}
__global__ __launch_bounds__(kBlock) void update_velocity_kernel(
    const uint8_t* __restrict__ cell, const double* __restrict__ phi, const float* __restrict__ p,
    const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ nu, float* __restrict__ nv,
    uint32_t* __restrict__ mu, uint32_t* __restrict__ mv, int nx, int ny, double scale, int ghost) {
    update_velocity_body(cell, phi, p, u, v, nu, nv, mu, mv, nx, ny, scale, ghost, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Extrapolation (Array2D<double>::extrapolate, include/Array2D.h:552-591).
// The reference's FIFO BFS from the known entries (mask 0) is layer-
// synchronous: an entry of layer L takes the mean of its neighbours of layer
// < L, taken in the order -x, +x, -y, +y (:570-590), all final before layer L
// starts.  On a full rectangle without obstacles the BFS layer of an entry is
// its L1 distance to the known set, so the layers come from a separable
// distance transform instead of a frontier:
//   rows:    d_row(i) = distance to the nearest known entry of the row (a
//            wave-wide prefix max / suffix min of known indices);
//   columns: d(j) = min_j' d_row(j') + |j - j'| (a prefix min of d_row - j,
//            then a suffix min of d + j over the result, in chunks per wave);
//   sort:    entries counted per layer, offsets scanned, scattered into layer
//            order; each entry's sources (neighbours of lower layer) are then
//            linked: the known ones by value, the others by their position in
//            the previous layer;
//   values:  one workgroup per array runs the layers in order, with the
//            previous layer's values in LDS: a layer costs one LDS round trip
//            and a barrier, and the global stores never wait.
// Masks end as the layer numbers (0 known, ~0 where the array has no known
// entry), as the reference leaves them.
// ---------------------------------------------------------------------------
constexpr uint32_t kFar = 1u << 29;  // distance of an entry with no known entry in its array
constexpr int kExtraColThreads = 1024;

struct ExtraArr {
    gptr<float> val;
    gptr<uint32_t> mask;
    gptr<int> order;   // entries in layer order
    gptr<int> pos;     // position of each entry in that order
    gptr<i4v> src;     // per ordered entry and neighbour (-x, +x, -y, +y): -1 none, -2 known, else position
    gptr<f4v> kv;      // per ordered entry: the known neighbours' values
    gptr<int> aux;     // layer d: count at [32 d], cursor at [32 d + 1] (one 128-byte line per layer,
                       // so that the atomics of different layers do not share a line); offsets after
    gptr<int> count;   // number of layers
    int w, h;
};
constexpr int kExtraLds = 4096;  // entries of a layer whose values stay in LDS
__host__ __device__ inline int extra_layers(int w, int h) { return w + h + 2; }
__host__ __device__ inline int extra_aux_len(int w, int h) { return 33 * extra_layers(w, h); }
__device__ __forceinline__ gptr<int> extra_off(const ExtraArr& X) { return X.aux + 32 * extra_layers(X.w, X.h); }

// Row pass; also clears the per-layer counts.
__device__ __forceinline__ void extrap_rows_body(ExtraArr A, ExtraArr B, int vb, int nvb) {
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    const int gt = vb * blockDim.x + threadIdx.x, nt = nvb * blockDim.x;
    for (int k = gt; k < extra_layers(A.w, A.h); k += nt) A.aux[32 * k] = 0;
    for (int k = gt; k < extra_layers(B.w, B.h); k += nt) B.aux[32 * k] = 0;
    const long nrows = (long)A.h + B.h;
    for (long r = (long)vb * wpb + (threadIdx.x >> 6); r < nrows; r += (long)nvb * wpb) {
        const ExtraArr& X = r < A.h ? A : B;
        const int j = (int)(r < A.h ? r : r - A.h);
        const gptr<uint32_t> m = X.mask + (long)j * X.w;
        // forward: index of the last known entry at or before i
        int carry = -(int)kFar;
        for (int i0 = 0; i0 < X.w; i0 += 64) {
            const int i = i0 + lane;
            int c = (i < X.w && m[i] == 0) ? i : -(int)kFar;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(c, o, 64);
                if (lane >= o) c = c > t ? c : t;
            }
            c = c > carry ? c : carry;
            carry = __shfl(c, 63, 64);
            if (i < X.w) m[i] = c < 0 ? kFar : (uint32_t)(i - c);  // known entries: 0
        }
        // backward: index of the next known entry at or after i
        carry = 2 * (int)kFar;
        for (int i0 = (X.w - 1) & ~63; i0 >= 0; i0 -= 64) {
            const int i = i0 + lane;
            const uint32_t df = i < X.w ? m[i] : kFar;
            int c = (i < X.w && df == 0) ? i : 2 * (int)kFar;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_down(c, o, 64);
                if (lane + o < 64) c = c < t ? c : t;
            }
            c = c < carry ? c : carry;
            carry = __shfl(c, 0, 64);
            if (i < X.w) {
                const uint32_t db = c >= (int)kFar ? kFar : (uint32_t)(c - i);
                m[i] = df < db ? df : db;
            }
        }
    }
}

// Column pass: a workgroup per 64 columns; each wave takes one contiguous
// chunk of rows.  Forward F(j) = j + min_{j' <= j} (d_row(j') - j'), then
// backward d(j) = min_{j' >= j} (F(j') + j') - j, each as chunk totals, an
// exclusive combination across waves in LDS, and a rescan.
__device__ __forceinline__ void extrap_cols_body(ExtraArr A, ExtraArr B, int vb, int nvb) {
    constexpr int NW = kExtraColThreads / 64;
    __shared__ int totl[NW][64];  // per wave and column: the chunk's total
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nbA = (A.w + 63) / 64, nbB = (B.w + 63) / 64;
    for (int blk = vb; blk < nbA + nbB; blk += nvb) {
        const ExtraArr& X = blk < nbA ? A : B;
        const int i = (blk < nbA ? blk : blk - nbA) * 64 + lane;
        const bool col = i < X.w;
        const int chunk = (X.h + NW - 1) / NW;
        const int j0 = wave * chunk, j1 = j0 + chunk < X.h ? j0 + chunk : X.h;
        const gptr<uint32_t> m = X.mask + (col ? i : 0);
        const long W = X.w;
        // forward totals
        int t = (int)kFar;
        if (col)
            for (int j = j0; j < j1; j++) {
                const int a = (int)m[j * W] - j;
                t = a < t ? a : t;
            }
        __syncthreads();  // the previous block's readers of totl are done
        totl[wave][lane] = t;
        __syncthreads();
        int run = (int)kFar;
        for (int v = 0; v < wave; v++) run = totl[v][lane] < run ? totl[v][lane] : run;
        if (col)
            for (int j = j0; j < j1; j++) {
                const int a = (int)m[j * W] - j;
                run = a < run ? a : run;
                const int F = j + run;
                m[j * W] = (uint32_t)(F < (int)kFar ? F : (int)kFar);
            }
        __syncthreads();
        // backward totals over F
        t = 3 * (int)kFar;
        if (col)
            for (int j = j1 - 1; j >= j0; j--) {
                const int b = (int)m[j * W] + j;
                t = b < t ? b : t;
            }
        totl[wave][lane] = t;
        __syncthreads();
        run = 3 * (int)kFar;
        for (int v = NW - 1; v > wave; v--) run = totl[v][lane] < run ? totl[v][lane] : run;
        if (col)
            for (int j = j1 - 1; j >= j0; j--) {
                const int b = (int)m[j * W] + j;
                run = b < run ? b : run;
                const int d = run - j;
                m[j * W] = (uint32_t)(d < (int)kFar ? d : (int)kFar);
            }
        __syncthreads();
    }
}

// Entries per layer, counted per 16 x 16 tile of faces: a tile's distances
// span at most 31 layers, so the workgroup adds its LDS counts with about 30
// global atomics instead of 256.  tile_pass(..., SCATTER) reserves each
// layer's range for the tile and writes the entries' places in layer order.
template <bool SCATTER>
__device__ __forceinline__ void extrap_tile_pass(ExtraArr A, ExtraArr B, int vb, int nvb) {
    __shared__ int cnt[32], base[32];
    __shared__ uint32_t dmin_s;
    const int tid = threadIdx.x, lx = tid & 15, ly = tid >> 4;
    const int txA = (A.w + 15) >> 4, tyA = (A.h + 15) >> 4, txB = (B.w + 15) >> 4, tyB = (B.h + 15) >> 4;
    const int ntA = txA * tyA, nt = ntA + txB * tyB;
    for (int t = vb; t < nt; t += nvb) {
        const bool inA = t < ntA;
        const ExtraArr& X = inA ? A : B;
        const int tt = inA ? t : t - ntA, tx = inA ? txA : txB;
        const int i = (tt % tx) * 16 + lx, j = (tt / tx) * 16 + ly;
        const bool in = i < X.w && j < X.h;
        const long f = (long)j * X.w + i;
        const uint32_t d = in ? X.mask[f] : 0u;
        const bool valid = d >= 1 && d < kFar;
        if (tid < 32) cnt[tid] = 0;
        if (tid == 0) dmin_s = kFar;
        __syncthreads();
        if (valid) atomicMin(&dmin_s, d);
        __syncthreads();
        const uint32_t dm = dmin_s;
        const int bin = (int)(d - dm);  // < 31 for a valid entry
        const int rank = valid ? atomicAdd(&cnt[bin], 1) : 0;
        __syncthreads();
        if (tid < 32 && cnt[tid] > 0) {
            const gptr<int> slot = X.aux + 32 * (dm + tid) + (SCATTER ? 1 : 0);
            const int b = __hip_atomic_fetch_add(slot, cnt[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (SCATTER) base[tid] = b;
        }
        if (SCATTER) {
            __syncthreads();
            if (valid) {
                const int k = base[bin] + rank;
                X.order[k] = (int)f;
                X.pos[f] = k;
            } else if (in && d >= kFar) {
                X.mask[f] = 0xFFFFFFFFu;  // no known entry reaches it
            }
        }
        __syncthreads();  // cnt and base are reused by the next tile
    }
}
__device__ __forceinline__ void extrap_hist_body(ExtraArr A, ExtraArr B, int vb, int nvb) {
    extrap_tile_pass<false>(A, B, vb, nvb);
}

// Exclusive offsets of the layers (workgroup 0: A, 1: B), the cursors, and
// the number of layers.
__device__ __forceinline__ void extrap_scan_body(ExtraArr A, ExtraArr B, int vb) {
    ExtraArr X = vb == 0 ? A : B;
    const int n = extra_layers(X.w, X.h);
    const gptr<int> off = extra_off(X);
    __shared__ int part[kBlock];
    __shared__ int last[kBlock];
    const int per = (n + kBlock - 1) / kBlock;
    const int k0 = threadIdx.x * per, k1 = k0 + per < n ? k0 + per : n;
    int sum = 0, hi = 0;
    for (int k = k0; k < k1; k++) {
        const int c = X.aux[32 * k];
        sum += c;
        if (c > 0) hi = k;
    }
    part[threadIdx.x] = sum;
    last[threadIdx.x] = hi;
    __syncthreads();
    if (threadIdx.x < 64) {  // one wave: 4 partials per lane, then a lane scan
        const int l = threadIdx.x;
        int v[4], tsum = 0, top = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = part[4 * l + q];
            tsum += v[q];
            top = last[4 * l + q] > top ? last[4 * l + q] : top;
        }
        int inc = tsum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(inc, o, 64);
            if (l >= o) inc += t;
            const int m = __shfl_xor(top, o, 64);
            top = m > top ? m : top;
        }
        int run = inc - tsum;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            part[4 * l + q] = run;
            run += v[q];
        }
        if (l == 0) *X.count = top;
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int k = k0; k < k1; k++) {
        const int c = X.aux[32 * k];
        off[k] = run;
        X.aux[32 * k + 1] = run;
        run += c;
    }
}

// Entries in layer order.  Entries no known entry reaches keep mask ~0.
__device__ __forceinline__ void extrap_scatter_body(ExtraArr A, ExtraArr B, int vb, int nvb) {
    extrap_tile_pass<true>(A, B, vb, nvb);
}

// Sources of each ordered entry: the neighbours of lower layer, in the order
// -x, +x, -y, +y (:575-588); known ones carry their value (constant while the
// layers run), the others their position in the order.
__device__ __forceinline__ void extrap_link_body(ExtraArr A, ExtraArr B, int vb, int nvb) {
    const long nA = (long)A.w * A.h, nB = (long)B.w * B.h;
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < nA + nB; e += (long)nvb * blockDim.x) {
        const ExtraArr& X = e < nA ? A : B;
        const long k = e < nA ? e : e - nA;
        const int n = extra_off(X)[extra_layers(X.w, X.h) - 1];  // the last offset: the number of ordered entries
        if (k >= n) continue;
        const int f = X.order[k];
        const uint32_t d = X.mask[f];
        const int i = f % X.w, j = f / X.w;
        const int nb[4] = {i > 0 ? f - 1 : -1, i < X.w - 1 ? f + 1 : -1, j > 0 ? f - X.w : -1,
                           j < X.h - 1 ? f + X.w : -1};
        int c[4];
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            c[q] = -1;
            v[q] = 0.0f;
            if (nb[q] < 0) continue;
            const uint32_t dq = X.mask[nb[q]];
            if (dq == 0) {
                c[q] = -2;
                v[q] = X.val[nb[q]];
            } else if (dq < d) {
                c[q] = X.pos[nb[q]];
            }
        }
        X.src[k] = i4v{c[0], c[1], c[2], c[3]};
        X.kv[k] = f4v{v[0], v[1], v[2], v[3]};
    }
}

// Values, layer by layer (workgroup 0: A, 1: B), sums in the reference's
// neighbour order.  A layer of at most kExtraLds entries keeps its values in
// LDS for the next layer; a larger one is read back from global memory.  A
// workgroup barrier orders the layers (on gfx950 it waits for LDS only; the
// workgroup's global accesses stay ordered on its CU), so the records of a
// thread's entries are loaded kExtraRing layers ahead into a register ring,
// and the layer offsets sit in LDS: a layer then costs an LDS round trip and
// a barrier.
constexpr int kExtraOffLds = 8192;
struct ExtraRec {
    i4v c;
    f4v kv;
    int f;
};
__device__ __forceinline__ float extrap_entry(const ExtraArr& X, const ExtraRec& r, bool prev_lds,
                                              const float* prevb, int pb) {
    const int cq[4] = {r.c.x, r.c.y, r.c.z, r.c.w};
    const float vq[4] = {r.kv.x, r.kv.y, r.kv.z, r.kv.w};
    float sum = 0.0f;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (cq[q] == -1) continue;
        float x;
        if (cq[q] == -2) x = vq[q];
        else if (prev_lds) x = prevb[cq[q] - pb];
        else x = X.val[X.order[cq[q]]];
        sum += x;
        cnt++;
    }
    return cnt == 0 ? 0.0f : sum / (float)cnt;
}
// Fast path: every layer of at most 2 x 1024 entries and the offsets in LDS.
// Each thread owns entries kb + tid and kb + tid + 1024 of a layer; their
// records are fetched kExtraRing layers ahead from clamped (always valid)
// positions, so no global load is ever waited for early.  The vector memory
// counter also counts stores, so waiting for a layer's records waits for the
// value stores issued before them too: four layers ahead leaves those stores
// time to complete (two: the pass took 6% longer on a 512^2 dam break).
constexpr int kExtraRing = 4;
__device__ __forceinline__ void extrap_values_fast(const ExtraArr& X, const int* offl, int nl, float (*buf)[kExtraLds]) {
    const int tid = threadIdx.x;
    const int last = offl[nl + 1] - 1;  // the last ordered entry
    auto fetch = [&](int L, ExtraRec (&r)[2]) {
        const int kb = offl[L <= nl ? L : nl];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int k = min(kb + tid + 1024 * e, last);
            r[e].c = X.src[k];
            r[e].kv = X.kv[k];
            r[e].f = X.order[k];
        }
    };
    ExtraRec ring[kExtraRing][2];
#pragma unroll
    for (int u = 0; u < kExtraRing; u++) fetch(1 + u, ring[u]);
    for (int L0 = 1; L0 <= nl; L0 += kExtraRing) {
#pragma unroll
        for (int u = 0; u < kExtraRing; u++) {
            const int L = L0 + u;
            if (L > nl) break;
            const int pb = offl[L - 1], kb = offl[L], ke = offl[L + 1];
            const float* prevb = buf[(L - 1) & 1];
            float* curb = buf[L & 1];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int k = kb + tid + 1024 * e;
                if (k < ke) {
                    const float val = extrap_entry(X, ring[u][e], L > 1, prevb, pb);
                    X.val[ring[u][e].f] = val;
                    curb[k - kb] = val;
                }
            }
            fetch(L + kExtraRing, ring[u]);
            __syncthreads();
        }
    }
}
__device__ __forceinline__ void extrap_values_body(ExtraArr A, ExtraArr B, int vb) {
    ExtraArr X = vb == 0 ? A : B;
    __shared__ float buf[2][kExtraLds];
    __shared__ int offl[kExtraOffLds];
    __shared__ int widest;
    const int nl = *X.count;
    if (nl == 0) return;
    const gptr<int> offg = extra_off(X);
    const int tid = threadIdx.x, bd = blockDim.x;
    const bool off_lds = nl + 2 <= kExtraOffLds;
    if (tid == 0) widest = 0;
    __syncthreads();
    if (off_lds) {
        int wmax = 0;
        for (int k = tid; k < nl + 2; k += bd) {
            offl[k] = offg[k];
            if (k >= 1 && k <= nl) wmax = max(wmax, offg[k + 1] - offg[k]);
        }
        atomicMax(&widest, wmax);
    }
    __syncthreads();
    if (off_lds && bd == 1024 && widest <= 2048) {
        extrap_values_fast(X, offl, nl, buf);
        return;
    }
    // general path: offsets and large layers from global memory
    for (int L = 1; L <= nl; L++) {
        const int pb = offg[L - 1], kb = offg[L], ke = offg[L + 1];
        const bool prev_lds = L > 1 && kb - pb <= kExtraLds;
        const bool cur_lds = ke - kb <= kExtraLds;
        const float* prevb = buf[(L - 1) & 1];
        float* curb = buf[L & 1];
        for (int k = kb + tid; k < ke; k += bd) {
            ExtraRec r;
            r.c = X.src[k];
            r.kv = X.kv[k];
            r.f = X.order[k];
            const float val = extrap_entry(X, r, prev_lds, prevb, pb);
            X.val[r.f] = val;
            if (cur_lds) curb[k - kb] = val;
        }
        __syncthreads();
    }
}
__global__ __launch_bounds__(kBlock) void extrap_rows_kernel(ExtraArr A, ExtraArr B) {
    extrap_rows_body(A, B, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kExtraColThreads) void extrap_cols_kernel(ExtraArr A, ExtraArr B) {
    extrap_cols_body(A, B, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void extrap_hist_kernel(ExtraArr A, ExtraArr B) {
    extrap_hist_body(A, B, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void extrap_scan_kernel(ExtraArr A, ExtraArr B) {
    extrap_scan_body(A, B, blockIdx.x);
}
__global__ __launch_bounds__(kBlock) void extrap_scatter_kernel(ExtraArr A, ExtraArr B) {
    extrap_scatter_body(A, B, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void extrap_link_kernel(ExtraArr A, ExtraArr B) {
    extrap_link_body(A, B, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(1024) void extrap_values_kernel(ExtraArr A, ExtraArr B) {
    extrap_values_body(A, B, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Particles.  Cell keys use the reference's truncation (int)(x/dx)
// (src/FluidSim2D.cpp:767) in FP64.  Sorting is a counting sort by cell key
// (histogram, exclusive scan, scatter) followed by an in-segment sort by
// particle index, so the order is (key, index): the order a stable sort of the
// particles by key produces, and the order in which the reference's serial
// loops meet the particles of one cell.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void particle_keys_body(const float* __restrict__ px,
                                                               const float* __restrict__ py, long np,
                                                               double dx, int nx, int ny, int* __restrict__ key,
                                                               int* __restrict__ counts, int vb, int nvb) {
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < np; e += (long)nvb * blockDim.x) {
        int x = (int)((double)px[e] / dx), y = (int)((double)py[e] / dx);
        x = clampi(x, 0, nx - 1);
        y = clampi(y, 0, ny - 1);
        const int k = y * nx + x;
        key[e] = k;
        atomicAdd(&counts[k], 1);
    }
}
__global__ __launch_bounds__(kBlock) void particle_keys_kernel(const float* __restrict__ px,
                                                               const float* __restrict__ py, long np,
                                                               double dx, int nx, int ny, int* __restrict__ key,
                                                               int* __restrict__ counts) {
    particle_keys_body(px, py, np, dx, nx, ny, key, counts, blockIdx.x, gridDim.x);
}

// Exclusive scan in three passes over 1024-element tiles.
__device__ __forceinline__ void scan_tiles_body(const int* __restrict__ in, int* __restrict__ out,
                                                            int* __restrict__ tile_sums, long n, int vb, int nvb) {
    __shared__ int sh[1024];
    const long base = (long)vb * 1024;
    for (int k = threadIdx.x; k < 1024; k += kBlock) sh[k] = base + k < n ? in[base + k] : 0;
    __syncthreads();
    if (threadIdx.x == 0) {  // 1024 adds: small next to the memory traffic of the tile
        int run = 0;
        for (int k = 0; k < 1024; k++) {
            int v = sh[k];
            sh[k] = run;
            run += v;
        }
        tile_sums[vb] = run;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 1024; k += kBlock)
        if (base + k < n) out[base + k] = sh[k];
}
__global__ __launch_bounds__(kBlock) void scan_tiles_kernel(const int* __restrict__ in, int* __restrict__ out,
                                                            int* __restrict__ tile_sums, long n) {
    scan_tiles_body(in, out, tile_sums, n, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void scan_sums_body(int* tile_sums, int ntiles, int vb, int nvb) {
    if (threadIdx.x == 0 && vb == 0) {
        int run = 0;
        for (int k = 0; k < ntiles; k++) {
            int v = tile_sums[k];
            tile_sums[k] = run;
            run += v;
        }
    }
}
__global__ void scan_sums_kernel(int* tile_sums, int ntiles) {
    scan_sums_body(tile_sums, ntiles, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void scan_add_body(int* __restrict__ out, const int* __restrict__ tile_sums,
                                                          long n, int vb, int nvb) {
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < n; e += (long)nvb * blockDim.x)
        out[e] += tile_sums[e / 1024];
}
__global__ __launch_bounds__(kBlock) void scan_add_kernel(int* __restrict__ out, const int* __restrict__ tile_sums,
                                                          long n) {
    scan_add_body(out, tile_sums, n, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void particle_scatter_body(const int* __restrict__ key, long np,
                                                                  int* __restrict__ cursor,
                                                                  int* __restrict__ sorted, int vb, int nvb) {
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < np; e += (long)nvb * blockDim.x) {
        int slot = atomicAdd(&cursor[key[e]], 1);
        sorted[slot] = (int)e;
    }
}
__global__ __launch_bounds__(kBlock) void particle_scatter_kernel(const int* __restrict__ key, long np,
                                                                  int* __restrict__ cursor,
                                                                  int* __restrict__ sorted) {
    particle_scatter_body(key, np, cursor, sorted, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void segment_sort_body(const int* __restrict__ start,
                                                              const int* __restrict__ counts,
                                                              int* __restrict__ sorted, long ncells, int vb, int nvb) {
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < ncells; c += (long)nvb * blockDim.x) {
        const int s0 = start[c], n = counts[c];
        for (int a = 1; a < n; a++) {
            int v = sorted[s0 + a], b = a - 1;
            while (b >= 0 && sorted[s0 + b] > v) {
                sorted[s0 + b + 1] = sorted[s0 + b];
                b--;
            }
            sorted[s0 + b + 1] = v;
        }
    }
}
__global__ __launch_bounds__(kBlock) void segment_sort_kernel(const int* __restrict__ start,
                                                              const int* __restrict__ counts,
                                                              int* __restrict__ sorted, long ncells) {
    segment_sort_body(start, counts, sorted, ncells, blockIdx.x, gridDim.x);
}

// FluidSim2D::transferVelocityToGrid (src/FluidSim2D.cpp:144-204) as a gather:
// every face sums the tent weights (Array2D::linearDistribute, Array2D.h:361-376)
// of the particles of the cells that can reach it, divides by the weight sum,
// and faces whose value is exactly 0 become unknown (:192-201).
__device__ __forceinline__ void p2g_body(const float* __restrict__ px, const float* __restrict__ py,
                                                     const float* __restrict__ pvx, const float* __restrict__ pvy,
                                                     const int* __restrict__ start, const int* __restrict__ counts,
                                                     const int* __restrict__ sorted, float* __restrict__ u,
                                                     float* __restrict__ v, uint32_t* __restrict__ mu,
                                                     uint32_t* __restrict__ mv, int nx, int ny, double dx, int vb, int nvb) {
    const long nU = (long)(nx + 1) * ny, nV = (long)nx * (ny + 1);
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < nU + nV; e += (long)nvb * blockDim.x) {
        const bool isU = e < nU;
        const long f = isU ? e : e - nU;
        const int W = isU ? nx + 1 : nx, H = isU ? ny : ny + 1;
        const int i = (int)(f % W), j = (int)(f / W);
        // cells whose particles reach this face: u: x in {i-1,i}, y in {j-1..j+1};
        // v: x in {i-1..i+1}, y in {j-1,j}
        const int cx0 = i - 1, cx1 = isU ? i : i + 1;
        const int cy0 = j - 1, cy1 = isU ? j + 1 : j;
        double sv = 0, sw = 0;
        for (int cy = cy0; cy <= cy1; cy++) {
            if (cy < 0 || cy >= ny) continue;
            for (int cx = cx0; cx <= cx1; cx++) {
                if (cx < 0 || cx >= nx) continue;
                const int c = cy * nx + cx, s0 = start[c], n = counts[c];
                for (int k = 0; k < n; k++) {
                    const int pe = sorted[s0 + k];
                    double gx = (double)px[pe] / dx, gy = (double)py[pe] / dx;
                    if (isU) gy -= 0.5;
                    else gx -= 0.5;
                    const int ui = (int)gx, uj = (int)gy;
                    const double ddx = gx - ui, ddy = gy - uj;
                    const int x1 = clampi(ui, 0, W - 1), x2 = clampi(ui + 1, 0, W - 1);
                    const int y1 = clampi(uj, 0, H - 1), y2 = clampi(uj + 1, 0, H - 1);
                    double wt = 0;
                    if (y1 == j) {
                        if (x1 == i) wt += (1 - ddx) * (1 - ddy);
                        if (x2 == i) wt += ddx * (1 - ddy);
                    }
                    if (y2 == j) {
                        if (x1 == i) wt += (1 - ddx) * ddy;
                        if (x2 == i) wt += ddx * ddy;
                    }
                    if (wt != 0) {
                        sv += wt * (double)(isU ? pvx[pe] : pvy[pe]);
                        sw += wt;
                    }
                }
            }
        }
        const double val = sw > 0 ? sv / sw : sv;
        const float fv = (float)val;
        if (isU) {
            u[f] = fv;
            mu[f] = fv == 0.0f ? 0xFFFFFFFFu : 0u;
        } else {
            v[f] = fv;
            mv[f] = fv == 0.0f ? 0xFFFFFFFFu : 0u;
        }
    }
}
__global__ __launch_bounds__(kBlock) void p2g_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                     const float* __restrict__ pvx, const float* __restrict__ pvy,
                                                     const int* __restrict__ start, const int* __restrict__ counts,
                                                     const int* __restrict__ sorted, float* __restrict__ u,
                                                     float* __restrict__ v, uint32_t* __restrict__ mu,
                                                     uint32_t* __restrict__ mv, int nx, int ny, double dx) {
    p2g_body(px, py, pvx, pvy, start, counts, sorted, u, v, mu, mv, nx, ny, dx, blockIdx.x, gridDim.x);
}

// FluidSim2D::updateParticleVelocities (src/FluidSim2D.cpp:552-568): PIC/FLIP
// blend of bilinear samples of newMac and of (newMac - mac).
__device__ __forceinline__ void bilinear_pair(const float* __restrict__ a_new, const float* __restrict__ a_old,
                                              int w, int h, double gx, double gy, float& pic, float& diff) {
    int ui = (int)gx, uj = (int)gy;
    float ddx = (float)(gx - ui), ddy = (float)(gy - uj);
    int x1 = clampi(ui, 0, w - 1), x2 = clampi(ui + 1, 0, w - 1);
    int y1 = clampi(uj, 0, h - 1), y2 = clampi(uj + 1, 0, h - 1);
    const long i11 = (long)y1 * w + x1, i21 = (long)y1 * w + x2, i12 = (long)y2 * w + x1, i22 = (long)y2 * w + x2;
    const float w11 = (1 - ddx) * (1 - ddy), w21 = ddx * (1 - ddy), w12 = (1 - ddx) * ddy, w22 = ddx * ddy;
    pic = a_new[i11] * w11 + a_new[i21] * w21 + a_new[i12] * w12 + a_new[i22] * w22;
    diff = (a_new[i11] - a_old[i11]) * w11 + (a_new[i21] - a_old[i21]) * w21 + (a_new[i12] - a_old[i12]) * w12 +
           (a_new[i22] - a_old[i22]) * w22;
}

__device__ __forceinline__ void g2p_body(const float* __restrict__ px, const float* __restrict__ py,
                                                     float* __restrict__ pvx, float* __restrict__ pvy, long np,
                                                     const float* __restrict__ nu, const float* __restrict__ nv,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     int nx, int ny, double dx, float alpha, int vb, int nvb) {
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < np; e += (long)nvb * blockDim.x) {
        const double gx = (double)px[e] / dx, gy = (double)py[e] / dx;
        float picu, du, picv, dv;
        bilinear_pair(nu, u, nx + 1, ny, gx, gy - 0.5, picu, du);
        bilinear_pair(nv, v, nx, ny + 1, gx - 0.5, gy, picv, dv);
        const float fu = pvx[e] + du, fv = pvy[e] + dv;
        pvx[e] = alpha * picu + (1 - alpha) * fu;
        pvy[e] = alpha * picv + (1 - alpha) * fv;
    }
}
__global__ __launch_bounds__(kBlock) void g2p_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                     float* __restrict__ pvx, float* __restrict__ pvy, long np,
                                                     const float* __restrict__ nu, const float* __restrict__ nv,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     int nx, int ny, double dx, float alpha) {
    g2p_body(px, py, pvx, pvy, np, nu, nv, u, v, nx, ny, dx, alpha, blockIdx.x, gridDim.x);
}

// FluidSim2D::applyAdvection (src/FluidSim2D.cpp:570-604): CFL scan, forward
// Ralston RK3 through the Catmull-Rom field, clampPos, NaN flag (on the
// position before the move, as the reference checks it).
// rk2: the midpoint rule x + dt v(x + dt/2 v(x)) instead of the reference's
// RK3 (the 512^2 PIC/FLIP dam-break workload's option).
__device__ __forceinline__ void particle_advect_body(float* __restrict__ px, float* __restrict__ py,
                                                                 const float* __restrict__ pvx,
                                                                 const float* __restrict__ pvy, long np, MacView m,
                                                                 double dt, ClampBox cb, Diag* dg, double* partials,
                                                                 int rk2, int vb, int nvb) {
    double cmax = 0;
    int nan = 0;
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < np; e += (long)nvb * blockDim.x) {
        const double c = ((double)pvx[e] + (double)pvy[e]) * dt / m.dx;
        cmax = c > cmax ? c : cmax;
        const double x = px[e], y = py[e];
        const double k1x = vel_interp_u(m, x, y), k1y = vel_interp_v(m, x, y);
        const double x2 = x + 0.5 * dt * k1x, y2 = y + 0.5 * dt * k1y;
        const double k2x = vel_interp_u(m, x2, y2), k2y = vel_interp_v(m, x2, y2);
        double nx_, ny_;
        if (rk2) {
            nx_ = x + dt * k2x;
            ny_ = y + dt * k2y;
        } else {
            const double x3 = x + 0.75 * dt * k2x, y3 = y + 0.75 * dt * k2y;
            const double k3x = vel_interp_u(m, x3, y3), k3y = vel_interp_v(m, x3, y3);
            nx_ = x + (2. / 9.) * dt * k1x;
            ny_ = y + (2. / 9.) * dt * k1y;
            nx_ = nx_ + (3. / 9.) * dt * k2x;
            ny_ = ny_ + (3. / 9.) * dt * k2y;
            nx_ = nx_ + (4. / 9.) * dt * k3x;
            ny_ = ny_ + (4. / 9.) * dt * k3y;
        }
        if (isnan(x) || isnan(y)) nan = 1;
        px[e] = (float)clampd(nx_, cb.lox, cb.hix);
        py[e] = (float)clampd(ny_, cb.loy, cb.hiy);
    }
    double bm = block_max(cmax);
    __shared__ int sh_nan;
    if (threadIdx.x == 0) sh_nan = 0;
    __syncthreads();
    if (nan) atomicOr(&sh_nan, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[vb] = bm;
        if (sh_nan) atomicOr(&dg->nan_flag, 1);
    }
}
__global__ __launch_bounds__(kBlock) void particle_advect_kernel(float* __restrict__ px, float* __restrict__ py,
                                                                 const float* __restrict__ pvx,
                                                                 const float* __restrict__ pvy, long np, MacView m,
                                                                 double dt, ClampBox cb, Diag* dg, double* partials,
                                                                 int rk2) {
    particle_advect_body(px, py, pvx, pvy, np, m, dt, cb, dg, partials, rk2, blockIdx.x, gridDim.x);
}

// Max of the per-block CFL partials (one block).
__device__ __forceinline__ void cfl_finalize_body(const double* partials, int n, Diag* dg, int vb, int nvb) {
    double t = max_partials(partials, n);
    if (threadIdx.x == 0) dg->cfl_max = t;
}
__global__ __launch_bounds__(kBlock) void cfl_finalize_kernel(const double* partials, int n, Diag* dg) {
    cfl_finalize_body(partials, n, dg, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Level set (src/FluidSim2D.cpp:752-920), in FP64 and without contraction so
// that every distance, comparison and sign equals the reference's double
// arithmetic on the same positions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double corner_dist(int i, int j, double x, double y, double dx, double dr) {
#pragma clang fp contract(off)
    const double ex = x - i * dx, ey = y - j * dx;
    return sqrt(ex * ex + ey * ey) - dr;
}

// Scatter-min of constructFromParticles (:764-773): the first particle (in
// index order) with the smallest distance to the cell's corner wins.
__device__ __forceinline__ void ls_seed_body(const float* __restrict__ px, const float* __restrict__ py,
                                                         const int* __restrict__ start, const int* __restrict__ counts,
                                                         const int* __restrict__ sorted, int* __restrict__ t,
                                                         int nx, long ncells, double dx, double dr, int vb, int nvb) {
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < ncells; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        double best = HUGE_VAL;
        int bt = -1;
        const int s0 = start[c], n = counts[c];
        for (int k = 0; k < n; k++) {
            const int e = sorted[s0 + k];
            const double d = corner_dist(i, j, px[e], py[e], dx, dr);
            if (d < best) {
                best = d;
                bt = e;
            }
        }
        t[c] = bt;
    }
}
__global__ __launch_bounds__(kBlock) void ls_seed_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                         const int* __restrict__ start, const int* __restrict__ counts,
                                                         const int* __restrict__ sorted, int* __restrict__ t,
                                                         int nx, long ncells, double dx, double dr) {
    ls_seed_body(px, py, start, counts, sorted, t, nx, ncells, dx, dr, blockIdx.x, gridDim.x);
}

// One directional Gauss-Seidel sweep of the closest-particle propagation
// (FluidSim2D.cpp:775-792).  The cell adopts, among its own entry and its four
// neighbours' entries taken in the order (cell, i-1, i+1, j-1, j+1), the
// first one with the strictly smallest corner distance: the lexicographic
// minimum of (distance, position in that order).  Three of the five do not
// depend on the sweep front (the cell's own entry and the two neighbours not
// yet updated in this sweep, all from the previous sweep's array), so
// ls_cand_sk_body merges them beforehand, in parallel, into one candidate; the
// sweep compares only the two candidates its front produces.  FX / FY: the
// sweep runs towards -x / -y.
struct LsCand {
    double d;    // corner distance; HUGE_VAL: no candidate
    int t, pos;  // particle (-1: none), position in the candidate order
    float x, y;  // the particle's position
};

// tr(i, j): the adopted particle of cell (i, j) after the previous pass
template <class TRead>
__device__ __forceinline__ LsCand ls_cand_cell(TRead tr, const float* __restrict__ px,
                                               const float* __restrict__ py, int nx, int ny, int i, int j, int fx,
                                               int fy, double dx, double dr) {
    LsCand best{HUGE_VAL, -1, 0, 0.0f, 0.0f};
    const int tc = tr(i, j);
    if (tc >= 0) {
        const float x = px[tc], y = py[tc];
        best = LsCand{corner_dist(i, j, x, y, dx, dr), tc, 0, x, y};
    }
    // not yet updated: (i+1, j) sweeping +x (position 2), (i-1, j) sweeping -x (position 1); likewise y
    const int ii = fx ? i - 1 : i + 1, jj = fy ? j - 1 : j + 1;
    if (ii >= 0 && ii < nx) {
        const int tn = tr(ii, j);
        if (tn >= 0) {
            const float x = px[tn], y = py[tn];
            const double d = corner_dist(i, j, x, y, dx, dr);
            if (d < best.d) best = LsCand{d, tn, fx ? 1 : 2, x, y};
        }
    }
    if (jj >= 0 && jj < ny) {
        const int tn = tr(i, jj);
        if (tn >= 0) {
            const float x = px[tn], y = py[tn];
            const double d = corner_dist(i, j, x, y, dx, dr);
            if (d < best.d) best = LsCand{d, tn, fy ? 3 : 4, x, y};
        }
    }
    return best;
}
// The closest-particle passes on the solver's sweep engine (fs2d_sweep.hpp).
// A pass in direction (fx, fy) is a forward sweep in the flipped coordinates
// (a, b) = (fx ? nx-1-i : i, fy ? ny-1-j : j).  Its candidates (those of
// ls_cand_cell) are laid out in SK4 order of (a, b) as three words: (t + 1) |
// pos << 29 (0: none), x, y.  The sweep recomputes a candidate's distance with
// the same corner_dist call, so every comparison sees the same value.  Words
// are moved, never computed on, so the bits survive the float arrays.  The
// SK4 arrays are the solver's scratch (r, z, q in; w out), free between
// projections.
constexpr int kCandPosShift = 29;  // particle indices below 2^29 - 1
// The passes stop early.  A pass (candidates and sweep) that changes no
// cell's adopted particle leaves every cell's particle at least as close as
// each of its four neighbours' (every cell of a pass compares all four; the
// direction only orders the visits, and ties never move a particle), so every
// later pass, in any direction, changes nothing either: the reference's
// remaining passes are the identity.  chg[k] != 0 records that pass k changed
// a cell, or may have (a sweep adoption that keeps the same particle counts
// too); the passes after the first unchanged one return at once, and the
// state is read from that pass's output.  chg holds 16 ints, cleared before
// pass 0.
__device__ __forceinline__ int ls_first_unchanged(const int* __restrict__ chg, int k) {
    for (int p = 0; p < k; p++)
        if (chg[p] == 0) return p;
    return -1;
}
__device__ __forceinline__ void ls_cand_sk_body(SkGeom g, const int* __restrict__ t_rm, const float* __restrict__ t_sk,
                                                int pfx, int pfy, const float* __restrict__ px,
                                                const float* __restrict__ py, float* __restrict__ w0,
                                                float* __restrict__ w1, float* __restrict__ w2, int fx, int fy,
                                                double dx, double dr, int* chg, int k, int vb, int nvb) {
    if (ls_first_unchanged(chg, k) >= 0) return;  // the same for every thread
    // SK4 entries of this pass, so the three stores coalesce; the sweep never
    // reads a padding entry's record (steps outside the grid reset its
    // state), so padding entries are skipped
    const int nx = g.nx, ny = g.ny;
    auto tr_rm = [&](int i, int j) { return t_rm[(long)j * nx + i]; };
    auto tr_sk = [&](int i, int j) {
        return (int)__float_as_uint(t_sk[g.idx(pfx ? nx - 1 - i : i, pfy ? ny - 1 - j : j)]);
    };
    const long n = g.size();
    bool changed = false;
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < n; e += (long)nvb * blockDim.x) {
        int a, b;
        if (!sk_cell(g, e, a, b)) continue;
        const int i = fx ? nx - 1 - a : a, j = fy ? ny - 1 - b : b;
        const LsCand best = t_rm ? ls_cand_cell(tr_rm, px, py, nx, ny, i, j, fx, fy, dx, dr)
                                 : ls_cand_cell(tr_sk, px, py, nx, ny, i, j, fx, fy, dx, dr);
        changed |= best.t != (t_rm ? tr_rm(i, j) : tr_sk(i, j));
        uint32_t r0 = 0;
        float rx = 0.0f, ry = 0.0f;
        if (best.t >= 0) {
            r0 = (uint32_t)(best.t + 1) | ((uint32_t)best.pos << kCandPosShift);
            rx = best.x;
            ry = best.y;
        }
        w0[e] = __uint_as_float(r0);
        w1[e] = rx;
        w2[e] = ry;
    }
    if (__syncthreads_or(changed) && threadIdx.x == 0) atomicOr(chg + k, 1);
}
__global__ __launch_bounds__(kBlock) void ls_cand_sk_kernel(SkGeom g, const int* __restrict__ t_rm,
                                                            const float* __restrict__ t_sk, int pfx, int pfy,
                                                            const float* __restrict__ px, const float* __restrict__ py,
                                                            float* __restrict__ w0, float* __restrict__ w1,
                                                            float* __restrict__ w2, int fx, int fy, double dx,
                                                            double dr, int* chg, int k) {
    ls_cand_sk_body(g, t_rm, t_sk, pfx, pfy, px, py, w0, w1, w2, fx, fy, dx, dr, chg, k, blockIdx.x, gridDim.x);
}
// The adopted indices back to row-major, from the output of the last pass
// that ran (SK4 order of its direction; pass 15 when every pass changed).
__device__ __forceinline__ void closest_t_to_rm_body(SkGeom g, const float* __restrict__ tk, int* __restrict__ t,
                                                     const int* __restrict__ chg, int vb, int nvb) {
    const int p = ls_first_unchanged(chg, 16);
    const int d = (p >= 0 ? p : 15) & 3, fx = d & 1, fy = d >> 1;
    const long n = (long)g.nx * g.ny;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % g.nx), j = (int)(c / g.nx);
        t[c] = (int)__float_as_uint(tk[g.idx(fx ? g.nx - 1 - i : i, fy ? g.ny - 1 - j : j)]);
    }
}
__global__ __launch_bounds__(kBlock) void closest_t_to_rm_kernel(SkGeom g, const float* __restrict__ tk,
                                                                 int* __restrict__ t, const int* __restrict__ chg) {
    closest_t_to_rm_body(g, tk, t, chg, blockIdx.x, gridDim.x);
}

// The sweep operator: carried state (t + 1, x, y) of the adopted particle (0:
// none, the engine's zero boundary); step q of lane l is cell (a, b) = (t0 + q
// - l, 64 s + l), and steps outside the grid leave the state empty: a
// neighbour outside the grid offers no particle.
template <int FX, int FY>
struct ClosestSweepOp {
    static constexpr int NW = 3, NIN = 3, NF = 5, NB = 0, NOUT = 1;
    const float* in[NIN];  // (t + 1) | pos << 29, x, y of the cell's candidate
    const uint8_t* bytes;
    float* out[NOUT];      // adopted particle index (bits; -1: none)
    int nx, ny;
    double dx, dr;
    int* chg;              // per-pass change flags (ls_first_unchanged)
    int k;                 // this pass
    bool changed = false;  // an adoption in this lane's cells
    __device__ bool skip(bool) const { return ls_first_unchanged(chg, k) >= 0; }
    // The loaders also compute the candidate's distance (the ring's words 3
    // and 4: its low and high halves), which takes one of the compute wave's
    // three FP64 corner distances per step off the chain.
    __device__ void xform(const float4 (&in)[NIN], float4 (&x)[NF], double&, int t0, int s) const {
        const int lane = threadIdx.x & 63;
        const int b = 64 * s + lane;
        const int j = FY ? ny - 1 - b : b;
        x[0] = in[0];
        x[1] = in[1];
        x[2] = in[2];
        float lo[4], hi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int a = t0 + q - lane;
            const int i = FX ? nx - 1 - a : a;
            const uint32_t r0 = __float_as_uint(f4get(in[0], q));
            const int bt = (int)(r0 & ((1u << kCandPosShift) - 1u)) - 1;
            const double bd = bt >= 0 ? corner_dist(i, j, f4get(in[1], q), f4get(in[2], q), dx, dr) : HUGE_VAL;
            uint32_t w[2];
            d2u(bd, w);
            lo[q] = __uint_as_float(w[0]);
            hi[q] = __uint_as_float(w[1]);
        }
        x[3] = make_float4(lo[0], lo[1], lo[2], lo[3]);
        x[4] = make_float4(hi[0], hi[1], hi[2], hi[3]);
    }
    __device__ void loader_done(int, int, double) const {}
    template <bool B, bool C>
    __device__ void group(const float4 (&x)[NF], uint32_t, uint32_t (&st)[NW], const uint4 (&hb)[NW],
                          float (&o)[NOUT][4], uint32_t (&prod)[NW][4], int t0, int s) {
        const int lane = threadIdx.x & 63;
        const int b = 64 * s + lane;
        const int j = FY ? ny - 1 - b : b;
        bool took = false;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t nb[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) nb[w] = nbr<B, C>(u4get(hb[w], q), st[w]);
            const int a = t0 + q - lane;
            const bool ing = a >= 0 && a < nx && b < ny;
            const int i = FX ? nx - 1 - a : a;
            const uint32_t r0 = __float_as_uint(f4get(x[0], q));
            int bt = (int)(r0 & ((1u << kCandPosShift) - 1u)) - 1;
            int bp = (int)(r0 >> kCandPosShift);
            float bx = f4get(x[1], q), by = f4get(x[2], q);
            const uint32_t bw[2] = {__float_as_uint(f4get(x[3], q)), __float_as_uint(f4get(x[4], q))};
            double bd = u2d(bw);  // the candidate's distance, from the loaders
            // the updated neighbours, (a-1, b) then (a, b-1): positions 1 / 3 sweeping
            // +x / +y, else 2 / 4; selects, not branches
            auto consider = [&](const uint32_t* sw, int pos) {
                const int tn = (int)sw[0] - 1;
                const float qx = u2f(sw[1]), qy = u2f(sw[2]);
                const double d = corner_dist(i, j, qx, qy, dx, dr);
                const bool take = (tn >= 0) & ((d < bd) | ((d == bd) & (pos < bp)));
                took |= take & ing;
                bd = take ? d : bd;
                bt = take ? tn : bt;
                bp = take ? pos : bp;
                bx = take ? qx : bx;
                by = take ? qy : by;
            };
            consider(st, FX ? 2 : 1);
            consider(nb, FY ? 4 : 3);
            st[0] = ing ? (uint32_t)(bt + 1) : 0u;
            st[1] = ing ? f2u(bx) : 0u;
            st[2] = ing ? f2u(by) : 0u;
#pragma unroll
            for (int w = 0; w < NW; w++) prod[w][q] = st[w];
            o[0][q] = __uint_as_float((uint32_t)bt);
        }
        changed |= took;
    }
    // wave 0 after the strip's last step: one flag write per strip that adopted
    __device__ void strip_done(int) const {
        if (__any(changed) && (threadIdx.x & 63) == 0) atomicOr(chg + k, 1);
    }
};

// phi from the adopted particle; then the surface marks (:803-834) and the
// interior fill (:836-842) of redistance, reading phi of the finished sweeps.
__device__ __forceinline__ void ls_phi_body(const int* __restrict__ t, const float* __restrict__ px,
                                                        const float* __restrict__ py, double* __restrict__ phi,
                                                        int nx, long ncells, double dx, double dr, int vb, int nvb) {
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < ncells; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        const int e = t[c];
        phi[c] = e >= 0 ? corner_dist(i, j, px[e], py[e], dx, dr) : HUGE_VAL;
    }
}
__global__ __launch_bounds__(kBlock) void ls_phi_kernel(const int* __restrict__ t, const float* __restrict__ px,
                                                        const float* __restrict__ py, double* __restrict__ phi,
                                                        int nx, long ncells, double dx, double dr) {
    ls_phi_body(t, px, py, phi, nx, ncells, dx, dr, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void ls_surface_body(const double* __restrict__ phi, double* __restrict__ out,
                                                            int nx, int ny, int vb, int nvb) {
    const long n = (long)nx * ny;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        const double p0 = phi[c];
        bool surf = false;
        // pairs (i,j)-(i+1,j) and (i,j)-(i,j+1) are examined for i < nx-1, j < ny-1
        if (i < nx - 1 && j < ny - 1) surf = surf || (p0 * phi[c + 1] < 0) || (p0 * phi[c + nx] < 0);
        if (i >= 1 && j < ny - 1) surf = surf || (phi[c - 1] * p0 < 0);
        if (j >= 1 && i < nx - 1) surf = surf || (phi[c - nx] * p0 < 0);
        out[c] = (!surf && p0 < 0) ? -HUGE_VAL : p0;
    }
}
__global__ __launch_bounds__(kBlock) void ls_surface_kernel(const double* __restrict__ phi, double* __restrict__ out,
                                                            int nx, int ny) {
    ls_surface_body(phi, out, nx, ny, blockIdx.x, gridDim.x);
}

// The eikonal update of one cell (:845-902) from its two already-updated
// neighbours: only a cell with phi < 0 changes.  FP64, no contraction, as the
// reference compiles it.
__device__ __forceinline__ double eik_update(double cur, double left, double below, double dx) {
#pragma clang fp contract(off)
    if (!(cur >= 0)) {
        const double A = fabs(left), B = fabs(below);
        const double phi0 = A < B ? A : B;
        const double phi1 = A > B ? A : B;
        double d = phi0 + dx;
        if (d > phi1) d = 0.5 * (phi0 + phi1 + sqrt(2 * dx * dx - (phi1 - phi0) * (phi1 - phi0)));
        if (d < -cur) cur = -d;
    }
    return cur;
}

// ---------------------------------------------------------------------------
// The eikonal passes on the solver's sweep engine (fs2d_sweep.hpp), whose
// loader and poller waves take the loads and hand-offs off the compute wave.
// A pass in direction (FX, FY) is a forward sweep in the flipped coordinates
// (a, b) = (FX ? nx-1-i : i, FY ? ny-1-j : j): its distances are laid out in
// SK4 order of (a, b) as low and high 32-bit words (bit moves only, no float
// arithmetic), with a flag byte per entry (a >= 1 and b >= 1: the cells the
// update may change; 0 on padding).  The SK4 arrays are the solver's scratch
// (z, q in; w, s1 out; cpl flags), free between projections.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eik_to_sk_body(SkGeom g, const double* __restrict__ phi, float* __restrict__ lo,
                                               float* __restrict__ hi, uint8_t* __restrict__ flag, int fx, int fy,
                                               int vb, int nvb) {
    const long n = g.size();
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < n; e += (long)nvb * blockDim.x) {
        int a, b;
        const bool in = sk_cell(g, e, a, b);  // flipped coordinates
        double v = 0.0;
        if (in) v = phi[(long)(fy ? g.ny - 1 - b : b) * g.nx + (fx ? g.nx - 1 - a : a)];
        uint32_t w[2];
        d2u(v, w);
        lo[e] = __uint_as_float(w[0]);
        hi[e] = __uint_as_float(w[1]);
        flag[e] = (uint8_t)(in && a >= 1 && b >= 1);
    }
}
__global__ __launch_bounds__(kBlock) void eik_to_sk_kernel(SkGeom g, const double* __restrict__ phi,
                                                           float* __restrict__ lo, float* __restrict__ hi,
                                                           uint8_t* __restrict__ flag, int fx, int fy) {
    eik_to_sk_body(g, phi, lo, hi, flag, fx, fy, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void eik_from_sk_body(SkGeom g, const float* __restrict__ lo, const float* __restrict__ hi,
                                                 double* __restrict__ phi, int fx, int fy, int vb, int nvb) {
    const long n = (long)g.nx * g.ny;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % g.nx), j = (int)(c / g.nx);
        const long e = g.idx(fx ? g.nx - 1 - i : i, fy ? g.ny - 1 - j : j);
        const uint32_t w[2] = {__float_as_uint(lo[e]), __float_as_uint(hi[e])};
        phi[c] = u2d(w);
    }
}
__global__ __launch_bounds__(kBlock) void eik_from_sk_kernel(SkGeom g, const float* __restrict__ lo,
                                                             const float* __restrict__ hi, double* __restrict__ phi,
                                                             int fx, int fy) {
    eik_from_sk_body(g, lo, hi, phi, fx, fy, blockIdx.x, gridDim.x);
}

// Between passes: the previous pass's output (SK4 order of (fx, fy)) moved to
// the next pass's input order (fx2, fy2), with its flags.
__device__ __forceinline__ void eik_sk_to_sk_body(SkGeom g, const float* __restrict__ lo, const float* __restrict__ hi,
                                                  float* __restrict__ lo2, float* __restrict__ hi2,
                                                  uint8_t* __restrict__ flag, int fx, int fy, int fx2, int fy2,
                                                  int vb, int nvb) {
    const long n = g.size();
    for (long e2 = vb * (long)blockDim.x + threadIdx.x; e2 < n; e2 += (long)nvb * blockDim.x) {
        int a2, b2;
        const bool in = sk_cell(g, e2, a2, b2);
        float vl = 0.0f, vh = 0.0f;
        if (in) {
            const int i = fx2 ? g.nx - 1 - a2 : a2, j = fy2 ? g.ny - 1 - b2 : b2;
            const long e = g.idx(fx ? g.nx - 1 - i : i, fy ? g.ny - 1 - j : j);
            vl = lo[e];
            vh = hi[e];
        }
        lo2[e2] = vl;
        hi2[e2] = vh;
        flag[e2] = (uint8_t)(in && a2 >= 1 && b2 >= 1);
    }
}
__global__ __launch_bounds__(kBlock) void eik_sk_to_sk_kernel(SkGeom g, const float* __restrict__ lo,
                                                              const float* __restrict__ hi, float* __restrict__ lo2,
                                                              float* __restrict__ hi2, uint8_t* __restrict__ flag,
                                                              int fx, int fy, int fx2, int fy2) {
    eik_sk_to_sk_body(g, lo, hi, lo2, hi2, flag, fx, fy, fx2, fy2, blockIdx.x, gridDim.x);
}

// The sweep operator: carried state phi (FP64) in two words; the row
// neighbour is the lane's own previous result, the column neighbour lane
// l-1's, moved by DPP (the edge lane's from the strip below).
struct EikSweepOp {
    static constexpr int NW = 2, NIN = 2, NF = 2, NB = 1, NOUT = 2;
    const float* in[NIN];  // low, high words of phi
    const uint8_t* bytes;  // 1 where the update may change the cell
    float* out[NOUT];      // low, high words of the result
    double dx;
    __device__ bool skip(bool) const { return false; }
    __device__ void xform(const float4 (&in)[NIN], float4 (&x)[NF], double&, int, int) const {
        x[0] = in[0];
        x[1] = in[1];
    }
    __device__ void loader_done(int, int, double) const {}
    template <bool B, bool C>
    __device__ void group(const float4 (&x)[NF], uint32_t b4, uint32_t (&st)[NW], const uint4 (&hb)[NW],
                          float (&o)[NOUT][4], uint32_t (&prod)[NW][4], int, int) const {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t nb[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) nb[w] = nbr<B, C>(u4get(hb[w], q), st[w]);
            const uint32_t cw[2] = {__float_as_uint(f4get(x[0], q)), __float_as_uint(f4get(x[1], q))};
            double cur = u2d(cw);
            if ((b4 >> (8 * q)) & 0xFFu) cur = eik_update(cur, u2d(st), u2d(nb), dx);
            uint32_t r[2];
            d2u(cur, r);
#pragma unroll
            for (int w = 0; w < NW; w++) {
                st[w] = r[w];
                o[w][q] = __uint_as_float(r[w]);
                prod[w][q] = r[w];
            }
        }
    }
    __device__ void strip_done(int) const {}
};

// Smoothing pass phi = min(phi, mean of 4 neighbours) on 1..n-2 (:904-919).
__device__ __forceinline__ void ls_smooth_body(const double* __restrict__ in, double* __restrict__ out,
                                                           int nx, int ny, int vb, int nvb) {
#pragma clang fp contract(off)
    const long n = (long)nx * ny;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        double v = in[c];
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
            const double avg = 0.25 * (in[c - 1] + in[c + 1] + in[c - nx] + in[c + nx]);
            if (avg < v) v = avg;
        }
        out[c] = v;
    }
}
__global__ __launch_bounds__(kBlock) void ls_smooth_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                           int nx, int ny) {
    ls_smooth_body(in, out, nx, ny, blockIdx.x, gridDim.x);
}

// Classification (:657-664): non-solid cells are FLUID iff phi < 0.
__device__ __forceinline__ void classify_body(uint8_t* __restrict__ cell, const double* __restrict__ phi, long n, int vb, int nvb) {
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x)
        if (cell[c] != SOLID) cell[c] = phi[c] < 0.0 ? FLUID : EMPTY;
}
__global__ void classify_kernel(uint8_t* __restrict__ cell, const double* __restrict__ phi, long n) {
    classify_body(cell, phi, n, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// createWaterLevelSet without USE_LEVEL_SET (src/FluidSim2D.cpp:665-706).
// ---------------------------------------------------------------------------
// FLUID cells become EMPTY (:667-674) and the flags are copied to `out`, the
// array the bubble fill writes.
__device__ __forceinline__ void pc_clear_body(uint8_t* __restrict__ cell, uint8_t* __restrict__ out, long n, int vb,
                                              int nvb) {
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const uint8_t f = cell[c] == FLUID ? EMPTY : cell[c];
        cell[c] = f;
        out[c] = f;
    }
}
__global__ __launch_bounds__(kBlock) void pc_clear_kernel(uint8_t* __restrict__ cell, uint8_t* __restrict__ out,
                                                          long n) {
    pc_clear_body(cell, out, n, blockIdx.x, gridDim.x);
}
// Each particle marks cell ((int)x, (int)y) of its position in metres (:676-683:
// the reference does not divide by dx) in both flag arrays, if EMPTY.  Racing
// particles store the same value.  Marks outside the grid (a domain larger
// than size metres) are skipped, where the reference writes out of bounds.
__device__ __forceinline__ void pc_mark_body(const float* __restrict__ px, const float* __restrict__ py, long np,
                                             uint8_t* __restrict__ cell, uint8_t* __restrict__ out, int nx, int ny,
                                             int vb, int nvb) {
    for (long e = vb * (long)blockDim.x + threadIdx.x; e < np; e += (long)nvb * blockDim.x) {
        const int x = (int)(double)px[e], y = (int)(double)py[e];
        if (x < 0 || x >= nx || y < 0 || y >= ny) continue;
        const long c = (long)y * nx + x;
        if (cell[c] == EMPTY) {
            cell[c] = FLUID;
            out[c] = FLUID;
        }
    }
}
__global__ __launch_bounds__(kBlock) void pc_mark_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                         long np, uint8_t* __restrict__ cell,
                                                         uint8_t* __restrict__ out, int nx, int ny) {
    pc_mark_body(px, py, np, cell, out, nx, ny, blockIdx.x, gridDim.x);
}
// Bubble fill (:685-704): i = 1..nx-2 outer, j = 1..ny-2 inner, in place: an
// EMPTY cell with 4 FLUID neighbours, or 3 FLUID and 1 SOLID, becomes FLUID.
// Fills never cascade: a cell fills only if none of its neighbours is EMPTY
// when it is visited, and an EMPTY neighbour visited earlier could not have
// filled either (this cell was still EMPTY then).  So the serial in-place pass
// equals one parallel pass over the flags before it: read `cell`, write `out`.
__device__ __forceinline__ void pc_fill_body(const uint8_t* __restrict__ cell, uint8_t* __restrict__ out, int nx,
                                             int ny, int vb, int nvb) {
    const long n = (long)nx * ny;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < n; c += (long)nvb * blockDim.x) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || cell[c] != EMPTY) continue;
        const uint8_t nb[4] = {cell[c - 1], cell[c + 1], cell[c - nx], cell[c + nx]};
        int fluid = 0, solid = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            fluid += nb[k] == FLUID;
            solid += nb[k] == SOLID;
        }
        if (fluid == 4 || (fluid == 3 && solid == 1)) out[c] = FLUID;
    }
}
__global__ __launch_bounds__(kBlock) void pc_fill_kernel(const uint8_t* __restrict__ cell, uint8_t* __restrict__ out,
                                                         int nx, int ny) {
    pc_fill_body(cell, out, nx, ny, blockIdx.x, gridDim.x);
}

// Diagnostics (:709-731): volume, grid energy (Catmull-Rom velocity at cell
// corners), particle energy; FP64 sums.
__device__ __forceinline__ void diagnostics_body(const uint8_t* __restrict__ cell, MacView m,
                                                             const float* __restrict__ px,
                                                             const float* __restrict__ py,
                                                             const float* __restrict__ pvx,
                                                             const float* __restrict__ pvy, long np, double rho,
                                                             double gx, double gy, double pmass,
                                                             double* partials, int vb, int nvb) {
    const long nc = (long)m.nx * m.ny;
    const double dx = m.dx, cm = rho * dx * dx;
    double vol = 0, eg = 0, ep = 0;
    for (long c = vb * (long)blockDim.x + threadIdx.x; c < nc + np; c += (long)nvb * blockDim.x) {
        if (c < nc) {
            if (cell[c] != FLUID) continue;
            const int i = (int)(c % m.nx), j = (int)(c / m.nx);
            const double x = i * dx, y = j * dx;
            const double vx = vel_interp_u(m, x, y), vy = vel_interp_v(m, x, y);
            vol += 1.0;
            eg += 0.5 * cm * (vx * vx + vy * vy);
            eg -= cm * (gx * x + gy * y);
        } else {
            const long e = c - nc;
            const double vx = pvx[e], vy = pvy[e];
            ep += 0.5 * pmass * (vx * vx + vy * vy);
            ep -= pmass * (gx * (double)px[e] + gy * (double)py[e]);
        }
    }
    double a = block_sum(vol), b = block_sum(eg), c2 = block_sum(ep);
    if (threadIdx.x == 0) {
        partials[3 * vb + 0] = a;
        partials[3 * vb + 1] = b;
        partials[3 * vb + 2] = c2;
    }
}
__global__ __launch_bounds__(kBlock) void diagnostics_kernel(const uint8_t* __restrict__ cell, MacView m,
                                                             const float* __restrict__ px,
                                                             const float* __restrict__ py,
                                                             const float* __restrict__ pvx,
                                                             const float* __restrict__ pvy, long np, double rho,
                                                             double gx, double gy, double pmass,
                                                             double* partials) {
    diagnostics_body(cell, m, px, py, pvx, pvy, np, rho, gx, gy, pmass, partials, blockIdx.x, gridDim.x);
}

// Sums of the per-block diagnostics partials in index order (one block).
__device__ __forceinline__ void diag_finalize_body(const double* partials, int n, double dx, Diag* dg, int vb, int nvb) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = threadIdx.x; k < n; k += kBlock) {
        s0 += partials[3 * k];
        s1 += partials[3 * k + 1];
        s2 += partials[3 * k + 2];
    }
    s0 = block_sum(s0);
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        dg->volume = s0 * dx * dx;
        dg->grid_energy = s1;
        dg->particle_energy = s2;
    }
}
__global__ __launch_bounds__(kBlock) void diag_finalize_kernel(const double* partials, int n, double dx, Diag* dg) {
    diag_finalize_body(partials, n, dx, dg, blockIdx.x, gridDim.x);
}

}  // namespace fs2d
