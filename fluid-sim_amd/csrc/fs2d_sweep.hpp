// fs2d_sweep.hpp -- warp-specialised wavefront passes of the pressure solve.
//
// The MIC(0) factor and both triangular solves of applyPreconditioner
// (src/FluidSim2D.cpp:364-421) are lexicographic recurrences: cell (i, j)
// needs (i-1, j) and (i, j-1) (forward) or (i+1, j) and (i, j+1) (backward).
//
// One workgroup owns a strip of 64 rows j = 64 s + l.  Its compute wave holds
// row l in lane l and at step t processes column i = t - l: the row neighbour
// is the lane's own previous result, the column neighbour is lane l-1's
// previous result moved by one DPP wave_shr:1 (wave_shl:1 backward), and the
// strip's edge lane receives the neighbouring strip's boundary row.
//
// A step of a solve is one DPP and three FMAs; a lone wave issues them at one
// VALU instruction per 4 clocks, so the design keeps every other cost off the
// compute wave:
//   wave 0          compute: per group of 4 steps, LDS reads for the next
//                   group issued before the current one is computed, DPP/FMA
//                   chain, one dwordx4 store per output array.  It never waits
//                   on a global load.
//   waves 1..3      loaders: stream the strip's inputs (SK4 layout, 1 KB per
//                   wave-wide dwordx4 load) into an LDS ring, kBatch groups in
//                   flight per loader.
//   wave 4          poller: reads the neighbouring strip's boundary row from
//                   tagged 64-bit words ({payload, epoch}, agent-scope relaxed
//                   atomics, see fs2d_wavefront.hpp) into an LDS ring indexed
//                   by step, so a group's four values are one aligned read.
// The producing lane's boundary values go through a 2-entry LDS outbox and are
// stored by 4*NW lanes one group later.  Waves signal each other through LDS
// counters.  Strips are taken from a 64-bit ticket (strip = ticket mod S,
// launch = ticket / S, epoch = launch + 1): a strip only ever waits on a strip
// owned by a workgroup that started earlier, and no reset is needed between
// launches.  Every wait is bounded; on expiry the workgroup raises the error
// flag and all its waves run to completion.
//
// SK4 layout: element (s, t, l) at ((s * T4 + t/4) * 64 + l) * 4 + t%4 (float
// arrays; byte arrays use the same element index), T4 = ceil((nx + 64) / 4).
// Padding entries hold zero coefficients, so every lane computes every step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fs2d_wavefront.hpp"

namespace fs2d {

struct SkGeom {
    int nx, ny, S, T4;  // S strips of 64 rows, T4 groups of 4 steps per strip
    __host__ __device__ long size() const { return (long)S * T4 * 256; }
    __host__ __device__ long idx(int i, int j) const {
        const int s = j >> 6, l = j & 63, t = i + l;
        return (((long)s * T4 + (t >> 2)) * 64 + l) * 4 + (t & 3);
    }
};

__host__ inline SkGeom make_geom(int nx, int ny) {
    SkGeom g;
    g.nx = nx;
    g.ny = ny;
    g.S = (ny + 63) / 64;
    g.T4 = (nx + 64 + 3) / 4;
    return g;
}

struct SkCtl {
    unsigned long long* ticket;
    unsigned long long* handoff;  // [S][nx][NW] tagged words
    int* err;
};

// Coupling bits of a FLUID cell: the off-diagonal entries of row (i, j) are
// -scaleA towards each coupled neighbour (FluidSim2D.cpp:266-301).
enum : uint32_t { CPL_L = 1, CPL_R = 2, CPL_D = 4, CPL_U = 8 };

constexpr int kLoaders = 3;
constexpr int kBatch = 8;                  // groups in flight per loader
constexpr int kRing = 32;                  // groups in the LDS ring
constexpr int kHandRing = 256;             // steps in the hand-off ring
constexpr int kFlagEvery = 4;              // groups between consumption flags
constexpr int kSweepThreads = 64 * (kLoaders + 2);
constexpr int kLdsSpin = 1 << 26;          // bounded LDS waits (~seconds)
static_assert(kRing >= kLoaders * kBatch + 2 * kFlagEvery, "ring too small for the loaders");

__device__ __forceinline__ float f4get(const float4& v, int q) {
    return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

// LDS flags between the waves of one workgroup.  Only LDS data is published
// through them, so a workgroup-scope release/acquire (which also waits for
// every outstanding global store) is more than needed: the writer waits for
// its own LDS operations (lgkmcnt) and LDS processes each wave's operations in
// order; the compiler barriers keep LDS accesses on their side of the flag.
__device__ __forceinline__ unsigned lds_ld_acq(unsigned* p) {
    unsigned v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    asm volatile("" ::: "memory");
    return v;
}
__device__ __forceinline__ void lds_st_rel(unsigned* p, unsigned v) {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Neighbour-row value: lane l-1's (forward) or lane l+1's (backward) register,
// `old` on the edge lane.
template <bool B>
__device__ __forceinline__ uint32_t nbr(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, B ? 0x130 /* wave_shl:1 */ : 0x138 /* wave_shr:1 */,
                                                 0xf, 0xf, false);
}

// Op interface:
//   NW    32-bit words of carried state, NF float input arrays, NB (0/1) byte
//         input array `bytes`, NOUT float output arrays
//   const float* in[NF]; const uint8_t* bytes; float* out[NOUT]
//   bool skip(bool first)   wave 0 only; first: strip order 0
//   void boundary(uint32_t*) state of out-of-grid neighbours
//   template <bool B> void group(const float4* x, uint32_t b4, uint32_t* st,
//        const uint32_t (*hb)[4], float (*o)[4], uint32_t (*prod)[4])
//        4 steps; step q uses component q of x[a] and byte q of b4; st is the
//        lane's state after the previous step, updated; hb[w][q] the edge
//        lane's neighbour word w at step q; o[n][q] output n; prod[w][q] state
//        word w after step q
//   void strip_done(int s)  wave 0, after the last step
template <class Op, bool BACKWARD>
__global__ __launch_bounds__(kSweepThreads) void sweep_kernel(Op op, SkGeom g, SkCtl c) {
    constexpr int NW = Op::NW, NF = Op::NF, NB = Op::NB, NOUT = Op::NOUT;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;

    __shared__ __align__(16) float4 ring_f[kRing + 1][NF][64];  // + a spare slot for tail groups
    __shared__ uint32_t ring_b[NB ? kRing + 1 : 1][64];
    __shared__ unsigned slot_tag[kRing];
    __shared__ __align__(16) uint32_t hand[NW][kHandRing];
    __shared__ __align__(16) uint32_t outbox[2][NW][4];
    __shared__ unsigned consumed_groups, hand_ready, hand_consumed, abort_flag;
    __shared__ unsigned long long sh_ticket;
    __shared__ int sh_skip;

    if (threadIdx.x == 0) {
        sh_ticket = atomicAdd(c.ticket, 1ull);
        consumed_groups = 0;
        hand_ready = 0;
        hand_consumed = 0;
        abort_flag = 0;
    }
    if (threadIdx.x < kRing) slot_tag[threadIdx.x] = 0;
    __syncthreads();
    const int order = (int)(sh_ticket % (unsigned long long)g.S);
    const unsigned epoch = (unsigned)(sh_ticket / (unsigned long long)g.S) + 1u;
    const int s = BACKWARD ? g.S - 1 - order : order;
    if (wave == 0) {
        int sk = __hip_atomic_load(c.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        if (!sk) sk = op.skip(order == 0);
        if (lane == 0) sh_skip = sk;
    }
    __syncthreads();
    if (sh_skip) return;

    const int T4 = g.T4, nx = g.nx;
    const long strip_base = (long)s * T4 * 256;
    const bool consume = BACKWARD ? (s + 1 < g.S) : (s > 0);
    const bool produce = BACKWARD ? (s > 0) : (s + 1 < g.S);
    const int edge_lane = BACKWARD ? 63 : 0;
    const int prod_lane = BACKWARD ? 0 : 63;
    auto raise = [&]() {
        atomicExch(c.err, 1);
        atomicExch(&abort_flag, 1u);
    };
    // Hand-off values are consumed in order k = 0, 1, ...: column k forward,
    // nx-1-k backward.  Largest k among a group's four steps (-1: none).
    auto group_kmax = [&](int t0) {
        if (BACKWARD) {
            const int lo = max(t0 - 63, 0), hi = min(t0 - 60, nx - 1);
            return lo <= hi ? nx - 1 - lo : -1;
        }
        return t0 < nx ? min(t0 + 3, nx - 1) : -1;
    };

    if (wave == 0) {
        // ------------------------------------------------------------- compute
        uint32_t st[NW], bnd[NW];
        op.boundary(bnd);
#pragma unroll
        for (int w = 0; w < NW; w++) st[w] = bnd[w];
        unsigned long long* dst = c.handoff + (size_t)s * nx * NW;
        bool bad = false;

        auto fetch = [&](int gi, float4* x, uint32_t& b4, uint32_t (*hb)[4], unsigned& tag, unsigned& hr) {
            const int slot = gi % kRing;
            const int t0 = (BACKWARD ? T4 - 1 - gi : gi) * 4;
            tag = lds_ld_acq(&slot_tag[slot]);
            hr = consume ? lds_ld_acq(&hand_ready) : 0u;
#pragma unroll
            for (int a = 0; a < NF; a++) x[a] = ring_f[slot][a][lane];
            b4 = NB ? ring_b[NB ? slot : 0][lane] : 0u;
            if (consume) {
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    const uint4 h = *(const uint4*)&hand[w][t0 & (kHandRing - 1)];
                    hb[w][0] = h.x;
                    hb[w][1] = h.y;
                    hb[w][2] = h.z;
                    hb[w][3] = h.w;
                }
            }
        };
        auto ready = [&](int gi, unsigned tag, unsigned hr) {
            const int km = consume ? group_kmax((BACKWARD ? T4 - 1 - gi : gi) * 4) : -1;
            return tag == (unsigned)gi + 1u && (km < 0 || (int)hr > km);
        };
        auto wait_ready = [&](int gi) {
            for (int spin = 0; !bad; spin++) {
                const unsigned tag = lds_ld_acq(&slot_tag[gi % kRing]);
                const unsigned hr = consume ? lds_ld_acq(&hand_ready) : 0u;
                if (ready(gi, tag, hr)) break;
                if (spin >= kLdsSpin || lds_ld_acq(&abort_flag)) {
                    bad = true;
                    raise();
                }
                __builtin_amdgcn_s_sleep(0);
            }
        };

        float4 xc[NF], xn[NF];
        uint32_t bc = 0, bn = 0;
        uint32_t hbc[NW][4], hbn[NW][4];
        unsigned tg, hr;
#pragma unroll
        for (int w = 0; w < NW; w++)
#pragma unroll
            for (int q = 0; q < 4; q++) hbc[w][q] = hbn[w][q] = bnd[w];
        wait_ready(0);
        fetch(0, xc, bc, hbc, tg, hr);
        int pend_t0 = -1;  // step of the group whose boundary words sit in the outbox
        for (int gi = 0; gi < T4; gi++) {
            const int t0 = (BACKWARD ? T4 - 1 - gi : gi) * 4;
            const bool more = gi + 1 < T4;
            if (more) fetch(gi + 1, xn, bn, hbn, tg, hr);
            uint32_t ob = 0;
            if (produce && pend_t0 >= 0 && lane < 4 * NW) ob = outbox[(gi - 1) & 1][lane >> 2][lane & 3];

            // steps whose edge column lies outside the grid see the boundary
            if (consume) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int col = t0 + q - edge_lane;
                    if (col < 0 || col >= nx)
#pragma unroll
                        for (int w = 0; w < NW; w++) hbc[w][q] = bnd[w];
                }
            }
            float o[NOUT > 0 ? NOUT : 1][4];
            uint32_t prod[NW][4];
            op.template group<BACKWARD>(xc, bc, st, hbc, o, prod);
#pragma unroll
            for (int n = 0; n < NOUT; n++)
                *(float4*)(op.out[n] + strip_base + (long)t0 * 64 + lane * 4) =
                    make_float4(o[n][0], o[n][1], o[n][2], o[n][3]);
            if (produce) {
                // word (w, q) of the previous group: lane 4 w + q
                if (pend_t0 >= 0 && lane < 4 * NW) {
                    const int col = pend_t0 + (lane & 3) - prod_lane;
                    if (col >= 0 && col < nx)
                        atomic_st64(dst + (size_t)col * NW + (lane >> 2),
                                    (unsigned long long)ob | ((unsigned long long)epoch << 32));
                }
                if (lane == prod_lane)
#pragma unroll
                    for (int w = 0; w < NW; w++)
                        *(uint4*)&outbox[gi & 1][w][0] = make_uint4(prod[w][0], prod[w][1], prod[w][2], prod[w][3]);
                pend_t0 = t0;
            }
            if (lane == 0 && ((gi % kFlagEvery) == kFlagEvery - 1 || !more)) {
                lds_st_rel(&consumed_groups, (unsigned)gi + 1u);
                const int km = group_kmax(t0);
                if (consume && km >= 0) lds_st_rel(&hand_consumed, (unsigned)km + 1u);
            }
            if (more) {
                if (!ready(gi + 1, tg, hr)) {
                    wait_ready(gi + 1);
                    fetch(gi + 1, xn, bn, hbn, tg, hr);
                }
#pragma unroll
                for (int a = 0; a < NF; a++) xc[a] = xn[a];
                bc = bn;
#pragma unroll
                for (int w = 0; w < NW; w++)
#pragma unroll
                    for (int q = 0; q < 4; q++) hbc[w][q] = hbn[w][q];
            }
        }
        if (produce && pend_t0 >= 0) {
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (lane < 4 * NW) {
                const uint32_t ob = outbox[(T4 - 1) & 1][lane >> 2][lane & 3];
                const int col = pend_t0 + (lane & 3) - prod_lane;
                if (col >= 0 && col < nx)
                    atomic_st64(dst + (size_t)col * NW + (lane >> 2),
                                (unsigned long long)ob | ((unsigned long long)epoch << 32));
            }
        }
        op.strip_done(s);
    } else if (wave <= kLoaders) {
        // ------------------------------------------------------------- loaders
        // A batch of kBatch groups is loaded into registers with all loads in
        // flight, then written to the ring.  Tail groups beyond the strip go to
        // a spare slot nobody reads, so every register index is a constant.
        const int me = wave - 1;
        bool bad = false;
        for (int gb = me * kBatch; gb < T4; gb += kLoaders * kBatch) {
            const int last = min(gb + kBatch, T4) - 1;
            // slot of group gi is free once group gi - kRing has been consumed
            for (int spin = 0; !bad && (int)lds_ld_acq(&consumed_groups) < last - kRing + 1; spin++) {
                if (spin >= kLdsSpin || lds_ld_acq(&abort_flag)) {
                    bad = true;
                    raise();
                }
                __builtin_amdgcn_s_sleep(1);
            }
            float4 v[kBatch * NF];
            uint32_t b[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; k++) {
                const int gi = min(gb + k, T4 - 1);
                const long base = strip_base + (long)(BACKWARD ? T4 - 1 - gi : gi) * 256 + lane * 4;
#pragma unroll
                for (int a = 0; a < NF; a++) v[k * NF + a] = *(const float4*)(op.in[a] + base);
                b[k] = NB ? *(const uint32_t*)(op.bytes + base) : 0u;
            }
#pragma unroll
            for (int k = 0; k < kBatch; k++) {
                const int slot = gb + k < T4 ? (gb + k) % kRing : kRing;
#pragma unroll
                for (int a = 0; a < NF; a++) ring_f[slot][a][lane] = v[k * NF + a];
                if (NB) ring_b[NB ? slot : 0][lane] = b[k];
            }
            if (lane == 0)
                for (int gi = gb; gi <= last; gi++) lds_st_rel(&slot_tag[gi % kRing], (unsigned)gi + 1u);
        }
    } else if (consume) {
        // -------------------------------------------------------------- poller
        const int src_strip = BACKWARD ? s + 1 : s - 1;
        const unsigned long long* src = c.handoff + (size_t)src_strip * nx * NW;
        bool bad = false;
        for (int k0 = 0; k0 < nx && !bad; k0 += 64) {
            const int k = k0 + lane;
            const bool mine = k < nx;
            const int col = BACKWARD ? nx - 1 - k : k;
            const int slot = (col + edge_lane) & (kHandRing - 1);  // = step of that column
            // ring entries of k0 .. k0+63 were used by k - kHandRing
            for (int spin = 0; !bad && (int)lds_ld_acq(&hand_consumed) < k0 + 64 - kHandRing; spin++) {
                if (spin >= kLdsSpin || lds_ld_acq(&abort_flag)) {
                    bad = true;
                    raise();
                }
                __builtin_amdgcn_s_sleep(1);
            }
            unsigned long long hv[NW];
            bool got = !mine;
#pragma unroll
            for (int w = 0; w < NW; w++) hv[w] = 0;
            for (int spin = 0; !bad; spin++) {
                if (!got) {
                    bool ok = true;
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        if ((unsigned)(hv[w] >> 32) != epoch) hv[w] = atomic_ld64(src + (size_t)col * NW + w);
                        ok = ok && (unsigned)(hv[w] >> 32) == epoch;
                    }
                    if (ok) {
#pragma unroll
                        for (int w = 0; w < NW; w++) hand[w][slot] = (uint32_t)hv[w];
                        got = true;
                    }
                }
                // publish the ready prefix of this chunk
                const unsigned long long rdy = __ballot(got);
                const int prefix = rdy == ~0ull ? 64 : __ffsll((long long)~rdy) - 1;
                if (lane == 0) lds_st_rel(&hand_ready, (unsigned)min(k0 + prefix, nx));
                if (prefix == 64) break;
                if (spin >= kSpinLimit || lds_ld_acq(&abort_flag)) {
                    bad = true;
                    if (lane == 0) raise();
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Operators.  scaleA = dt / (rho dx^2) is the magnitude of every off-diagonal
// (FluidSim2D.cpp:261, 277, 294).
// ---------------------------------------------------------------------------

// Scalars of one projection, each written by exactly one workgroup of the
// kernel that produces it and read by later kernels.
struct PcgScalars {
    double rhs_norm, r_norm, sigma, tol;
    int iter, done, pad0, pad1;
};

__device__ __forceinline__ double wave_max_of(const double* a, int n) {
    double m = 0;
    for (int k = threadIdx.x & 63; k < n; k += 64) m = a[k] > m ? a[k] : m;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double x = __shfl_xor(m, o, 64);
        m = x > m ? x : m;
    }
    return m;
}
__device__ __forceinline__ double wave_sum_of(const double* a, int n) {
    double t = 0;
    for (int k = threadIdx.x & 63; k < n; k += 64) t += a[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    return t;
}

// MIC(0) factor (:364-388) and the solve coefficients.  State: precon (FP64)
// and the cell's coupling bits; the neighbours' Ax, Ay are -scaleA times their
// bits.  Outputs (Lt = diag(1/P) + lower(A) diag(P), M = Lt Lt^T):
//   P,  fx = scaleA c_L P P_L,  fy = scaleA c_D P P_D   (q = P r + fx q_L + fy q_D)
//       gx = scaleA c_R P^2,    gy = scaleA c_U P^2      (z = P q + gx z_R + gy z_U)
// The first workgroup also finishes |rhs|_inf and initialises the solver
// (:424-430, 447-448).
struct FactorOp {
    static constexpr int NW = 3, NF = 1, NB = 1, NOUT = 5;
    const float* in[NF];   // Adiag
    const uint8_t* bytes;  // coupling bits
    float* out[NOUT];      // P, fx, fy, gx, gy
    double tau, sigma, sA;
    const double* rhs_partials;
    int n_partials, max_iters;
    PcgScalars* sc;
    __device__ bool skip(bool first) const {
        if (first) {
            const double rn = wave_max_of(rhs_partials, n_partials);
            if ((threadIdx.x & 63) == 0) {
                sc->rhs_norm = rn;
                sc->r_norm = rn;
                sc->iter = 0;
                sc->sigma = 0;
                sc->done = (rn <= 1e-12) || (max_iters <= 0);
            }
        }
        return false;
    }
    __device__ void boundary(uint32_t* b) const { b[0] = b[1] = b[2] = 0; }
    template <bool B>
    __device__ void group(const float4* x, uint32_t b4, uint32_t* st, const uint32_t (*hb)[4], float (*o)[4],
                          uint32_t (*prod)[4]) const {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t nb[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) nb[w] = nbr<B>(hb[w][q], st[w]);
            const uint32_t f = (b4 >> (8 * q)) & 0xffu;
            const double A0 = f4get(x[0], q);
            const double PL = u2d(st), PB = u2d(nb);
            const uint32_t fL = st[2], fB = nb[2];
            // Ax(i-1,j), Ay(i-1,j), Ay(i,j-1), Ax(i,j-1)
            const double axl = (f & CPL_L) ? -sA : 0.0, ayl = (fL & CPL_U) ? -sA : 0.0;
            const double ayb = (f & CPL_D) ? -sA : 0.0, axb = (fB & CPL_R) ? -sA : 0.0;
            double Pc = 0;
            if (A0 != 0.0) {  // FLUID cells have a non-zero diagonal, all others zero
                const double t1 = axl * PL, t2 = ayb * PB;
                double e = A0 - t1 * t1 - t2 * t2 - tau * (axl * ayl * PL * PL + ayb * axb * PB * PB);
                if (e < sigma * A0) e = A0;
                Pc = 1.0 / sqrt(e);
            }
            o[0][q] = (float)Pc;
            o[1][q] = (float)(-axl * Pc * PL);
            o[2][q] = (float)(-ayb * Pc * PB);
            o[3][q] = (f & CPL_R) ? (float)(sA * Pc * Pc) : 0.0f;
            o[4][q] = (f & CPL_U) ? (float)(sA * Pc * Pc) : 0.0f;
            d2u(Pc, st);
            st[2] = f;
#pragma unroll
            for (int w = 0; w < NW; w++) prod[w][q] = st[w];
        }
    }
    __device__ void strip_done(int) const {}
};

// Forward solve q = Lt^-1 r (:400-409), q = P r + fx q_L + fy q_D.
// Accumulates q.q per strip: z.r = r.M^-1 r = |Lt^-1 r|^2 = q.q is sigma
// (:428, :457) without reading r again later.  In iteration mode the first
// workgroup finishes |r|_inf and applies the stopping test (:453) before any
// preconditioning, as the reference does.
struct FwdOp {
    static constexpr int NW = 1, NF = 4, NB = 0, NOUT = 1;
    const float* in[NF];  // r, P, fx, fy
    const uint8_t* bytes;
    float* out[NOUT];     // q
    PcgScalars* sc;
    const double* rmax_partials;
    double* qq_partials;  // [S]
    int n_partials, check;
    double acc = 0;
    __device__ bool skip(bool first) {
        if (sc->done) return true;
        if (!check) return false;
        const double rn = wave_max_of(rmax_partials, n_partials);
        const bool conv = rn <= sc->tol * sc->rhs_norm;
        if (first && (threadIdx.x & 63) == 0) {
            sc->r_norm = rn;
            if (conv) sc->done = 1;
        }
        return conv;
    }
    __device__ void boundary(uint32_t* b) const { b[0] = 0; }
    template <bool B>
    __device__ void group(const float4* x, uint32_t, uint32_t* st, const uint32_t (*hb)[4], float (*o)[4],
                          uint32_t (*prod)[4]) {
        float qp = u2f(st[0]), sq = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float qd = u2f(nbr<B>(hb[0][q], f2u(qp)));
            const float qn = fmaf(f4get(x[3], q), qd, fmaf(f4get(x[2], q), qp, f4get(x[1], q) * f4get(x[0], q)));
            o[0][q] = qn;
            prod[0][q] = f2u(qn);
            sq = fmaf(qn, qn, sq);
            qp = qn;
        }
        acc += (double)sq;
        st[0] = f2u(qp);
    }
    __device__ void strip_done(int s) {
        double t = acc;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if ((threadIdx.x & 63) == 0) qq_partials[s] = t;
    }
};

// Backward solve z = Lt^-T q (:410-419), z = P q + gx z_R + gy z_U; steps run
// q = 3, 2, 1, 0 within a group.  The first workgroup records the iteration
// count (:462).
struct BwdOp {
    static constexpr int NW = 1, NF = 4, NB = 0, NOUT = 1;
    const float* in[NF];  // q, P, gx, gy
    const uint8_t* bytes;
    float* out[NOUT];     // z
    PcgScalars* sc;
    int next_iter;        // iterations completed after this solve (0: initial solve)
    __device__ bool skip(bool first) const {
        if (sc->done) return true;
        if (first && (threadIdx.x & 63) == 0 && next_iter > 0) sc->iter = next_iter;
        return false;
    }
    __device__ void boundary(uint32_t* b) const { b[0] = 0; }
    template <bool B>
    __device__ void group(const float4* x, uint32_t, uint32_t* st, const uint32_t (*hb)[4], float (*o)[4],
                          uint32_t (*prod)[4]) const {
        float zp = u2f(st[0]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = 3 - i;
            const float zu = u2f(nbr<B>(hb[0][q], f2u(zp)));
            const float zn = fmaf(f4get(x[3], q), zu, fmaf(f4get(x[2], q), zp, f4get(x[1], q) * f4get(x[0], q)));
            o[0][q] = zn;
            prod[0][q] = f2u(zn);
            zp = zn;
        }
        st[0] = f2u(zp);
    }
    __device__ void strip_done(int) const {}
};

}  // namespace fs2d
