// fs2d_sweep.hpp -- warp-specialised wavefront passes: the pressure solve's
// recurrences and the level set's Gauss-Seidel sweeps.
//
// The MIC(0) factor and both triangular solves of applyPreconditioner
// (src/FluidSim2D.cpp:364-421) are lexicographic recurrences: cell (i, j)
// needs (i-1, j) and (i, j-1) (forward) or (i+1, j) and (i, j+1) (backward).
// So are the closest-particle and eikonal sweeps of the level set
// (src/FluidSim2D.cpp:775-792, 845-902), each a forward sweep in coordinates
// mirrored by its direction.
//
// One workgroup owns a strip of 64 rows j = 64 s + l.  Its compute wave holds
// row l in lane l and at step t processes column i = t - l: the row neighbour
// is the lane's own previous result, the column neighbour is lane l-1's
// previous result moved by one DPP wave_shr:1 (wave_shl:1 backward), and the
// strip's edge lane receives the neighbouring strip's boundary row through the
// DPP `old` operand.
//
// A solve step is one DPP and two FMAs.  A lone wave issues about one
// instruction per 4 clocks whatever its type, so the compute wave does nothing
// else it can hand to another wave:
//   wave 0          compute: supergroups of 16 steps; the LDS reads of the
//                   next supergroup are issued while the current one computes
//                   (two register buffers, no copies); readiness flags are read
//                   one supergroup ahead; one dwordx4 store per output array
//                   and 4 steps; no global loads and no waits on them.
//   waves 1..L      loaders: stream the strip's inputs (SK4 layout, 1 KB per
//                   wave-wide dwordx4 load), apply the operator's per-cell
//                   transform (e.g. P*r) and fill an LDS ring of supergroups.
//   wave L+1        poller: reads the neighbouring strip's boundary row from
//                   tagged 64-bit words ({payload, epoch}, agent-scope relaxed
//                   atomics, see fs2d_gmem.hpp) into an LDS ring indexed
//                   by step; steps whose edge column is outside the grid get
//                   the boundary value.
// The producing lane's 16 boundary values per supergroup go through a 2-entry
// LDS outbox and are stored as tagged words by 16 lanes after the first
// sub-group of the next supergroup, into a padded row (no range tests).
//
// Waves signal each other through LDS counters without s_waitcnt: LDS executes
// one wave's operations in order, so a flag written after data (or after the
// reads that free a slot) is seen only after them; compiler barriers keep the
// accesses on their side of the flag.
//
// Strips are taken from a 64-bit ticket (strip = ticket mod S, launch =
// ticket / S, epoch = launch + 1): a strip only ever waits on a strip owned by
// a workgroup that started earlier, and no reset is needed between launches.
// Every wait is bounded; on expiry the workgroup raises the error flag and all
// its waves run to completion.
//
// SK4 layout: element (s, t, l) at ((s * T4 + t/4) * 64 + l) * 4 + t%4 (float
// arrays; byte arrays use the same element index), T4 a multiple of 8 (an even
// number of supergroups) with 4 T4 >= nx + 64.  Padding entries hold zero
// coefficients, so every lane computes every step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fs2d_gmem.hpp"

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
    g.T4 = (nx + 64 + 31) / 32 * 8;
    return g;
}

struct SkCtl {
    unsigned long long* ticket;
    unsigned long long* handoff;  // per strip: [kHandPadLo + nx + kHandPadHi][NW] tagged words + 64 scratch words
    int* err;
};
// A strip's hand-off row holds columns -kHandPadLo .. nx + kHandPadHi - 1:
// every column a producing lane can name (-79 .. nx + 110, since 4 T4 <=
// nx + 95), so the stores need no range test; only columns 0 .. nx-1 are read.
constexpr int kHandPadLo = 80, kHandPadHi = 112;
__host__ __device__ inline size_t sk_hand_cols(int nx) { return (size_t)nx + kHandPadLo + kHandPadHi; }
__host__ __device__ inline size_t sk_handoff_stride(int nx, int nw) { return sk_hand_cols(nx) * nw + 64; }

// Coupling bits of a FLUID cell: the off-diagonal entries of row (i, j) are
// -scaleA towards each coupled neighbour (FluidSim2D.cpp:266-301).
enum : uint32_t { CPL_L = 1, CPL_R = 2, CPL_D = 4, CPL_U = 8 };

constexpr int kSub = 4;         // SK4 groups per supergroup (16 steps)
constexpr int kLoaders = 3;
constexpr int kBatch = 1;       // supergroups per loader batch (2 queues more loads ahead of the
                                // pollers' and costs 3% at 2048^2)
constexpr int kRing = 8;        // supergroups in the LDS ring of a single launch (a power of two)
constexpr int kHandRing = 256;  // steps in the hand-off ring
constexpr int kPollers = 2;     // poller waves: their polls interleave in time
constexpr int kSweepThreads = 64 * (1 + kLoaders + kPollers);
constexpr int kLdsSpin = 1 << 26;  // bounded LDS waits (~seconds)
// A ring of kLoaders * kBatch + 2 supergroups lets every loader keep its batch
// in flight.  A smaller ring (batched launches, where more workgroups per CU
// matter more) only makes loaders wait for free slots: a loader waits for the
// consumption of supergroups below its own, which other loaders fill without
// waiting on it, so any power of two >= 2 is deadlock-free.
static_assert(kRing >= kLoaders * kBatch + 2, "ring too small for the loaders");

__device__ __forceinline__ float f4get(const float4& v, int q) {
    return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}
__device__ __forceinline__ uint32_t u4get(const uint4& v, int q) {
    return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

__device__ __forceinline__ unsigned lds_ld(unsigned* p) {
    unsigned v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    asm volatile("" ::: "memory");
    return v;
}
__device__ __forceinline__ void lds_st(unsigned* p, unsigned v) {
    asm volatile("" ::: "memory");
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_max(unsigned* p, unsigned v) {
    asm volatile("" ::: "memory");
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned uniform(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

// Neighbour-row value: lane l-1's (forward) or lane l+1's (backward) register.
// The edge lane gets `edge` (C: a strip upstream) or zero (no strip upstream;
// every operator's boundary state is zero).
template <bool B, bool C>
__device__ __forceinline__ uint32_t nbr(uint32_t edge, uint32_t v) {
    constexpr int ctl = B ? 0x130 /* wave_shl:1 */ : 0x138 /* wave_shr:1 */;
    if (C) return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, ctl, 0xf, 0xf, false);
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, ctl, 0xf, 0xf, true);
}

// Operators whose outputs other workgroups read within the same launch
// declare kOutThrough: their stores then carry device scope (SC1), so each is
// written through to memory, and a completion flag needs only those stores'
// own completion (vmcnt) before it, not a write-back of the whole L2.
template <class Op>
constexpr auto out_through_(int) -> decltype(Op::kOutThrough) {
    return Op::kOutThrough;
}
template <class Op>
constexpr bool out_through_(long) {
    return false;
}
template <class Op>
constexpr bool out_through() {
    return out_through_<Op>(0);
}
// Operators whose inputs other workgroups of the same launch produce declare
// kWaitInputs and wait_inputs(s, err): each loader wave calls it before its
// first load of strip s.
template <class Op>
constexpr auto wait_inputs_(int) -> decltype(Op::kWaitInputs) {
    return Op::kWaitInputs;
}
template <class Op>
constexpr bool wait_inputs_(long) {
    return false;
}
template <class Op>
constexpr bool waits_inputs() {
    return wait_inputs_<Op>(0);
}
constexpr int kBufferCfg = 0x00020000;  // raw buffer resource word 3 of gfx9-family targets
constexpr int kCpolSc1 = 16;            // cache policy operand: SC1 (device scope)
constexpr unsigned kWaitVm0 = 0x0F70u;  // s_waitcnt vmcnt(0), gfx9 encoding

template <class Op, int RING = kRing>
struct SweepLds {
    static_assert(RING >= 2 && (RING & (RING - 1)) == 0, "ring slots are taken modulo RING");
    float4 ring_f[RING][kSub][Op::NF][64];
    uint32_t ring_b[Op::NB ? RING : 1][kSub][64];
    uint32_t hand[Op::NW][kHandRing];
    uint32_t outbox[2][Op::NW][16];
    unsigned slot_tag[RING];
    unsigned consumed_sg, hand_ready, abort_flag;
};

// Op interface:
//   NW    32-bit words of carried state; NIN float input arrays in[]; NF float
//         arrays in the ring; NB (0/1) byte input array `bytes`; NOUT float
//         output arrays out[]
//   bool skip(bool first)      wave 0 only; first: strip order 0
//   void xform(const float4 (&in)[NIN], float4 (&x)[NF], double& acc, int t, int s)
//         loader: ring values of one SK4 group from its inputs; t and s as
//         for group()
//   void loader_done(int s, int loader, double acc)
//   template <bool B, bool C> void group(const float4 (&x)[NF], uint32_t b4,
//        uint32_t (&st)[NW], const uint4 (&hb)[NW], float (&o)[NOUT][4],
//        uint32_t (&prod)[NW][4], int t, int s)
//         4 steps; step q uses component q of x[a], byte q of b4 and, on the
//         edge lane of a consuming strip, component q of hb[w]; st is the
//         lane's state after the previous step, updated; o[n][q] output n;
//         prod[w][q] state word w after step q; t is the group's first step
//         and s the strip, so step q of lane l is cell (t + q - l, 64 s + l)
//   void strip_done(int s)     wave 0, after the last step
template <class Op, bool B, bool C, bool P, int RING>
__device__ __forceinline__ void sweep_compute(Op& op, const SkGeom& g, SweepLds<Op, RING>& L, int s, unsigned epoch,
                                              unsigned long long* dst, int* err) {
    constexpr int NW = Op::NW, NF = Op::NF, NB = Op::NB, NOUT = Op::NOUT;
    const int lane = threadIdx.x & 63;
    const int NSG = g.T4 / kSub, nx = g.nx;
    const int T = g.T4 * 4;
    const int prod_lane = B ? 0 : 63;
    const long strip_base = (long)s * g.T4 * 256;
    constexpr bool THROUGH = out_through<Op>();
    __amdgpu_buffer_rsrc_t out_rs[NOUT];
    if (THROUGH)
#pragma unroll
        for (int n = 0; n < NOUT; n++)
            out_rs[n] = __builtin_amdgcn_make_buffer_rsrc(op.out[n], 0, (int)(g.size() * 4), kBufferCfg);

    struct Buf {
        float4 x[kSub][NF];
        uint32_t b[kSub];
        uint4 h[kSub][NW];
    };
    // first step of supergroup sg (processing order)
    auto sg_t0 = [&](int sg) { return B ? T - 16 * (sg + 1) : 16 * sg; };
    // x data of a supergroup is in the ring (loaders); hand values of its
    // steps are in the hand ring (poller)
    auto x_ready = [&](int sg, unsigned tag) { return tag == (unsigned)sg + 1u; };
    auto h_ready = [&](int sg, unsigned hr) { return hr >= 16u * (unsigned)(sg + 1); };
    auto fetch_x = [&](int sg, Buf& f) {
        const int slot = sg & (RING - 1);
#pragma unroll
        for (int j = 0; j < kSub; j++) {
#pragma unroll
            for (int a = 0; a < NF; a++) f.x[j][a] = L.ring_f[slot][j][a][lane];
            f.b[j] = NB ? L.ring_b[NB ? slot : 0][j][lane] : 0u;
        }
    };
    // LDS bases in a VGPR (mbcnt_lo(0, 0) is 0 in every lane but not known to
    // be uniform), so that the hand ring's and tags' constant offsets fold
    // into the ds_read offset fields instead of one address move per read
    const unsigned zero_v = __builtin_amdgcn_mbcnt_lo(0u, 0u);
    const uint32_t* const hand_v = &L.hand[0][0] + zero_v;
    unsigned* const tag_v = &L.slot_tag[0] + zero_v;
    // a supergroup's 16 steps sit at (t0 & 255) .. +15 of the hand ring (no wrap)
    auto fetch_h = [&](int sg, Buf& f) {
        const int h0 = sg_t0(sg) & (kHandRing - 1);
#pragma unroll
        for (int j = 0; j < kSub; j++)
#pragma unroll
            for (int w = 0; w < NW; w++)
                f.h[j][w] = C ? *(const uint4*)(hand_v + w * kHandRing + h0 + 4 * j) : make_uint4(0, 0, 0, 0);
    };
    bool bad = false;
    auto give_up = [&](int spin) {
        if (spin >= kLdsSpin || lds_ld(&L.abort_flag)) {
            bad = true;
            atomicExch(err, 1);
            atomicExch(&L.abort_flag, 1u);
        }
        __builtin_amdgcn_s_sleep(0);
    };
    auto wait_x = [&](int sg) {
        for (int spin = 0; !bad && !x_ready(sg, uniform(lds_ld(&L.slot_tag[sg & (RING - 1)]))); spin++) give_up(spin);
    };
    auto wait_h = [&](int sg) {
        for (int spin = 0; !bad && !h_ready(sg, uniform(lds_ld(&L.hand_ready))); spin++) give_up(spin);
    };

    uint32_t st[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) st[w] = 0u;
    // Hand-off stores: at supergroup sg, lane 16 w + q stores word w of step q
    // of supergroup sg-1 (at sg = 0 a column left of the grid, never read).
    // Each such lane walks 16 columns per supergroup through the padded row;
    // lanes past 16 NW keep one scratch word each.  No per-supergroup address
    // arithmetic, no range test.
    const bool hreal = lane < 16 * NW;
    const int hcol0 = sg_t0(-1) + (lane & 15) - prod_lane;
    unsigned long long* hp = hreal ? dst + (size_t)(hcol0 + kHandPadLo) * NW + (lane >> 4)
                                   : dst + sk_hand_cols(nx) * NW + lane;
    const long hinc = hreal ? (B ? -16L : 16L) * NW : 0L;

    // One supergroup: compute `cur` while the LDS reads of the next one go
    // into `nxt` speculatively: the readiness flags are read just before the
    // data (LDS executes a wave's operations in order), and checked after
    // the compute; only a failed check branches (wait, then read again).  The
    // prefetch index is clamped, so the last supergroup re-reads its own slot
    // and every iteration has the same shape.
    // LDS order per supergroup: an explicit wait at the top for all but the
    // previous supergroup's outbox and consumed_sg writes (the data of `cur`
    // arrived a supergroup ago, so it costs nothing), the outbox read, the
    // first sub-group, then all reads of the next supergroup.  The reads then
    // never exceed lgkmcnt's range (15) ahead of a use, so no wait for a
    // young read ever lands on the FMA chain.
    // s_waitcnt immediates (gfx9 encoding: lgkmcnt in bits 8-11, vmcnt and
    // expcnt at their maxima): lgkmcnt(young writes) and lgkmcnt(0)
    constexpr unsigned kYoung = P ? NW * kSub + 1 : 1;
    static_assert(kYoung <= 15, "outbox writes exceed lgkmcnt's range");
    constexpr unsigned kWaitOld = 0xC07Fu | (kYoung << 8);
    constexpr unsigned kWaitAll = 0xC07Fu;
    auto step = [&](int sg, Buf& cur, Buf& nxt) {
        const int t0 = sg_t0(sg);
        const int sn = min(sg + 1, NSG - 1);
        __builtin_amdgcn_s_waitcnt(kWaitOld);
        uint32_t ob = 0;
        if (P) ob = L.outbox[(sg - 1) & 1][min(lane >> 4, NW - 1)][lane & 15];
        __builtin_amdgcn_sched_barrier(0);
        unsigned tagv = 0, hrv = 0;

        uint32_t prod[kSub][NW][4];
#pragma unroll
        for (int k = 0; k < kSub; k++) {
            const int j = B ? kSub - 1 - k : k;
            float o[NOUT][4];
            op.template group<B, C>(cur.x[j], cur.b[j], st, cur.h[j], o, prod[j], t0 + 4 * j, s);
#pragma unroll
            for (int n = 0; n < NOUT; n++) {
                if (THROUGH) {
                    typedef float f4 __attribute__((ext_vector_type(4)));
                    const f4 v = {o[n][0], o[n][1], o[n][2], o[n][3]};
                    __builtin_amdgcn_raw_buffer_store_b128(v, out_rs[n], (j * 256 + lane * 4) * 4,
                                                           (int)((strip_base + (long)t0 * 64) * 4), kCpolSc1);
                } else {
                    float* base = op.out[n] + strip_base + (long)t0 * 64;  // uniform
                    *(float4*)(base + j * 256 + lane * 4) = make_float4(o[n][0], o[n][1], o[n][2], o[n][3]);
                }
            }
            if (k == 0) {
                // the previous supergroup's boundary values leave once the
                // outbox read has had a sub-group's time to return
                if (P) {
                    atomic_st64(hp, (unsigned long long)ob | ((unsigned long long)epoch << 32));
                    hp += hinc;
                }
                __builtin_amdgcn_sched_barrier(0);
                tagv = lds_ld(tag_v + (sn & (RING - 1)));
                if (C) hrv = lds_ld(&L.hand_ready);
                fetch_h(sn, nxt);
                fetch_x(sn, nxt);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (lane == prod_lane) {
            if (P)
#pragma unroll
                for (int w = 0; w < NW; w++)
#pragma unroll
                    for (int j = 0; j < kSub; j++)
                        *(uint4*)&L.outbox[sg & 1][w][4 * j] =
                            make_uint4(prod[j][w][0], prod[j][w][1], prod[j][w][2], prod[j][w][3]);
            lds_st(&L.consumed_sg, (unsigned)sg + 1u);
        }
        const unsigned sn1 = (unsigned)sn + 1u;
        const unsigned tv = uniform(tagv), hv = C ? uniform(hrv) : 0u;
        // nonzero unless the slot holds supergroup sn and its hand values are in
        // (hand_ready < 2^31, so the difference's sign bit is hv < 16 sn1)
        if ((tv ^ sn1) | (C ? (hv - 16u * sn1) >> 31 : 0u)) {
            if (tv != sn1) {
                wait_x(sn);
                fetch_x(sn, nxt);
            }
            if (C && hv < 16u * sn1) {
                wait_h(sn);
                fetch_h(sn, nxt);
            }
            __builtin_amdgcn_s_waitcnt(kWaitAll);  // re-read data in registers, as on the fast path
        }
    };

    Buf b0, b1;
    wait_x(0);
    if (C) wait_h(0);
    fetch_h(0, b0);
    fetch_x(0, b0);
    __builtin_amdgcn_s_waitcnt(kWaitAll);  // the first supergroup's data in registers
    for (int sg = 0; sg < NSG; sg += 2) {  // NSG is even
        step(sg, b0, b1);
        step(sg + 1, b1, b0);
    }
    if (P && hreal) {
        const uint32_t ob = L.outbox[(NSG - 1) & 1][min(lane >> 4, NW - 1)][lane & 15];
        atomic_st64(hp, (unsigned long long)ob | ((unsigned long long)epoch << 32));
    }
    op.strip_done(s);
}

// Workgroup prologue: take a ticket and clear the LDS counters.
template <class Op, int RING>
__device__ __forceinline__ unsigned long long sweep_ticket(SweepLds<Op, RING>& L, unsigned long long* ticket) {
    __shared__ unsigned long long sh_ticket;
    if (threadIdx.x == 0) {
        sh_ticket = atomicAdd(ticket, 1ull);
        L.consumed_sg = 0;
        L.hand_ready = 0;
        L.abort_flag = 0;
    }
    if (threadIdx.x < RING) L.slot_tag[threadIdx.x] = 0;
    __syncthreads();
    return sh_ticket;
}

// One strip of a pass: `order` is the strip's position in the pass (strip
// order for forward passes, reversed for backward ones), `epoch` tags its
// hand-off words.
template <class Op, bool BACKWARD, bool PREFETCH, int RING>
__device__ __forceinline__ void sweep_run(Op& op, const SkGeom& g, const SkCtl& c, int order, unsigned epoch,
                                          SweepLds<Op, RING>& L) {
    constexpr int NW = Op::NW, NIN = Op::NIN, NF = Op::NF, NB = Op::NB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __shared__ int sh_skip;
    const int s = BACKWARD ? g.S - 1 - order : order;
    const int NSG = g.T4 / kSub, T = g.T4 * 4, nx = g.nx;

    // A loader's batch of supergroups: all loads in flight together; indices
    // past the strip are clamped (their ring writes are skipped later), so
    // every register index stays a constant.
    const long strip_base = (long)s * g.T4 * 256 + lane * 4;
    float4 v[kBatch][kSub][NIN];
    uint32_t vb[kBatch][kSub];
    auto load_batch = [&](int sb) {
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int sg = min(sb + k, NSG - 1);
            const int t0 = BACKWARD ? T - 16 * (sg + 1) : 16 * sg;
#pragma unroll
            for (int j = 0; j < kSub; j++) {
                const long e = strip_base + (long)(t0 + 4 * j) * 64;
#pragma unroll
                for (int a = 0; a < NIN; a++) v[k][j][a] = *(const float4*)(op.in[a] + e);
                vb[k][j] = NB ? *(const uint32_t*)(op.bytes + e) : 0u;
            }
        }
    };
    // With PREFETCH a loader's first batch goes out before the skip test,
    // whose reads it overlaps (loads only: a skipped pass discards them).
    if (wave >= 1 && wave <= kLoaders) {
        if constexpr (waits_inputs<Op>()) op.wait_inputs(s, c.err);
        if (PREFETCH) load_batch((wave - 1) * kBatch);
    }
    if (wave == 0) {
        int sk = __hip_atomic_load(c.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        if (!sk) sk = op.skip(order == 0);
        if (lane == 0) sh_skip = sk;
    }
    __syncthreads();
    if (sh_skip) return;

    const bool consume = BACKWARD ? (s + 1 < g.S) : (s > 0);
    const bool produce = BACKWARD ? (s > 0) : (s + 1 < g.S);
    const int edge_lane = BACKWARD ? 63 : 0;
    auto raise = [&]() {
        atomicExch(c.err, 1);
        atomicExch(&L.abort_flag, 1u);
    };

    if (wave == 0) {
        // the compute wave is the critical path: issue priority over the
        // loaders and pollers that share its SIMD
        __builtin_amdgcn_s_setprio(3);
        unsigned long long* dst = c.handoff + (size_t)s * sk_handoff_stride(nx, NW);
        if (consume && produce)
            sweep_compute<Op, BACKWARD, true, true>(op, g, L, s, epoch, dst, c.err);
        else if (consume)
            sweep_compute<Op, BACKWARD, true, false>(op, g, L, s, epoch, dst, c.err);
        else if (produce)
            sweep_compute<Op, BACKWARD, false, true>(op, g, L, s, epoch, dst, c.err);
        else
            sweep_compute<Op, BACKWARD, false, false>(op, g, L, s, epoch, dst, c.err);
    } else if (wave <= kLoaders) {
        // ------------------------------------------------------------- loaders
        // Loader m takes batches of kBatch supergroups, m-th of every kLoaders
        // (load_batch above); ring writes past the strip are skipped (a
        // uniform branch).
        const int me = wave - 1;
        bool bad = false;
        double acc = 0.0;
        for (int sb = me * kBatch; sb < NSG; sb += kLoaders * kBatch) {
            const int last = min(sb + kBatch, NSG) - 1;
            // slot of supergroup sg is free once sg - RING has been consumed
            for (int spin = 0; !bad && (int)lds_ld(&L.consumed_sg) < last - RING + 1; spin++) {
                if (spin >= kLdsSpin || lds_ld(&L.abort_flag)) {
                    bad = true;
                    raise();
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!PREFETCH || sb != me * kBatch) load_batch(sb);  // with PREFETCH the first batch is loaded
#pragma unroll
            for (int k = 0; k < kBatch; k++) {
                if (sb + k < NSG) {
                    const int slot = (sb + k) & (RING - 1);
                    const int t0 = BACKWARD ? T - 16 * (sb + k + 1) : 16 * (sb + k);
#pragma unroll
                    for (int j = 0; j < kSub; j++) {
                        float4 x[NF];
                        op.xform(v[k][j], x, acc, t0 + 4 * j, s);
#pragma unroll
                        for (int a = 0; a < NF; a++) L.ring_f[slot][j][a][lane] = x[a];
                        if (NB) L.ring_b[NB ? slot : 0][j][lane] = vb[k][j];
                    }
                }
            }
            if (lane == 0)
                for (int sg = sb; sg <= last; sg++) lds_st(&L.slot_tag[sg & (RING - 1)], (unsigned)sg + 1u);
        }
        op.loader_done(s, me, acc);
    } else if (consume) {
        // ------------------------------------------------------------- pollers
        // Steps in consumption order kk = 0 .. T-1 (step t = kk forward,
        // T-1-kk backward); the edge lane reads column t - edge_lane.  The
        // pollers run the same loop, the second one half a poll behind, so a
        // hand-off word is seen about half a memory round trip sooner; both
        // write identical values and hand_ready only grows (LDS max).
        const int src_strip = BACKWARD ? s + 1 : s - 1;
        if (wave == kLoaders + 2) __builtin_amdgcn_s_sleep(4);
        const unsigned long long* src = c.handoff + (size_t)src_strip * sk_handoff_stride(nx, NW);
        bool bad = false;
        for (int k0 = 0; k0 < T && !bad; k0 += 64) {
            const int kk = k0 + lane;
            const int t = BACKWARD ? T - 1 - kk : kk;
            const int col = t - edge_lane;
            const bool in_grid = kk < T && col >= 0 && col < nx;
            // ring entries of k0 .. k0+63 were used by k - kHandRing
            for (int spin = 0; !bad && 16 * (int)lds_ld(&L.consumed_sg) < k0 + 64 - kHandRing; spin++) {
                if (spin >= kLdsSpin || lds_ld(&L.abort_flag)) {
                    bad = true;
                    raise();
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!in_grid && kk < T)
#pragma unroll
                for (int w = 0; w < NW; w++) L.hand[w][t & (kHandRing - 1)] = 0u;  // boundary state
            unsigned long long hv[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) hv[w] = 0;
            bool got = !in_grid;
            for (int spin = 0; !bad; spin++) {
                if (!got) {
                    bool ok = true;
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        if ((unsigned)(hv[w] >> 32) != epoch)
                            hv[w] = atomic_ld64(src + (size_t)(col + kHandPadLo) * NW + w);
                        ok = ok && (unsigned)(hv[w] >> 32) == epoch;
                    }
                    if (ok) {
#pragma unroll
                        for (int w = 0; w < NW; w++) L.hand[w][t & (kHandRing - 1)] = (uint32_t)hv[w];
                        got = true;
                    }
                }
                // publish the ready prefix of this chunk
                const unsigned long long rdy = __ballot(got);
                const int prefix = rdy == ~0ull ? 64 : __ffsll((long long)~rdy) - 1;
                if (lane == 0) lds_max(&L.hand_ready, (unsigned)min(k0 + prefix, T));
                if (prefix == 64) break;
                if (spin >= kSpinLimit || lds_ld(&L.abort_flag)) {
                    bad = true;
                    if (lane == 0) raise();
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// Supergroups in a single launch's ring: kRing, or 4 for operators with more
// than 3 ring arrays (8 slots of 5 arrays would need 160 KB of LDS).
template <class Op>
constexpr int sweep_ring() {
    return Op::NF > 3 ? 4 : kRing;
}

template <class Op, bool BACKWARD>
__global__ __launch_bounds__(kSweepThreads) void sweep_kernel(Op op, SkGeom g, SkCtl c) {
    __shared__ __align__(16) SweepLds<Op, sweep_ring<Op>()> L;
    const unsigned long long ticket = sweep_ticket(L, c.ticket);
    const int order = (int)(ticket % (unsigned long long)g.S);
    const unsigned epoch = (unsigned)(ticket / (unsigned long long)g.S) + 1u;
    sweep_run<Op, BACKWARD, true>(op, g, c, order, epoch, L);
}

// ---------------------------------------------------------------------------
// Operators.  scaleA = dt / (rho dx^2) is the magnitude of every off-diagonal
// (FluidSim2D.cpp:261, 277, 294).
// ---------------------------------------------------------------------------

// Scalars of one projection, each written by exactly one workgroup of the
// kernel that produces it and read by later kernels.
struct PcgScalars {
    double rhs_norm, r_norm, sigma, tol;
    double ws;  // w.s of the current iteration when a separate reduction formed it
    int iter, done;
    float beta;  // beta of the current iteration when a separate reduction formed it
    int pad1;
};

__device__ __forceinline__ double wave_max_of(const double* a, int n) {
    // lane k takes a[k], a[k + 64], ... in that order, with 8 loads in flight
    double m = 0;
    for (int k0 = threadIdx.x & 63; k0 < n; k0 += 64 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = k0 + 64 * u < n ? a[k0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (k0 + 64 * u < n) m = v[u] > m ? v[u] : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double x = __shfl_xor(m, o, 64);
        m = x > m ? x : m;
    }
    return m;
}
// wave_max_of with every load of a lane in flight at once (n <= 64 W): one
// memory round trip instead of n / 512.
template <int W>
__device__ __forceinline__ double wave_max_of_wide(const double* a, int n) {
    const int lane = threadIdx.x & 63;
    double v[W];
#pragma unroll
    for (int u = 0; u < W; u++) v[u] = lane + 64 * u < n ? a[lane + 64 * u] : 0.0;
    double m = 0;
#pragma unroll
    for (int u = 0; u < W; u++) m = v[u] > m ? v[u] : m;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double x = __shfl_xor(m, o, 64);
        m = x > m ? x : m;
    }
    return m;
}
// Two wave_sum_of sums with every load of both in flight together (n1 <=
// 64 W1, n2 <= 64 W2): one memory round trip instead of several; the same
// additions in the same order.
template <int W1, int W2>
__device__ __forceinline__ void wave_sum_of_wide2(const double* a, int n1, const double* b, int n2, double& sa,
                                                  double& sb) {
    const int lane = threadIdx.x & 63;
    double va[W1], vb[W2];
#pragma unroll
    for (int u = 0; u < W1; u++) va[u] = lane + 64 * u < n1 ? a[lane + 64 * u] : 0.0;
#pragma unroll
    for (int u = 0; u < W2; u++) vb[u] = lane + 64 * u < n2 ? b[lane + 64 * u] : 0.0;
    double ta = 0, tb = 0;
#pragma unroll
    for (int u = 0; u < W1; u++)
        if (lane + 64 * u < n1) ta += va[u];
#pragma unroll
    for (int u = 0; u < W2; u++)
        if (lane + 64 * u < n2) tb += vb[u];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ta += __shfl_xor(ta, o, 64);
        tb += __shfl_xor(tb, o, 64);
    }
    sa = ta;
    sb = tb;
}
__device__ __forceinline__ double wave_sum_of(const double* a, int n) {
    double t = 0;
    for (int k0 = threadIdx.x & 63; k0 < n; k0 += 64 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = k0 + 64 * u < n ? a[k0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (k0 + 64 * u < n) t += v[u];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    return t;
}

// MIC(0) factor (:364-388) and the solve coefficients, in solver units (the
// matrix divided by scaleA: the neighbours' Ax, Ay are -1 times their coupling
// bits).  State: precon (FP64) and the cell's coupling bits.  Outputs (Lt =
// diag(1/P) + lower(A) diag(P), M = Lt Lt^T):
//   P,  fx = c_L P P_L,  fy = c_D P P_D   (q = P r + fx q_L + fy q_D)
//       gx = c_R P^2,    gy = c_U P^2      (z = P q + gx z_R + gy z_U)
// Products are rounded one by one, as the reference's expression reads
// (no contraction, the same in every launch path).  The first workgroup also
// finishes |rhs|_inf and initialises the solver (:424-430, 447-448).
struct FactorOp {
    static constexpr int NW = 3, NIN = 1, NF = 1, NB = 1, NOUT = 5;
    const float* in[NIN];  // Adiag
    const uint8_t* bytes;  // coupling bits
    float* out[NOUT];      // P, fx, fy, gx, gy
    double tau, sigma;
    const double* rhs_partials;
    int n_partials, max_iters;
    PcgScalars* sc;
    double rhs_zero;  // |rhs|_inf at or below which no iteration runs (1e-12, :448, in the rhs' units)
    __device__ bool skip(bool first) const {
        if (first) {
            const double rn = wave_max_of(rhs_partials, n_partials);
            if ((threadIdx.x & 63) == 0) {
                sc->rhs_norm = rn;
                sc->r_norm = rn;
                sc->iter = 0;
                sc->sigma = 0;
                sc->done = (rn <= rhs_zero) || (max_iters <= 0);
            }
        }
        return false;
    }
    __device__ void xform(const float4 (&in)[NIN], float4 (&x)[NF], double&, int, int) const { x[0] = in[0]; }
    __device__ void loader_done(int, int, double) const {}
    template <bool B, bool C>
    __device__ void group(const float4 (&x)[NF], uint32_t b4, uint32_t (&st)[NW], const uint4 (&hb)[NW],
                          float (&o)[NOUT][4], uint32_t (&prod)[NW][4], int, int) const {
#pragma clang fp contract(off)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t nb[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) nb[w] = nbr<B, C>(u4get(hb[w], q), st[w]);
            const uint32_t f = (b4 >> (8 * q)) & 0xffu;
            const double A0 = f4get(x[0], q);
            const double PL = u2d(st), PB = u2d(nb);
            const uint32_t fL = st[2], fB = nb[2];
            // Ax(i-1,j), Ay(i-1,j), Ay(i,j-1), Ax(i,j-1)
            const double axl = (f & CPL_L) ? -1.0 : 0.0, ayl = (fL & CPL_U) ? -1.0 : 0.0;
            const double ayb = (f & CPL_D) ? -1.0 : 0.0, axb = (fB & CPL_R) ? -1.0 : 0.0;
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
            o[3][q] = (f & CPL_R) ? (float)(Pc * Pc) : 0.0f;
            o[4][q] = (f & CPL_U) ? (float)(Pc * Pc) : 0.0f;
            d2u(Pc, st);
            st[2] = f;
#pragma unroll
            for (int w = 0; w < NW; w++) prod[w][q] = st[w];
        }
    }
    __device__ void strip_done(int) const {}
};

// Forward solve q = Lt^-1 r (:400-409), q = P r + fx q_L + fy q_D; the loaders
// form P r.
//   FUSED false: in iteration mode the first workgroup finishes |r|_inf and
//     applies the stopping test (:453) before any preconditioning, as the
//     reference does.
//   FUSED true: the solve's strips inside fwd_axpy_kernel, after the same
//     launch's p/r update.  Strip s's loaders first wait for the update's runs
//     over strip s (run_flags[vb] = the launch's epoch, published after the
//     runs' device-scope r stores); the stopping test moves to the backward
//     solve's launch, so the last solve of a projection runs once more than
//     needed, writing only q.
template <bool FUSED>
struct FwdOpT {
    static constexpr int NW = 1, NIN = 4, NF = 3, NB = 0, NOUT = 1;
    static constexpr bool kWaitInputs = FUSED;
    const float* in[NIN];  // r, P, fx, fy
    const uint8_t* bytes;
    float* out[NOUT];      // q
    PcgScalars* sc;
    const double* rmax_partials;
    int n_partials, check;
    const unsigned* run_flags;  // FUSED: [n_partials] p/r update runs
    SkGeom geom;                // FUSED: the runs' strips
    unsigned epoch;
    __device__ bool skip(bool first) {
        if (sc->done) return true;
        if (FUSED || !check) return false;
        const double rn = wave_max_of(rmax_partials, n_partials);
        const bool conv = rn <= sc->tol * sc->rhs_norm;
        if (first && (threadIdx.x & 63) == 0) {
            sc->r_norm = rn;
            if (conv) sc->done = 1;
        }
        return conv;
    }
    // Runs of the p/r update are whole 256-entry groups, the n / 256 groups
    // split evenly over n_partials runs (pcg_run); strip s is groups
    // s T4 .. (s+1) T4 - 1.
    __device__ void wait_inputs(int s, int* err) const {
        if (!FUSED || sc->done) return;
        const long ng = geom.size() >> 8, nr = n_partials;
        const long g0 = (long)s * geom.T4, g1 = g0 + geom.T4;  // the strip's groups
        // runs vb with ng vb / nr < g1 and ng (vb + 1) / nr > g0
        const int lo = (int)(g0 * nr / ng), hi = (int)min(nr - 1, (g1 * nr + ng - 1) / ng);
        const int lane = threadIdx.x & 63;
        for (int v0 = lo; v0 <= hi; v0 += 64) {
            const int vb = v0 + lane;
            for (int spin = 0;; spin++) {
                const bool ok = vb > hi || __hip_atomic_load(run_flags + vb, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT) == epoch;
                if (__all(ok)) break;
                if (spin >= kSpinLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    if (lane == 0) atomicExch(err, 1);
                    return;
                }
                __builtin_amdgcn_s_sleep(4);
            }
        }
    }
    __device__ void xform(const float4 (&in)[NIN], float4 (&x)[NF], double&, int, int) const {
        x[0] = make_float4(in[1].x * in[0].x, in[1].y * in[0].y, in[1].z * in[0].z, in[1].w * in[0].w);
        x[1] = in[2];
        x[2] = in[3];
    }
    __device__ void loader_done(int, int, double) const {}
    template <bool B, bool C>
    __device__ void group(const float4 (&x)[NF], uint32_t, uint32_t (&st)[NW], const uint4 (&hb)[NW],
                          float (&o)[NOUT][4], uint32_t (&prod)[NW][4], int, int) const {
        float qp = u2f(st[0]);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // the row term first: the DPP then reads qp one instruction later
            const float t1 = fmaf(f4get(x[1], q), qp, f4get(x[0], q));
            const float qd = u2f(nbr<B, C>(u4get(hb[0], q), f2u(qp)));
            const float qn = fmaf(f4get(x[2], q), qd, t1);
            o[0][q] = qn;
            prod[0][q] = f2u(qn);
            qp = qn;
        }
        st[0] = f2u(qp);
    }
    __device__ void strip_done(int) const {}
};
using FwdOp = FwdOpT<false>;
using FwdFusedOp = FwdOpT<true>;

// Backward solve z = Lt^-T q (:410-419), z = P q + gx z_R + gy z_U; steps run
// q = 3, 2, 1, 0 within a group; the loaders form P q.  The first workgroup
// records the iteration count (:462).
//   FUSED false: the loaders also sum q.q per strip: z.r = r.M^-1 r =
//     |Lt^-1 r|^2 = q.q is sigma (:428, :457), into qq_partials[S][kLoaders].
//   FUSED true: the solve's strips inside bwd_spmv_kernel; sigma is summed by
//     that launch's own workgroups (in the loaders' order), and each strip's
//     compute wave publishes its completion for the SpMV workgroups of the
//     same launch: z is stored with device scope (kOutThrough), so once those
//     stores have completed (vmcnt 0) the flag may follow.
__device__ __forceinline__ double qq4(const float4& q) {
    return (double)fmaf(q.w, q.w, fmaf(q.z, q.z, fmaf(q.y, q.y, q.x * q.x)));
}
template <bool FUSED>
struct BwdOpT {
    static constexpr int NW = 1, NIN = 4, NF = 3, NB = 0, NOUT = 1;
    static constexpr bool kOutThrough = FUSED;
    const float* in[NIN];  // q, P, gx, gy
    const uint8_t* bytes;
    float* out[NOUT];      // z
    PcgScalars* sc;
    double* qq_partials;   // [S][kLoaders] (FUSED false)
    int next_iter;         // iterations completed after this solve (0: initial solve)
    unsigned* strip_flags; // [S] (FUSED true): the launch's epoch once the strip's z is stored
    unsigned epoch;
    int check;                    // FUSED false: the iteration's stopping test (:453) first, as
                                  // FwdOp's, when the p/r update ran with the forward solves
                                  // (b_fwd_axpy_kernel)
    const double* rmax_partials;  // |r|_inf partials of the iteration's p/r update
    int n_partials;
    __device__ bool skip(bool first) const {
        if (sc->done) return true;
        // fused: the launch's SpMV workgroups apply the stopping test and
        // record the count; the strips run regardless (the solve of the
        // iteration that converges writes only z and sigma, which no later
        // launch reads)
        if (FUSED) return false;
        if (check) {
            const double rn = n_partials <= 64 * 32 ? wave_max_of_wide<32>(rmax_partials, n_partials)
                                                    : wave_max_of(rmax_partials, n_partials);
            const bool conv = rn <= sc->tol * sc->rhs_norm;
            if (first && (threadIdx.x & 63) == 0) {
                sc->r_norm = rn;
                if (conv) sc->done = 1;
            }
            if (conv) return true;
        }
        if (first && (threadIdx.x & 63) == 0 && next_iter > 0) sc->iter = next_iter;
        return false;
    }
    __device__ void xform(const float4 (&in)[NIN], float4 (&x)[NF], double& acc, int, int) const {
        const float4 q = in[0], P = in[1];
        x[0] = make_float4(P.x * q.x, P.y * q.y, P.z * q.z, P.w * q.w);
        x[1] = in[2];
        x[2] = in[3];
        if (!FUSED) acc += qq4(q);
    }
    __device__ void loader_done(int s, int me, double acc) const {
        if (FUSED) return;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((threadIdx.x & 63) == 0) qq_partials[s * kLoaders + me] = acc;
    }
    template <bool B, bool C>
    __device__ void group(const float4 (&x)[NF], uint32_t, uint32_t (&st)[NW], const uint4 (&hb)[NW],
                          float (&o)[NOUT][4], uint32_t (&prod)[NW][4], int, int) const {
        float zp = u2f(st[0]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = 3 - i;
            const float t1 = fmaf(f4get(x[1], q), zp, f4get(x[0], q));
            const float zu = u2f(nbr<B, C>(u4get(hb[0], q), f2u(zp)));
            const float zn = fmaf(f4get(x[2], q), zu, t1);
            o[0][q] = zn;
            prod[0][q] = f2u(zn);
            zp = zn;
        }
        st[0] = f2u(zp);
    }
    __device__ void strip_done(int s) const {
        if (!FUSED) return;
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(kWaitVm0);  // every z store of the strip is in memory
        asm volatile("" ::: "memory");
        if ((threadIdx.x & 63) == 0)
            __hip_atomic_store(strip_flags + s, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};
using BwdOp = BwdOpT<false>;
using BwdFusedOp = BwdOpT<true>;

}  // namespace fs2d
