// fs2d_wavefront.hpp -- skewed-strip wavefront engine for 2D lexicographic
// recurrences on gfx950.
//
// A recurrence of the form  X(a, b) = F(X(a-1, b), X(a, b-1), inputs(a, b))
// over an NA x NB grid has anti-diagonal a+b as its only parallel axis.  Four
// passes of the reference have this shape: the MIC(0) factor and the forward /
// backward triangular solves of the preconditioner (src/FluidSim2D.cpp:364-421),
// and the Gauss-Seidel sweeps of the level set (include/FluidSim2D.h:178-203,
// src/FluidSim2D.cpp:775-792, 845-902).  Sweeps in the other three directions
// are this canonical one with the axes mirrored by the operator.
//
// Mapping: one wave64 owns a strip of 64 consecutive rows b = 64*s + lane.  At
// step t lane l processes column a = t - l, so each lane walks its own row
// left to right one step behind the lane below it:
//   * X(a-1, b) is the lane's own result of step t-1 (a register);
//   * X(a, b-1) is lane l-1's result of step t-1, moved by one DPP
//     `wave_shr:1` (no LDS, no barrier);
//   * lane 0 gets X(a, 64*s-1) from the strip below through a hand-off row in
//     global memory.  Each entry is one 64-bit word {payload, epoch} written
//     and read with agent-scope relaxed atomics, so a consumer that sees the
//     current epoch also sees the payload; no fences are needed.
// Strips are taken in order from an atomic ticket, so strip s-1 is always
// owned by a workgroup that started earlier and never waits on strip s: the
// kernel cannot deadlock however many workgroups are resident.  Every wait is
// bounded (kSpinLimit); on expiry the kernel raises an error flag and runs to
// completion, so every wave exits.  The last workgroup to finish resets the
// ticket and advances the epoch, which keeps the launch free of host-side
// per-launch arguments (replayable from a graph).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fs2d {

constexpr int kSpinLimit = 1 << 16;  // polls of one memory round trip each: well under 1 s

// Pointers typed as global memory.  Accesses through them compile to GLOBAL
// instructions, counted by vmcnt alone.  A pointer read from memory (the
// batched kernels' descriptors) is otherwise generic, and FLAT accesses also
// count in lgkmcnt, so every LDS or scalar-load wait would wait for them too.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__host__ __device__ inline gptr<T> G(T* p) {
    return (gptr<T>)p;
}
// int4 / float4 layouts as builtin vectors, which load and store through gptr
// (the HIP vector classes' copy operators expect generic pointers)
typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// ctl[0] ticket, ctl[1] finished workgroups, ctl[2] epoch (starts at 1).
struct WaveCtl {
    unsigned long long* handoff;  // [strips][NA][NW]
    unsigned* ctl;
    int* err;
};

__device__ __forceinline__ unsigned long long atomic_ld64(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_ld64(gptr<const unsigned long long> p) {
    return __hip_atomic_load(const_cast<gptr<unsigned long long>>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st64(gptr<unsigned long long> p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned readlane_u32(unsigned v, int lane) {
    return (unsigned)__builtin_amdgcn_readlane((int)v, lane);
}

// Op interface:
//   static constexpr int NW;                       32-bit words of carried state
//   bool skip() const;                             whole pass is a no-op
//   void boundary(uint32_t s[NW]) const;           state of out-of-grid neighbours
//   struct Rec;                                    a cell's inputs
//   Rec load(int a, int b, bool ok) const;         loads of (a, b) (ok: in the grid; else a
//                                                  branch-free load of a clamped address)
//   void compute(int a, int b, const Rec& r, const uint32_t left[NW],
//                const uint32_t below[NW], uint32_t out[NW]);
//                                                  new state of (a, b); stores results
//   void strip_done(int strip);                    once per strip, whole wave
//   void last_done(int nstrips);                   by the last workgroup, whole wave
// A cell's inputs never depend on the carried state, so the engine loads them
// kWavePrefetch steps before compute() needs them, in a register ring whose
// slot k is loaded and consumed at steps of the same residue (no copies): a
// step waits only for its own record and the DPP of its neighbour's result.
constexpr int kWavePrefetch = 8;
// Workgroup prologue: ticket, epoch and the sticky error flag.
struct WaveTicket {
    unsigned ticket, epoch;
    int err;
};
__device__ __forceinline__ WaveTicket wave_ticket(const WaveCtl& c) {
    __shared__ unsigned sh_strip, sh_epoch;
    __shared__ int sh_err;
    if (threadIdx.x == 0) {
        sh_strip = atomicAdd(&c.ctl[0], 1u);
        sh_epoch = __hip_atomic_load(&c.ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh_err = __hip_atomic_load(c.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return WaveTicket{sh_strip, sh_epoch, sh_err};
}

// Strip s of an NA x NB pass; hand-off rows of strip k at handoff + k NA NW.
// Lane 0 needs the row below's value of column t at step t; it loads that
// hand-off word kWavePrefetch steps ahead with the cells' inputs (so no wait
// ever drains the prefetch stream) and re-polls only a word whose epoch is not
// yet current.  Strip 0 reads its own (never consumed) row instead: the
// address stays valid and the value is unused.
template <class Op>
__device__ __forceinline__ void wavefront_strip(Op& op, int NA, int NB, int s, unsigned epoch,
                                                unsigned long long* handoff, int* err) {
    constexpr int NW = Op::NW;
    constexpr int D = kWavePrefetch;
    const int lane = threadIdx.x;
    const int nstrips = (NB + 63) >> 6;
    const int b = s * 64 + lane;
    const int rows = min(64, NB - s * 64);
    const bool row_ok = lane < rows;
    uint32_t bnd[NW], left[NW], out[NW];
    op.boundary(bnd);
#pragma unroll
    for (int w = 0; w < NW; w++) left[w] = out[w] = bnd[w];

    const gptr<const unsigned long long> src = G(handoff) + (size_t)(s > 0 ? s - 1 : 0) * NA * NW;
    const gptr<unsigned long long> dst = G(handoff) + (size_t)s * NA * NW;
    const bool produce = (s + 1 < nstrips);
    const bool consume = s > 0;
    bool failed = false;
    auto in_grid = [&](int a) { return a >= 0 && a < NA && row_ok; };
    // Every lane loads (a clamped, in-bounds address) and lane 0's words are
    // the ones used: branch-free loads keep the compiler's wait counts exact.
    auto hand = [&](int col, unsigned long long (&h)[NW]) {
        const int c = min(col, NA - 1);
#pragma unroll
        for (int w = 0; w < NW; w++) h[w] = atomic_ld64(src + (size_t)c * NW + w);
    };

    const int T = NA + rows - 1;
    typename Op::Rec rq[D];
    unsigned long long hq[D][NW];
#pragma unroll
    for (int k = 0; k < D; k++) {
        rq[k] = op.load(k - lane, b, in_grid(k - lane));
        hand(k, hq[k]);
    }
    for (int t0 = 0; t0 < T; t0 += D) {
#pragma unroll
        for (int kq = 0; kq < D; kq++) {
            const int t = t0 + kq;
            if (t >= T) break;  // uniform
            const int a = t - lane;
            const typename Op::Rec cur = rq[kq];
            unsigned long long hw[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) hw[w] = hq[kq][w];
            rq[kq] = op.load(t + D - lane, b, in_grid(t + D - lane));  // step t + D
            hand(t + D, hq[kq]);

            uint32_t bl0[NW];
            if (!consume || t >= NA) {
#pragma unroll
                for (int w = 0; w < NW; w++) bl0[w] = bnd[w];
            } else {
                // check first: the re-poll loop (whose entry waits for every
                // outstanding load) runs only when the word is not yet current
                auto current = [&]() {
                    bool ok = true;
#pragma unroll
                    for (int w = 0; w < NW; w++) ok = ok && (readlane_u32((unsigned)(hw[w] >> 32), 0) == epoch);
                    return ok;
                };
                if (!current() && !failed) {
                    int spin = 0;
                    do {
                        if (spin++ >= kSpinLimit) {
                            failed = true;
                            if (lane == 0) atomicExch(err, 1);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
#pragma unroll
                        for (int w = 0; w < NW; w++)
                            if (lane == 0) hw[w] = atomic_ld64(src + (size_t)t * NW + w);
                    } while (!current());
                }
#pragma unroll
                for (int w = 0; w < NW; w++) bl0[w] = (uint32_t)hw[w];
            }
            uint32_t below[NW];
#pragma unroll
            for (int w = 0; w < NW; w++)
                below[w] = (uint32_t)__builtin_amdgcn_update_dpp((int)bl0[w], (int)out[w], 0x138 /* wave_shr:1 */,
                                                                 0xf, 0xf, false);
            if (in_grid(a)) {
                op.compute(a, b, cur, left, below, out);
#pragma unroll
                for (int w = 0; w < NW; w++) left[w] = out[w];
                if (produce && lane == 63) {
#pragma unroll
                    for (int w = 0; w < NW; w++)
                        atomic_st64(dst + (size_t)a * NW + w,
                                    (unsigned long long)out[w] | ((unsigned long long)epoch << 32));
                }
            }
        }
    }
    op.strip_done(s);
}

// Completion protocol: the last of `total` workgroups resets the ticket and
// advances the epoch.  Returns true in that workgroup.
__device__ __forceinline__ bool wave_finish(const WaveCtl& c, unsigned total, unsigned epoch) {
    __syncthreads();
    __threadfence();
    __shared__ unsigned sh_last;
    if (threadIdx.x == 0) {
        unsigned done = atomicAdd(&c.ctl[1], 1u);
        sh_last = (done == total - 1u);
    }
    __syncthreads();
    if (sh_last) {
        __threadfence();
        if (threadIdx.x == 0) {
            unsigned next = epoch + 1u;
            if (next == 0u) next = 1u;
            c.ctl[0] = 0u;
            c.ctl[1] = 0u;
            __hip_atomic_store(&c.ctl[2], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
        }
    }
    return sh_last != 0;
}

template <class Op>
__global__ __launch_bounds__(64) void wavefront_kernel(Op op, int NA, int NB, WaveCtl c) {
    const int nstrips = (NB + 63) >> 6;
    const WaveTicket tk = wave_ticket(c);
    // After an expired wait (sticky error) later launches do no work at all.
    const bool skip = op.skip() || tk.err != 0;
    if (!skip && (int)tk.ticket < nstrips) wavefront_strip(op, NA, NB, (int)tk.ticket, tk.epoch, c.handoff, c.err);
    if (wave_finish(c, (unsigned)nstrips, tk.epoch) && !skip) op.last_done(nstrips);
}

// Bit casts for carried state.
__device__ __forceinline__ uint32_t f2u(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float u2f(uint32_t x) { return __uint_as_float(x); }
__device__ __forceinline__ void d2u(double x, uint32_t* w) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    w[0] = (uint32_t)b;
    w[1] = (uint32_t)(b >> 32);
}
__device__ __forceinline__ double u2d(const uint32_t* w) {
    return __longlong_as_double((long long)(((unsigned long long)w[1] << 32) | w[0]));
}

}  // namespace fs2d
