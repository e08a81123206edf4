// fs2d_render.hpp -- the demo renderer's per-frame buffers built on the device
// (fs2d_render_buffers of include/fs2d.h).  Included at the end of
// fs2d_capi.hip (it uses the handle and its helpers).
//
// FluidRenderer2D::updateBuffers (demo/FluidRenderer2D.cpp:435-486) walks the
// simulation on the host every frame: FLUID and SOLID cell corners, a
// velocity line per cell centre, the cells with non-zero pressure and their
// sigmoid, a velocity line per particle and the sigmoid of phi, pushed in
// row-major order (FluidSim2D::iterate, include/FluidSim2D.h:153-159).  Here
// the same lists are written straight into device buffers (a renderer's
// mapped vertex buffers) from the handle's device state:
//   - the three cell lists are stream compactions in row-major order: a count
//     per 256-cell block, one workgroup's scan of the block counts, then each
//     block's own prefix of its flags (ballot + popcount) for the positions;
//   - the velocity lines sample the MAC grid with the advection's Catmull-Rom
//     velInterp (vel_interp_u / vel_interp_v);
//   - positions are formed in FP64 and rounded to FP32, as the reference's
//     (float)(i * dx); particle lines in FP32, as (float)x + (float)dt * (float)vel.
// aml::sigmoid (the un-vendored altmath library) is taken as the logistic
// 1 / (1 + e^-x).
#pragma once

namespace fs2d {

constexpr int kRenderBlock = 256;

__device__ __forceinline__ float logistic(float x) { return 1.0f / (1.0f + expf(-x)); }

// Exclusive prefix of a flag over a 256-thread block (row-major order) and
// the block's total.
__device__ __forceinline__ int render_flag_scan(bool flag, int* wsum, int& total) {
    const unsigned long long m = __ballot(flag);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();  // earlier readers of wsum are done
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kRenderBlock / 64; w++) {
        if (w < wave) before += wsum[w];
        total += wsum[w];
    }
    return before + in_wave;
}

// Per block: FLUID, SOLID and p != 0 counts, at cnt[0 nb + b], [nb + b], [2 nb + b].
__global__ __launch_bounds__(kRenderBlock) void render_count_kernel(const uint8_t* __restrict__ cell,
                                                                    const float* __restrict__ p, long n,
                                                                    int* __restrict__ cnt, int nb) {
    __shared__ int wsum[kRenderBlock / 64];
    const long c = (long)blockIdx.x * kRenderBlock + threadIdx.x;
    const uint8_t t = c < n ? cell[c] : (uint8_t)EMPTY;
    const bool hp = c < n && p[c] != 0.0f;
    int tf, ts, tp;
    (void)render_flag_scan(t == FLUID, wsum, tf);
    (void)render_flag_scan(t == SOLID, wsum, ts);
    (void)render_flag_scan(hp, wsum, tp);
    if (threadIdx.x == 0) {
        cnt[blockIdx.x] = tf;
        cnt[nb + blockIdx.x] = ts;
        cnt[2 * nb + blockIdx.x] = tp;
    }
}

// One workgroup: the block counts become exclusive offsets (in place), the
// three totals go to tot[0..2].
__global__ __launch_bounds__(kBlock) void render_scan_kernel(int* __restrict__ cnt, int nb, int* __restrict__ tot) {
    for (int a = 0; a < 3; a++) {
        int* x = cnt + (long)a * nb;
        int carry = 0;
        for (int base = 0; base < nb; base += kBlock) {
            const int i = base + threadIdx.x;
            const int v = i < nb ? x[i] : 0;
            int total;
            const int ex = block_excl_scan(v, total);
            if (i < nb) x[i] = carry + ex;
            carry += total;
        }
        if (threadIdx.x == 0) tot[a] = carry;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRenderBlock) void render_scatter_kernel(
    const uint8_t* __restrict__ cell, const float* __restrict__ p, int nx, long n, double dx,
    const int* __restrict__ off, int nb, float2* __restrict__ water, float2* __restrict__ solid,
    float2* __restrict__ pcell, float* __restrict__ pval) {
    __shared__ int wsum[kRenderBlock / 64];
    const long c = (long)blockIdx.x * kRenderBlock + threadIdx.x;
    const uint8_t t = c < n ? cell[c] : (uint8_t)EMPTY;
    const float pc = c < n ? p[c] : 0.0f;
    const bool hp = c < n && pc != 0.0f;
    int unused;
    const int kf = render_flag_scan(t == FLUID, wsum, unused);
    const int ks = render_flag_scan(t == SOLID, wsum, unused);
    const int kp = render_flag_scan(hp, wsum, unused);
    if (c >= n) return;
    const long i = c % nx, j = c / nx;
    const float2 pos = make_float2((float)((double)i * dx), (float)((double)j * dx));
    if (t == FLUID && water) water[off[blockIdx.x] + kf] = pos;
    if (t == SOLID && solid) solid[off[nb + blockIdx.x] + ks] = pos;
    if (hp) {
        const long k = off[2 * nb + blockIdx.x] + kp;
        if (pcell) pcell[k] = pos;
        if (pval) pval[k] = logistic(0.01f * pc);
    }
}

// Cell-centre velocity lines over i < nx - 1, j < ny - 1 (FluidRenderer2D.cpp:452-461).
__global__ __launch_bounds__(kBlock) void render_cell_vels_kernel(MacView m, double dt, float4* __restrict__ out) {
    const int w = m.nx - 1;
    const long n = (long)w * (m.ny - 1);
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        const long i = c % w, j = c / w;
        const double x = ((double)i + 0.5) * m.dx, y = ((double)j + 0.5) * m.dx;
        const double vx = vel_interp_u(m, x, y), vy = vel_interp_v(m, x, y);
        out[c] = make_float4((float)x, (float)y, (float)(x + vx * dt), (float)(y + vy * dt));
    }
}

// Particle velocity lines (FluidRenderer2D.cpp:471-479).
__global__ __launch_bounds__(kBlock) void render_particles_kernel(MacView m, const float* __restrict__ px,
                                                                  const float* __restrict__ py, long np, float dtf,
                                                                  float4* __restrict__ out) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < np; e += (long)gridDim.x * blockDim.x) {
        const float x = px[e], y = py[e];
        const float vx = vel_interp_u(m, x, y), vy = vel_interp_v(m, x, y);
        out[e] = make_float4(x, y, x + dtf * vx, y + dtf * vy);
    }
}

// sigmoid(100 phi) per cell (FluidRenderer2D.cpp:481-485): 100.0f * phi is a
// double product, rounded to float before the sigmoid.
__global__ __launch_bounds__(kBlock) void render_phi_kernel(const double* __restrict__ phi, long n,
                                                            float* __restrict__ out) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
        out[c] = logistic((float)(100.0f * phi[c]));
}

}  // namespace fs2d

using namespace fs2d;

extern "C" fs2d_status fs2d_render_buffers(fs2d_sim* s, const fs2d_render_targets* t, fs2d_render_counts* out) {
    if (!s || !t) return FS2D_ERR_ARG;
    const long n = n_c(s);
    const int nb = (int)((n + kRenderBlock - 1) / kRenderBlock);
    fs2d_render_counts c{};
    const bool cells = t->water_cells || t->solid_cells || t->pressure_cells || t->pressure_values;
    if (cells) {
        if (!s->render_cnt) TRY(dalloc(s, &s->render_cnt, 3 * (size_t)nb + 3));
        int* tot = s->render_cnt + 3 * (size_t)nb;
        hipLaunchKernelGGL(render_count_kernel, dim3(nb), dim3(kRenderBlock), 0, s->stream, s->cell, s->p, n,
                           s->render_cnt, nb);
        hipLaunchKernelGGL(render_scan_kernel, dim3(1), dim3(kBlock), 0, s->stream, s->render_cnt, nb, tot);
        hipLaunchKernelGGL(render_scatter_kernel, dim3(nb), dim3(kRenderBlock), 0, s->stream, s->cell, s->p, s->nx,
                           n, s->dx, s->render_cnt, nb, (float2*)t->water_cells, (float2*)t->solid_cells,
                           (float2*)t->pressure_cells, t->pressure_values);
        int h[3] = {0, 0, 0};
        TRY(d2h(s, h, tot, sizeof h));
        c.water_cells = h[0];
        c.solid_cells = h[1];
        c.pressure_cells = h[2];
    }
    if (t->cell_vel_lines) {
        const long nl = (long)(s->nx - 1) * (s->ny - 1);
        hipLaunchKernelGGL(render_cell_vels_kernel, dim3(blocks_for(nl)), dim3(kBlock), 0, s->stream, mac_view(s),
                           s->dt, (float4*)t->cell_vel_lines);
        c.cell_vel_lines = nl;
    }
    if (t->particle_lines && s->np > 0)
        hipLaunchKernelGGL(render_particles_kernel, dim3(blocks_for(s->np)), dim3(kBlock), 0, s->stream, mac_view(s),
                           s->px, s->py, s->np, (float)s->dt, (float4*)t->particle_lines);
    c.particles = s->np;
    if (t->phi_values) {
        hipLaunchKernelGGL(render_phi_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s->stream, s->phi, n,
                           t->phi_values);
        c.phi_values = n;
    }
    HIPCHK(s, hipGetLastError());
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (out) *out = c;
    return FS2D_OK;
}
