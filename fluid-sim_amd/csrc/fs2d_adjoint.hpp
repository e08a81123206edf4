// fs2d_adjoint.hpp -- gfx950 kernels of the adjoint (vector-Jacobian product)
// of one step of the all-fluid semi-Lagrangian mode.
//
// The step is: out-of-place SL advection of u, v and the passive density
// (sl_advect_kernel), applyGravity (a constant: identity on adjoints),
// applyProjection (assemble_sk_kernel + the PCG) and updateVelocity with its
// extrapolation (update_velocity_kernel, extrap_*_kernel), on static cell
// flags without EMPTY cells, so A is the symmetric 5-point Laplacian.  The
// reference has no adjoint; each kernel here is the transpose of the forward
// kernel it names, applied in reverse stage order.
//
// Scatter-adds of the transposed Catmull-Rom samples go to 64-bit fixed-point
// accumulators through integer atomics.  Integer addition is associative, so a
// sum does not depend on the order in which waves arrive and the adjoint is
// bit-reproducible.  The scale is 2^E, E chosen from the largest seed of the
// launch and the number of terms so that no sum can overflow (every Catmull-Rom
// weight has magnitude <= 1).
#pragma once

#include "fs2d_kernels.hpp"

namespace fs2d {

struct AdjSl {
    MacView m;               // the step's input velocity
    const float* dens;       // the step's input density; nullptr without density
    const float* gu;         // adjoints of the SL outputs (seeds)
    const float* gv;
    const float* gd;
    unsigned long long* au;  // fixed-point accumulators of the input adjoints
    unsigned long long* av;
    unsigned long long* ad;
    unsigned* gmax;          // float bits of the largest |seed| of any sample
    double dt;
    ClampBox cb;
    int log2_terms;          // >= log2(number of scattered terms)
};

// Fixed-point scale of a launch: 2^(62 - log2_terms - e) with max|seed| < 2^e.
__device__ __forceinline__ double adj_scale(const unsigned* gmax, int log2_terms) {
    const float G = __uint_as_float(*gmax);
    if (!(G > 0.0f)) return 1.0;
    int e;
    frexpf(G, &e);
    return ldexp(1.0, 62 - log2_terms - e);
}

// Catmull-Rom weights of catmull_rom() (fs2d_device.hpp) and their derivatives.
__device__ __forceinline__ void cr_weights(float f, float w[4], float dw[4]) {
    const float f2 = f * f, f3 = f2 * f;
    w[0] = -0.5f * f3 + f2 - 0.5f * f;
    w[1] = 1.5f * f3 - 2.5f * f2 + 1.0f;
    w[2] = -1.5f * f3 + 2.0f * f2 + 0.5f * f;
    w[3] = 0.5f * f3 - 0.5f * f2;
    dw[0] = -1.5f * f2 + 2.0f * f - 0.5f;
    dw[1] = 4.5f * f2 - 5.0f * f;
    dw[2] = -4.5f * f2 + 4.0f * f + 0.5f;
    dw[3] = 1.5f * f2 - f;
}

__device__ __forceinline__ void acc_add(unsigned long long* p, double x) {
    __hip_atomic_fetch_add(p, (unsigned long long)__double2ll_rn(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Transpose of one catmull_rom() sample of the w x h array `a` at grid
// coordinates (gx, gy), seeded with g.  SCATTER adds g * weight to each tap's
// accumulator (scale S); GRAD adds d sample / d(gx, gy) (without g) to (sx, sy).
template <bool SCATTER, bool GRAD>
__device__ __forceinline__ void cr_adjoint(const float* __restrict__ a, unsigned long long* acc, int w, int h,
                                           double gx, double gy, float g, double S, float& sx, float& sy) {
    const int x = (int)gx, y = (int)gy;
    if (x < 0 || x >= w || y < 0 || y >= h) return;  // the sample is 0 there
    const float fx = (float)(gx - (double)x), fy = (float)(gy - (double)y);
    float wu[4], du[4], wv[4], dv[4];
    cr_weights(fx, wu, du);
    cr_weights(fy, wv, dv);
    const int xs[4] = {clampi(x - 1, 0, w - 1), clampi(x, 0, w - 1), clampi(x + 1, 0, w - 1), clampi(x + 2, 0, w - 1)};
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const long row = (long)clampi(y - 1 + r, 0, h - 1) * w;
        if (GRAD) {
            const float a0 = a[row + xs[0]], a1 = a[row + xs[1]], a2 = a[row + xs[2]], a3 = a[row + xs[3]];
            const float rs = (wu[0] * a0 + wu[1] * a1) + (wu[2] * a2 + wu[3] * a3);
            const float drs = (du[0] * a0 + du[1] * a1) + (du[2] * a2 + du[3] * a3);
            sx += wv[r] * drs;
            sy += dv[r] * rs;
        }
        if (SCATTER) {
#pragma unroll
            for (int c = 0; c < 4; c++) acc_add(acc + row + xs[c], (double)(g * (wu[c] * wv[r])) * S);
        }
    }
}

// Transpose of vel_interp_u (WHICH 0), vel_interp_v (1) or the density sample
// of sl_advect_kernel (2) at the position (px, py) in metres, seeded with g:
// scatters into acc and adds g * d sample / d(px, py) to (bx, by).  The clamps
// to [1e-6, n - 1 - 1e-6] pass no gradient where they are active.
template <int WHICH, bool SCATTER, bool GRAD>
__device__ __forceinline__ void interp_adjoint(const float* __restrict__ a, unsigned long long* acc, int nx, int ny,
                                               double dx, double px, double py, float g, double S, float& bx,
                                               float& by) {
    // grid units as vel_interp_u / vel_interp_v / the SL density sample form them
    const double inv_dx = 1.0 / dx;
    double gx = px * inv_dx, gy = py * inv_dx;
    if (WHICH == 0) gy -= 0.5;
    else if (WHICH == 1) gx -= 0.5;
    else { gx -= 0.5; gy -= 0.5; }
    const double hx = (double)(nx - 1) - 1e-6, hy = (double)(ny - 1) - 1e-6;
    const int w = WHICH == 0 ? nx + 1 : nx, h = WHICH == 1 ? ny + 1 : ny;
    float sx = 0.0f, sy = 0.0f;
    cr_adjoint<SCATTER, GRAD>(a, acc, w, h, clampd(gx, 1e-6, hx), clampd(gy, 1e-6, hy), g, S, sx, sy);
    if (GRAD) {
        const float inv = (float)(1.0 / dx);
        if (gx >= 1e-6 && gx <= hx) bx += g * sx * inv;
        if (gy >= 1e-6 && gy <= hy) by += g * sy * inv;
    }
}

// Transpose of one output of sl_advect_kernel: the sample of field `kind`
// (0 u, 1 v, 2 density) at the departure point of (x0, y0), seeded with g.
// The departure point is recomputed with backtrace()'s arithmetic; the adjoint
// runs back through its clamp, the Ralston combination and the three RK
// stages.  Returns the largest |seed| of the seven samples.
template <bool SCATTER>
__device__ __forceinline__ float adj_point(const AdjSl& a, int kind, double x0, double y0, float g, double S) {
    const MacView& m = a.m;
    const double dt = a.dt;
    const double k1x = vel_interp_u(m, x0, y0), k1y = vel_interp_v(m, x0, y0);
    const double x2 = x0 - 0.5 * dt * k1x, y2 = y0 - 0.5 * dt * k1y;
    const double k2x = vel_interp_u(m, x2, y2), k2y = vel_interp_v(m, x2, y2);
    const double x3 = x0 - 0.75 * dt * k2x, y3 = y0 - 0.75 * dt * k2y;
    const double k3x = vel_interp_u(m, x3, y3), k3y = vel_interp_v(m, x3, y3);
    double sx = (2. / 9.) * dt * k1x + (3. / 9.) * dt * k2x;
    double sy = (2. / 9.) * dt * k1y + (3. / 9.) * dt * k2y;
    sx = sx + (4. / 9.) * dt * k3x;
    sy = sy + (4. / 9.) * dt * k3y;
    const double rx = x0 - sx, ry = y0 - sy;
    const double ox = clampd(rx, a.cb.lox, a.cb.hix), oy = clampd(ry, a.cb.loy, a.cb.hiy);
    const int nx = m.nx, ny = m.ny;
    const double dx = m.dx;

    float Xbx = 0.0f, Xby = 0.0f;
    if (kind == 0) interp_adjoint<0, SCATTER, true>(m.u, a.au, nx, ny, dx, ox, oy, g, S, Xbx, Xby);
    else if (kind == 1) interp_adjoint<1, SCATTER, true>(m.v, a.av, nx, ny, dx, ox, oy, g, S, Xbx, Xby);
    else interp_adjoint<2, SCATTER, true>(a.dens, a.ad, nx, ny, dx, ox, oy, g, S, Xbx, Xby);
    if (!(rx >= a.cb.lox && rx <= a.cb.hix)) Xbx = 0.0f;
    if (!(ry >= a.cb.loy && ry <= a.cb.hiy)) Xby = 0.0f;

    const float c1 = (float)(-(2. / 9.) * dt), c2 = (float)(-(3. / 9.) * dt), c3 = (float)(-(4. / 9.) * dt);
    const float h2 = (float)(-0.5 * dt), h3 = (float)(-0.75 * dt);
    const float k3bx = c3 * Xbx, k3by = c3 * Xby;
    float x3bx = 0.0f, x3by = 0.0f;
    interp_adjoint<0, SCATTER, true>(m.u, a.au, nx, ny, dx, x3, y3, k3bx, S, x3bx, x3by);
    interp_adjoint<1, SCATTER, true>(m.v, a.av, nx, ny, dx, x3, y3, k3by, S, x3bx, x3by);
    const float k2bx = c2 * Xbx + h3 * x3bx, k2by = c2 * Xby + h3 * x3by;
    float x2bx = 0.0f, x2by = 0.0f;
    interp_adjoint<0, SCATTER, true>(m.u, a.au, nx, ny, dx, x2, y2, k2bx, S, x2bx, x2by);
    interp_adjoint<1, SCATTER, true>(m.v, a.av, nx, ny, dx, x2, y2, k2by, S, x2bx, x2by);
    const float k1bx = c1 * Xbx + h2 * x2bx, k1by = c1 * Xby + h2 * x2by;
    float unused_x = 0.0f, unused_y = 0.0f;  // (x0, y0) is a constant
    interp_adjoint<0, SCATTER, false>(m.u, a.au, nx, ny, dx, x0, y0, k1bx, S, unused_x, unused_y);
    interp_adjoint<1, SCATTER, false>(m.v, a.av, nx, ny, dx, x0, y0, k1by, S, unused_x, unused_y);
    float mx = fabsf(g);
    mx = fmaxf(mx, fmaxf(fabsf(k3bx), fabsf(k3by)));
    mx = fmaxf(mx, fmaxf(fabsf(k2bx), fabsf(k2by)));
    mx = fmaxf(mx, fmaxf(fabsf(k1bx), fabsf(k1by)));
    return mx;
}

// Transpose of sl_advect_kernel.  Pass SCATTER = false finds the largest seed
// (-> the fixed-point scale), pass SCATTER = true scatters; both walk the same
// outputs with the same arithmetic.
template <bool SCATTER>
__global__ __launch_bounds__(kBlock) void adj_sl_kernel(AdjSl a) {
    const int nx = a.m.nx, ny = a.m.ny;
    const long nu = (long)(nx + 1) * ny, nv = (long)nx * (ny + 1);
    const long nd = a.dens ? (long)nx * ny : 0;
    const long n = nu + nv + nd;
    const double S = SCATTER ? adj_scale(a.gmax, a.log2_terms) : 0.0;
    const double dx = a.m.dx;
    float mx = 0.0f;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        int kind, i, j;
        float g;
        double x0, y0;
        if (e < nu) {
            kind = 0;
            i = (int)(e % (nx + 1)); j = (int)(e / (nx + 1));
            g = a.gu[e];
            x0 = (double)i * dx; y0 = ((double)j + 0.5) * dx;
        } else if (e < nu + nv) {
            const long f = e - nu;
            kind = 1;
            i = (int)(f % nx); j = (int)(f / nx);
            g = a.gv[f];
            x0 = ((double)i + 0.5) * dx; y0 = (double)j * dx;
        } else {
            const long f = e - nu - nv;
            kind = 2;
            i = (int)(f % nx); j = (int)(f / nx);
            g = a.gd[f];
            x0 = ((double)i + 0.5) * dx; y0 = ((double)j + 0.5) * dx;
        }
        if (g == 0.0f) continue;
        mx = fmaxf(mx, adj_point<SCATTER>(a, kind, x0, y0, g, S));
    }
    if (!SCATTER) {
        __shared__ float sh[kBlock / 64];
        mx = wave_max(mx);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            float t = 0.0f;
            for (int k = 0; k < kBlock / 64; k++) t = fmaxf(t, sh[k]);
            if (t > 0.0f) atomicMax(a.gmax, __float_as_uint(t));
        }
    }
}

// Accumulators -> FP32 adjoints (u, v, density outputs), clearing the
// accumulators for the next launch.
__global__ __launch_bounds__(kBlock) void adj_fixed_to_float_kernel(AdjSl a, float* __restrict__ ou,
                                                                    float* __restrict__ ov, float* __restrict__ od) {
    const int nx = a.m.nx, ny = a.m.ny;
    const long nu = (long)(nx + 1) * ny, nv = (long)nx * (ny + 1);
    const long nd = a.dens ? (long)nx * ny : 0;
    const double inv = 1.0 / adj_scale(a.gmax, a.log2_terms);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nu + nv + nd;
         e += (long)gridDim.x * blockDim.x) {
        unsigned long long* acc;
        float* out;
        long f;
        if (e < nu) { acc = a.au; out = ou; f = e; }
        else if (e < nu + nv) { acc = a.av; out = ov; f = e - nu; }
        else { acc = a.ad; out = od; f = e - nu - nv; }
        out[f] = (float)((double)(long long)acc[f] * inv);
        acc[f] = 0ull;
    }
}

// ---------------------------------------------------------------------------
// Transpose of extrapolate() on static masks.  Stage L (L = Lmax .. 1) hands
// the adjoint of every layer-L entry in equal parts to the layer-(L-1)
// neighbours it averaged, written as a gather per receiving entry with its
// sources in a fixed order (deterministic); the extrapolated entries are then
// zeroed, since extrapolation overwrote their values.  The lists come from the
// host, which layers the faces exactly as extrap_layer1/extrap_layers do.
// One workgroup per array.
// ---------------------------------------------------------------------------
struct ExtraAdj {
    float* g;
    const int* tgt;        // receiving entries, grouped by stage
    const int4* src;       // up to four sources per receiver (-1: none)
    const float4* wt;      // 1 / (number of lower-layer neighbours) of each source
    const int* stage_off;  // nstages + 1 offsets into tgt
    int nstages;
    const int* zero;       // extrapolated entries
    int nzero;
};

__global__ __launch_bounds__(1024) void adj_extrap_kernel(ExtraAdj A, ExtraAdj B) {
    const ExtraAdj X = blockIdx.x == 0 ? A : B;
    for (int st = 0; st < X.nstages; st++) {
        for (int k = X.stage_off[st] + (int)threadIdx.x; k < X.stage_off[st + 1]; k += blockDim.x) {
            const int t = X.tgt[k];
            const int4 s = X.src[k];
            const float4 w = X.wt[k];
            float acc = X.g[t];
            if (s.x >= 0) acc += X.g[s.x] * w.x;
            if (s.y >= 0) acc += X.g[s.y] * w.y;
            if (s.z >= 0) acc += X.g[s.z] * w.z;
            if (s.w >= 0) acc += X.g[s.w] * w.w;
            X.g[t] = acc;
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < X.nzero; k += blockDim.x) X.g[X.zero[k]] = 0.0f;
}

// Transpose of updateVelocity's pressure gradient (nu -= scale (p(i) - p(i-1))
// on faces between two FLUID cells): p-bar of every FLUID cell, written in SK4
// order as the right-hand side of the adjoint solve, with p = s = 0 and the
// |rhs|_inf partials exactly as assemble_sk_kernel leaves them.
__global__ __launch_bounds__(kBlock) void adj_pressure_rhs_kernel(
    SkGeom g, const uint8_t* __restrict__ cell, const float* __restrict__ gu, const float* __restrict__ gv,
    float* __restrict__ r, float* __restrict__ p, float* __restrict__ s0, double scale, double* partials) {
    const int nx = g.nx;
    double m = 0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < g.size(); e += (long)gridDim.x * blockDim.x) {
        int i, j;
        double rhs = 0;
        if (sk_cell(g, e, i, j)) {
            const long c = (long)j * nx + i;
            if (cell[c] == FLUID) {
                const long iu = (long)j * (nx + 1) + i;
                if (cell[c - 1] == FLUID) rhs -= scale * (double)gu[iu];
                if (cell[c + 1] == FLUID) rhs += scale * (double)gu[iu + 1];
                if (cell[c - nx] == FLUID) rhs -= scale * (double)gv[c];
                if (cell[c + nx] == FLUID) rhs += scale * (double)gv[c + nx];
            }
        }
        r[e] = (float)rhs;
        p[e] = 0.0f;
        s0[e] = 0.0f;
        const double a = fabs((double)(float)rhs);
        m = a > m ? a : m;
    }
    double bm = block_max(m);
    if (threadIdx.x == 0) partials[blockIdx.x] = bm;
}

// Transpose of update_velocity_kernel's face update and of assemble_sk_kernel's
// rhs = -(1/dx) div u (faces next to SOLID cancel there), in place on (gu, gv):
// a face next to FLUID and SOLID was set to 0 and passes nothing; every other
// face passes its adjoint through; lam = A^-1 p-bar adds sr * (R^T lam).
__global__ __launch_bounds__(kBlock) void adj_update_velocity_kernel(const uint8_t* __restrict__ cell,
                                                                     const float* __restrict__ lam,
                                                                     float* __restrict__ gu, float* __restrict__ gv,
                                                                     int nx, int ny, double sr) {
    const long nU = (long)(nx + 1) * ny, nV = (long)nx * (ny + 1);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nU + nV; e += (long)gridDim.x * blockDim.x) {
        if (e < nU) {
            const int i = (int)(e % (nx + 1)), j = (int)(e / (nx + 1));
            const long c = (long)j * nx + i;  // cell (i, j) when i < nx
            const uint8_t cc = i < nx ? cell[c] : SOLID;
            const uint8_t cl = i > 0 ? cell[c - 1] : SOLID;
            double acc = gu[e];
            if (i < nx && ((i > 0 && cl == FLUID) || cc == FLUID) && (i == 0 || cl == SOLID || cc == SOLID)) acc = 0;
            if (cc == FLUID && cl != SOLID) acc += sr * (double)lam[c];
            if (cl == FLUID && cc != SOLID) acc -= sr * (double)lam[c - 1];
            gu[e] = (float)acc;
        } else {
            const long f = e - nU;
            const int i = (int)(f % nx), j = (int)(f / nx);
            const long c = (long)j * nx + i;  // cell (i, j) when j < ny
            const uint8_t cc = j < ny ? cell[c] : SOLID;
            const uint8_t cb = j > 0 ? cell[c - nx] : SOLID;
            double acc = gv[f];
            if (j < ny && ((j > 0 && cb == FLUID) || cc == FLUID) && (j == 0 || cb == SOLID || cc == SOLID)) acc = 0;
            if (cc == FLUID && cb != SOLID) acc += sr * (double)lam[c];
            if (cb == FLUID && cc != SOLID) acc -= sr * (double)lam[c - nx];
            gv[f] = (float)acc;
        }
    }
}

}  // namespace fs2d
