// fs2d_device.hpp -- device helpers shared by the gfx950 kernels.
//
// Positions are carried in FP64 (metres, as in the reference) and every sample
// of an FP32 field is taken at an FP64 position: the truncation to a cell index
// and the fractional offset then match the reference's double arithmetic, and
// only the weights and taps are FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fs2d {

enum : uint8_t { EMPTY = 0, FLUID = 1, SOLID = 2 };

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ double clampd(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

// Catmull-Rom (Keys a = -1/2) weights at offset f (include/Array2D.h:265-273),
// every product and sum rounded on its own (no contraction), so that each
// sampler below gives the same bits.
__device__ __forceinline__ void cr_weights(float f, float w[4]) {
#pragma clang fp contract(off)
    const float f2 = f * f, f3 = f2 * f;
    w[0] = -0.5f * f3 + f2 - 0.5f * f;
    w[1] = 1.5f * f3 - 2.5f * f2 + 1.0f;
    w[2] = -1.5f * f3 + 2.0f * f2 + 0.5f * f;
    w[3] = 0.5f * f3 - 0.5f * f2;
}
// Sum of the 4 x 4 taps t[row][col] in the reference's AVX order
// (Array2D.h:310-326): row dots, then (r0 + r2) + (r1 + r3).
__device__ __forceinline__ float cr_combine(const float (&t)[4][4], const float (&wu)[4], const float (&wv)[4]) {
#pragma clang fp contract(off)
    float f[4];
#pragma unroll
    for (int r = 0; r < 4; r++) f[r] = ((wu[0] * t[r][0] + wu[1] * t[r][1]) + (wu[2] * t[r][2] + wu[3] * t[r][3])) * wv[r];
    return (f[0] + f[2]) + (f[1] + f[3]);
}
// The taps of a sample at grid coordinates (gx, gy) of a w x h array: the
// index truncated, the taps clamped to the array (Array2D.h:244-283).
struct CrTaps {
    int xs[4], ys[4];
    float wu[4], wv[4];
};
__device__ __forceinline__ bool cr_taps(int w, int h, double gx, double gy, CrTaps& T) {
    const int x = (int)gx, y = (int)gy;
    if (x < 0 || x >= w || y < 0 || y >= h) return false;  // the sample is 0 (Array2D.h:245-248)
    cr_weights((float)(gx - (double)x), T.wu);
    cr_weights((float)(gy - (double)y), T.wv);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        T.xs[k] = clampi(x - 1 + k, 0, w - 1);
        T.ys[k] = clampi(y - 1 + k, 0, h - 1);
    }
    return true;
}

// Catmull-Rom sample of a w x h row-major FP32 array at grid coordinates
// (gx, gy): Array2D<double>::triCubic (include/Array2D.h:244-327).
__device__ __forceinline__ float catmull_rom(const float* __restrict__ a, int w, int h, double gx, double gy) {
    CrTaps T;
    if (!cr_taps(w, h, gx, gy, T)) return 0.0f;
    float t[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) t[r][c] = a[(size_t)T.ys[r] * w + T.xs[c]];
    return cr_combine(t, T.wu, T.wv);
}

// A window of a w x h array copied to LDS: columns x0 .. x0 + ww - 1, rows
// y0 .. y0 + wh - 1, row-major with stride ww.
struct LdsWin {
    const float* s;
    int x0, y0, ww, wh;
};
// catmull_rom() reading its taps from the window when all 16 lie inside it,
// from the array otherwise; the same bits either way.
__device__ __forceinline__ float catmull_rom_win(const LdsWin& W, const float* __restrict__ a, int w, int h,
                                                 double gx, double gy) {
    CrTaps T;
    if (!cr_taps(w, h, gx, gy, T)) return 0.0f;
    float t[4][4];
    const bool in = T.xs[0] >= W.x0 && T.xs[3] < W.x0 + W.ww && T.ys[0] >= W.y0 && T.ys[3] < W.y0 + W.wh;
    if (in) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float* row = W.s + (T.ys[r] - W.y0) * W.ww - W.x0;
#pragma unroll
            for (int c = 0; c < 4; c++) t[r][c] = row[T.xs[c]];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) t[r][c] = a[(size_t)T.ys[r] * w + T.xs[c]];
    }
    return cr_combine(t, T.wu, T.wv);
}

// Bilinear gather of Array2D<double>::linearExtract (include/Array2D.h:402-420)
// at grid coordinates (gx, gy) with clamped corners.
__device__ __forceinline__ float bilinear(const float* __restrict__ a, int w, int h, double gx, double gy) {
    int ui = (int)gx, uj = (int)gy;
    float ddx = (float)(gx - ui), ddy = (float)(gy - uj);
    int x1 = clampi(ui, 0, w - 1), x2 = clampi(ui + 1, 0, w - 1);
    int y1 = clampi(uj, 0, h - 1), y2 = clampi(uj + 1, 0, h - 1);
    float value = 0.0f;
    value += a[(size_t)y1 * w + x1] * (1 - ddx) * (1 - ddy);
    value += a[(size_t)y1 * w + x2] * ddx * (1 - ddy);
    value += a[(size_t)y2 * w + x1] * (1 - ddx) * ddy;
    value += a[(size_t)y2 * w + x2] * ddx * ddy;
    return value;
}

// A staggered MAC grid: u is (nx+1) x ny, v is nx x (ny+1) (MACGrid2D.h:19-25).
// inv_dx = 1 / dx: positions go to grid units by one multiplication (the
// reference divides, MACGrid2D.h:81-82; the two agree to an ulp, exactly for
// power-of-two dx).
struct MacView {
    const float* __restrict__ u;
    const float* __restrict__ v;
    int nx, ny;
    double dx, inv_dx;
};
__host__ __device__ inline MacView make_mac(const float* u, const float* v, int nx, int ny, double dx) {
    return MacView{u, v, nx, ny, dx, 1.0 / dx};
}

// MACGrid2D::velInterpU / velInterpV (include/MACGrid2D.h:80-98): grid
// coordinates of a position in metres, shifted to the staggered axis and
// clamped to [1e-6, n - 1 - 1e-6].
__device__ __forceinline__ void u_coords(const MacView& m, double px, double py, double& gx, double& gy) {
    gx = clampd(px * m.inv_dx, 1e-6, (double)(m.nx - 1) - 1e-6);
    gy = clampd(py * m.inv_dx - 0.5, 1e-6, (double)(m.ny - 1) - 1e-6);
}
__device__ __forceinline__ void v_coords(const MacView& m, double px, double py, double& gx, double& gy) {
    gx = clampd(px * m.inv_dx - 0.5, 1e-6, (double)(m.nx - 1) - 1e-6);
    gy = clampd(py * m.inv_dx, 1e-6, (double)(m.ny - 1) - 1e-6);
}
__device__ __forceinline__ float vel_interp_u(const MacView& m, double px, double py) {
    double gx, gy;
    u_coords(m, px, py, gx, gy);
    return catmull_rom(m.u, m.nx + 1, m.ny, gx, gy);
}
__device__ __forceinline__ float vel_interp_v(const MacView& m, double px, double py) {
    double gx, gy;
    v_coords(m, px, py, gx, gy);
    return catmull_rom(m.v, m.nx, m.ny + 1, gx, gy);
}

// FluidSim2D::clampPos bounds (src/FluidSim2D.cpp:645-651).
struct ClampBox {
    double lox, hix, loy, hiy;
};
__host__ __device__ inline ClampBox clamp_box(int nx, int ny, double dx) {
    const double off = 1e-3;
    return ClampBox{(1.0 + off) * dx, (nx - 1.0 - off) * dx, (1.0 + off) * dx, (ny - 1.0 - off) * dx};
}

// Semi-Lagrangian departure point: backward Ralston RK3 then clampPos
// (src/FluidSim2D.cpp:212-217), with the velocity samplers of S (S.u, S.v at
// a position in metres).
template <class S>
__device__ __forceinline__ void backtrace_with(const S& smp, double dt, const ClampBox& cb, double xp, double yp,
                                               double& ox, double& oy) {
    double k1x = smp.u(xp, yp), k1y = smp.v(xp, yp);
    double x2 = xp - 0.5 * dt * k1x, y2 = yp - 0.5 * dt * k1y;
    double k2x = smp.u(x2, y2), k2y = smp.v(x2, y2);
    double x3 = xp - 0.75 * dt * k2x, y3 = yp - 0.75 * dt * k2y;
    double k3x = smp.u(x3, y3), k3y = smp.v(x3, y3);
    double sx = (2. / 9.) * dt * k1x + (3. / 9.) * dt * k2x;
    double sy = (2. / 9.) * dt * k1y + (3. / 9.) * dt * k2y;
    sx = sx + (4. / 9.) * dt * k3x;
    sy = sy + (4. / 9.) * dt * k3y;
    ox = clampd(xp - sx, cb.lox, cb.hix);
    oy = clampd(yp - sy, cb.loy, cb.hiy);
}
struct GlobalSampler {
    MacView m;
    __device__ float u(double px, double py) const { return vel_interp_u(m, px, py); }
    __device__ float v(double px, double py) const { return vel_interp_v(m, px, py); }
};
// Samplers reading u and v through LDS windows (catmull_rom_win): the same
// bits as GlobalSampler.
struct WinSampler {
    MacView m;
    LdsWin U, V;
    __device__ float u(double px, double py) const {
        double gx, gy;
        u_coords(m, px, py, gx, gy);
        return catmull_rom_win(U, m.u, m.nx + 1, m.ny, gx, gy);
    }
    __device__ float v(double px, double py) const {
        double gx, gy;
        v_coords(m, px, py, gx, gy);
        return catmull_rom_win(V, m.v, m.nx, m.ny + 1, gx, gy);
    }
};
__device__ __forceinline__ void backtrace(const MacView& m, double dt, const ClampBox& cb, double xp, double yp,
                                          double& ox, double& oy) {
    backtrace_with(GlobalSampler{m}, dt, cb, xp, yp, ox, oy);
}

}  // namespace fs2d
