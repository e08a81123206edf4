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

// Catmull-Rom (Keys a = -1/2) sample of a w x h row-major FP32 array at grid
// coordinates (gx, gy): Array2D<double>::triCubic (include/Array2D.h:244-327).
// The index is truncated, the taps clamped to the array, out-of-range bases
// return 0, and the sums follow the reference's AVX order.
__device__ __forceinline__ float catmull_rom(const float* __restrict__ a, int w, int h, double gx,
                                             double gy) {
    int x = (int)gx, y = (int)gy;
    if (x < 0 || x >= w || y < 0 || y >= h) return 0.0f;
    float fx = (float)(gx - (double)x), fy = (float)(gy - (double)y);
    float fx2 = fx * fx, fx3 = fx2 * fx, fy2 = fy * fy, fy3 = fy2 * fy;
    float u0 = -0.5f * fx3 + fx2 - 0.5f * fx;
    float u1 = 1.5f * fx3 - 2.5f * fx2 + 1.0f;
    float u2 = -1.5f * fx3 + 2.0f * fx2 + 0.5f * fx;
    float u3 = 0.5f * fx3 - 0.5f * fx2;
    float v[4];
    v[0] = -0.5f * fy3 + fy2 - 0.5f * fy;
    v[1] = 1.5f * fy3 - 2.5f * fy2 + 1.0f;
    v[2] = -1.5f * fy3 + 2.0f * fy2 + 0.5f * fy;
    v[3] = 0.5f * fy3 - 0.5f * fy2;
    const int x0 = clampi(x - 1, 0, w - 1), x1 = clampi(x, 0, w - 1);
    const int x2 = clampi(x + 1, 0, w - 1), x3 = clampi(x + 2, 0, w - 1);
    float f[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float* row = a + (size_t)clampi(y - 1 + r, 0, h - 1) * w;
        f[r] = ((u0 * row[x0] + u1 * row[x1]) + (u2 * row[x2] + u3 * row[x3])) * v[r];
    }
    return (f[0] + f[2]) + (f[1] + f[3]);
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
struct MacView {
    const float* __restrict__ u;
    const float* __restrict__ v;
    int nx, ny;
    double dx;
};

// MACGrid2D::velInterpU / velInterpV (include/MACGrid2D.h:80-98).
__device__ __forceinline__ float vel_interp_u(const MacView& m, double px, double py) {
    px /= m.dx;
    py /= m.dx;
    py -= 0.5;
    px = clampd(px, 1e-6, (double)(m.nx - 1) - 1e-6);
    py = clampd(py, 1e-6, (double)(m.ny - 1) - 1e-6);
    return catmull_rom(m.u, m.nx + 1, m.ny, px, py);
}
__device__ __forceinline__ float vel_interp_v(const MacView& m, double px, double py) {
    px /= m.dx;
    py /= m.dx;
    px -= 0.5;
    px = clampd(px, 1e-6, (double)(m.nx - 1) - 1e-6);
    py = clampd(py, 1e-6, (double)(m.ny - 1) - 1e-6);
    return catmull_rom(m.v, m.nx, m.ny + 1, px, py);
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
// (src/FluidSim2D.cpp:212-217).
__device__ __forceinline__ void backtrace(const MacView& m, double dt, const ClampBox& cb, double xp,
                                          double yp, double& ox, double& oy) {
    double k1x = vel_interp_u(m, xp, yp), k1y = vel_interp_v(m, xp, yp);
    double x2 = xp - 0.5 * dt * k1x, y2 = yp - 0.5 * dt * k1y;
    double k2x = vel_interp_u(m, x2, y2), k2y = vel_interp_v(m, x2, y2);
    double x3 = xp - 0.75 * dt * k2x, y3 = yp - 0.75 * dt * k2y;
    double k3x = vel_interp_u(m, x3, y3), k3y = vel_interp_v(m, x3, y3);
    double sx = (2. / 9.) * dt * k1x + (3. / 9.) * dt * k2x;
    double sy = (2. / 9.) * dt * k1y + (3. / 9.) * dt * k2y;
    sx = sx + (4. / 9.) * dt * k3x;
    sy = sy + (4. / 9.) * dt * k3y;
    ox = clampd(xp - sx, cb.lox, cb.hix);
    oy = clampd(yp - sy, cb.loy, cb.hiy);
}

}  // namespace fs2d
