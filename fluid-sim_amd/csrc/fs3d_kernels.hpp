// fs3d_kernels.hpp -- gfx950 kernels of the 3D MAC grid (include/MACGrid3D.h,
// include/Array3D.h).
//
// Layout: u (nx+1) x ny x nz, v nx x (ny+1) x nz, w nx x ny x (nz+1), FP32,
// element (i, j, k) of an NX x NY x NZ array at [k*NY*NX + j*NX + i]
// (Array3D.h:32-37).  Cell and face kernels walk their output in that order;
// each XCD takes one contiguous eighth of it (fs2d::xcd_loop), so the
// neighbouring slab a stencil reads (k+1) is mostly in the same XCD's L2.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fs2d_kernels.hpp"

namespace fs3d {

using fs2d::kBlock;

struct Grid3 {
    const float* u;
    const float* v;
    const float* w;
    int nx, ny, nz;
    __device__ float U(int i, int j, int k) const { return u[((long)k * ny + j) * (nx + 1) + i]; }
    __device__ float V(int i, int j, int k) const { return v[((long)k * (ny + 1) + j) * nx + i]; }
    __device__ float W(int i, int j, int k) const { return w[((long)k * ny + j) * nx + i]; }
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Catmull-Rom factors of Array3D::triCubic (Array3D.h:91-104), as written.
template <class T>
__device__ __forceinline__ void cr4(T d, T (&f)[4]) {
    const T d2 = d * d, d3 = d2 * d;
    f[0] = T(-0.5) * d3 + d2 - T(0.5) * d;
    f[1] = T(1.5) * d3 - T(2.5) * d2 + T(1);
    f[2] = T(-1.5) * d3 + T(2) * d2 + T(0.5) * d;
    f[3] = T(0.5) * d3 - T(0.5) * d2;
}

// Array3D::triCubic (Array3D.h:75-126) of one NX x NY x NZ array at (px, py, pz).
// REF: taps read as (*this)(zp, yp, xp), i.e. element xp*NY*NX + yp*NX + zp
// (:117), and the sum clamped at zero (:123); a tap past the end of the array
// makes the sample NaN and sets `oob`.  !REF: taps (xp, yp, zp), signed sum.
// Sums run in the reference's loop order: r_j over i, q_k over j, vox over k.
template <class T, bool REF>
__device__ __forceinline__ T tricubic(const float* __restrict__ a, int NX, int NY, int NZ, T px, T py, T pz,
                                      bool& oob) {
    const int x = (int)px, y = (int)py, z = (int)pz;  // (int): truncation, as :77
    if (x < 0 || x >= NX || y < 0 || y >= NY || z < 0 || z >= NZ) return T(0);  // :78-79
    T fu[4], fv[4], fw[4];
    cr4<T>(px - (T)x, fu);
    cr4<T>(py - (T)y, fv);
    cr4<T>(pz - (T)z, fw);
    const long size = (long)NX * NY * NZ;
    int xs[4];
#pragma unroll
    for (int i = 0; i < 4; i++) xs[i] = clampi(x + i - 1, 0, NX - 1);
    T vox = 0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int zp = clampi(z + k - 1, 0, NZ - 1);
        T q = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int yp = clampi(y + j - 1, 0, NY - 1);
            T r = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                long e;
                if (REF) {
                    e = (long)xs[i] * NY * NX + (long)yp * NX + zp;
                    bad |= e >= size;
                    e = e < size ? e : size - 1;  // the load stays in bounds; the sample is dropped below
                } else {
                    e = ((long)zp * NY + yp) * NX + xs[i];
                }
                r += fu[i] * (T)a[e];
            }
            q += fv[j] * r;
        }
        vox += fw[k] * q;
    }
    if (REF) {
        if (bad) {
            oob = true;
            return (T)NAN;
        }
        return vox < 0 ? T(0) : vox;
    }
    return vox;
}

// MACGrid3D::velInterp (MACGrid3D.h:143-163) at n points, xyz interleaved.
// One thread per point; a sample whose taps left the array counts once in
// *range_errors.
template <class T, bool REF>
__global__ __launch_bounds__(kBlock) void vel_interp_kernel(Grid3 g, const T* __restrict__ pos, long n,
                                                            T* __restrict__ out, unsigned* range_errors) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        const T x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
        bool oob = false;
        const T a = tricubic<T, REF>(g.u, g.nx + 1, g.ny, g.nz, x, y - T(0.5), z - T(0.5), oob);  // velInterpU :143-147
        const T b = tricubic<T, REF>(g.v, g.nx, g.ny + 1, g.nz, x - T(0.5), y, z - T(0.5), oob);  // velInterpV :149-153
        const T c = tricubic<T, REF>(g.w, g.nx, g.ny, g.nz + 1, x - T(0.5), y - T(0.5), z, oob);  // velInterpW :155-159
        out[3 * p] = a;
        out[3 * p + 1] = b;
        out[3 * p + 2] = c;
        if (REF && oob) atomicAdd(range_errors, 1u);
    }
}

// MACGrid3D::velDiv (MACGrid3D.h:165-167), summed left to right as written.
__global__ __launch_bounds__(kBlock) void vel_div_kernel(Grid3 g, float* __restrict__ out) {
    const long n = (long)g.nx * g.ny * g.nz;
    const fs2d::XcdLoop it = fs2d::xcd_loop(n, blockIdx.x, gridDim.x);
    for (long c = it.e; c < it.end; c += it.stride) {
        const int i = (int)(c % g.nx);
        const long jk = c / g.nx;
        const int j = (int)(jk % g.ny), k = (int)(jk / g.ny);
        float d = g.U(i + 1, j, k) - g.U(i, j, k);
        d = d + g.V(i, j + 1, k);
        d = d - g.V(i, j, k);
        d = d + g.W(i, j, k + 1);
        out[c] = d - g.W(i, j, k);
    }
}

// MACGrid3D::vel (MACGrid3D.h:73-78): the cell-centre average of each component.
__global__ __launch_bounds__(kBlock) void vel_cell_kernel(Grid3 g, float* __restrict__ out) {
    const long n = (long)g.nx * g.ny * g.nz;
    const fs2d::XcdLoop it = fs2d::xcd_loop(n, blockIdx.x, gridDim.x);
    for (long c = it.e; c < it.end; c += it.stride) {
        const int i = (int)(c % g.nx);
        const long jk = c / g.nx;
        const int j = (int)(jk % g.ny), k = (int)(jk / g.ny);
        out[3 * c] = 0.5f * (g.U(i + 1, j, k) + g.U(i, j, k));
        out[3 * c + 1] = 0.5f * (g.V(i, j + 1, k) + g.V(i, j, k));
        out[3 * c + 2] = 0.5f * (g.W(i, j, k + 1) + g.W(i, j, k));
    }
}

// MACGrid3D::velU / velV / velW (MACGrid3D.h:80-141): the face's own
// component and the quarter-sums of the other two around it; at the first
// and last face of the axis only the two terms on the grid side remain (and
// are still divided by four).
template <int WHICH>
__global__ __launch_bounds__(kBlock) void vel_face_kernel(Grid3 g, float* __restrict__ out) {
    const int NX = g.nx + (WHICH == 0), NY = g.ny + (WHICH == 1), NZ = g.nz + (WHICH == 2);
    const long n = (long)NX * NY * NZ;
    const fs2d::XcdLoop it = fs2d::xcd_loop(n, blockIdx.x, gridDim.x);
    for (long c = it.e; c < it.end; c += it.stride) {
        const int i = (int)(c % NX);
        const long jk = c / NX;
        const int j = (int)(jk % NY), k = (int)(jk / NY);
        float a, b, own;
        if (WHICH == 0) {  // velU (:80-99)
            own = g.U(i, j, k);
            if (i > 0 && i < g.nx) {
                a = 0.25f * (g.V(i - 1, j, k) + g.V(i - 1, j + 1, k) + g.V(i, j, k) + g.V(i, j + 1, k));
                b = 0.25f * (g.W(i - 1, j, k) + g.W(i - 1, j, k + 1) + g.W(i, j, k) + g.W(i, j, k + 1));
            } else if (i == 0) {
                a = 0.25f * (g.V(i, j, k) + g.V(i, j + 1, k));
                b = 0.25f * (g.W(i, j, k) + g.W(i, j, k + 1));
            } else {
                a = 0.25f * (g.V(i - 1, j, k) + g.V(i - 1, j + 1, k));
                b = 0.25f * (g.W(i - 1, j, k) + g.W(i - 1, j, k + 1));
            }
            out[3 * c] = own;
            out[3 * c + 1] = a;
            out[3 * c + 2] = b;
        } else if (WHICH == 1) {  // velV (:101-120)
            own = g.V(i, j, k);
            if (j > 0 && j < g.ny) {
                a = 0.25f * (g.U(i, j - 1, k) + g.U(i, j, k) + g.U(i + 1, j - 1, k) + g.U(i + 1, j, k));
                b = 0.25f * (g.W(i, j - 1, k) + g.W(i, j - 1, k + 1) + g.W(i, j, k) + g.W(i, j, k + 1));
            } else if (j == 0) {
                a = 0.25f * (g.U(i, j, k) + g.U(i + 1, j, k));
                b = 0.25f * (g.W(i, j, k) + g.W(i, j, k + 1));
            } else {
                a = 0.25f * (g.U(i, j - 1, k) + g.U(i + 1, j - 1, k));
                b = 0.25f * (g.W(i, j - 1, k) + g.W(i, j - 1, k + 1));
            }
            out[3 * c] = a;
            out[3 * c + 1] = own;
            out[3 * c + 2] = b;
        } else {  // velW (:122-141)
            own = g.W(i, j, k);
            if (k > 0 && k < g.nz) {
                a = 0.25f * (g.U(i, j, k - 1) + g.U(i + 1, j, k - 1) + g.U(i, j, k) + g.U(i + 1, j, k));
                b = 0.25f * (g.V(i, j, k - 1) + g.V(i, j + 1, k - 1) + g.V(i, j, k) + g.V(i, j + 1, k));
            } else if (k == 0) {
                a = 0.25f * (g.U(i, j, k) + g.U(i + 1, j, k));
                b = 0.25f * (g.V(i, j, k) + g.V(i, j + 1, k));
            } else {
                a = 0.25f * (g.U(i, j, k - 1) + g.U(i + 1, j, k - 1));
                b = 0.25f * (g.V(i, j, k - 1) + g.V(i, j + 1, k - 1));
            }
            out[3 * c] = a;
            out[3 * c + 1] = b;
            out[3 * c + 2] = own;
        }
    }
}

// Array3D::innerProduct / infiniteNorm (Array3D.h:55-72): per-block FP64
// partials in a fixed order, then one block combines them in index order.
template <bool MAX>
__global__ __launch_bounds__(kBlock) void array_reduce_kernel(const float* __restrict__ a,
                                                              const float* __restrict__ b, long n,
                                                              double* __restrict__ partials) {
    double acc = 0;
    const fs2d::XcdLoop it = fs2d::xcd_loop(n, blockIdx.x, gridDim.x);
    for (long e = it.e; e < it.end; e += it.stride) {
        if (MAX) {
            const double x = fabs((double)a[e]);
            acc = x > acc ? x : acc;
        } else {
            acc += (double)a[e] * (double)b[e];
        }
    }
    const double r = MAX ? fs2d::block_max(acc) : fs2d::block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}
template <bool MAX>
__global__ __launch_bounds__(64) void array_finalize_kernel(const double* __restrict__ partials, int n,
                                                            double* __restrict__ out) {
    const double r = MAX ? fs2d::wave_max_of(partials, n) : fs2d::wave_sum_of(partials, n);
    if (threadIdx.x == 0) *out = r;
}

// Array3D::setMultiplyAdd (Array3D.h:48-53): dst = a + b c, in FP64, rounded once.
__global__ __launch_bounds__(kBlock) void array_multiply_add_kernel(float* dst, const float* a, double b,
                                                                    const float* c, long n) {
    const fs2d::XcdLoop it = fs2d::xcd_loop(n, blockIdx.x, gridDim.x);
    for (long e = it.e; e < it.end; e += it.stride) dst[e] = (float)((double)a[e] + b * (double)c[e]);
}

}  // namespace fs3d
