// fs3d_oracle.cpp -- CPU restatement of the reference's 3D grid classes
// (include/Array3D.h, include/MACGrid3D.h), FP64.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's checks as the
// parity oracle of the 3D grid kernels; the product library never links it.
//
// The reference's 3D classes are not used by its simulator (README.md:10, 25).
// Two quirks are restated as written, behind a switch:
//   * Array3D::triCubic reads its taps as (*this)(zp, yp, xp), i.e. the flat
//     element xp*NY*NX + yp*NX + zp, while clamping xp to [0, NX-1] and zp to
//     [0, NZ-1] (Array3D.h:106-121, :117); for some shapes that element lies
//     past the end of the array, which the oracle reports instead of reading;
//   * the result is clamped to be non-negative (Array3D.h:123).
// mode 0 is the reference; mode 1 ("velocity") reads taps as (xp, yp, zp) and
// returns the signed sum.
//
// Array3D's loops over j run to NY inclusive (Array3D.h:41, 50, 58, 67), which
// reads one row past each slab and past the end of the array; the restatement
// loops over the array's own NY rows.

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace {

// Array3D::operator() (Array3D.h:32-37): element (i, j, k) at [k*NY*NX + j*NX + i]
struct Arr {
    const double* d;
    int NX, NY, NZ;
    long size() const { return (long)NX * NY * NZ; }
    double at(int i, int j, int k) const { return d[(long)k * NY * NX + (long)j * NX + i]; }
};

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Array3D::triCubic (Array3D.h:75-126).  *oob is set when a tap's element lies
// outside the array (mode 0 only); the value is then NaN.
double tricubic(const Arr& a, double px, double py, double pz, int mode, int* oob) {
    const int x = (int)px, y = (int)py, z = (int)pz;
    if (x < 0 || x >= a.NX || y < 0 || y >= a.NY || z < 0 || z >= a.NZ) return 0;
    const double dx = px - (double)x, dy = py - (double)y, dz = pz - (double)z;
    auto cube = [](double t) { return t * t * t; };
    auto sqr = [](double t) { return t * t; };
    double u[4], v[4], w[4], r[4], q[4];
    double vox = 0;
    // Catmull-Rom factors as written (:91-104); 0.5f is exactly 0.5
    u[0] = -0.5 * cube(dx) + sqr(dx) - 0.5 * dx;
    u[1] = 1.5 * cube(dx) - 2.5 * sqr(dx) + 1;
    u[2] = -1.5 * cube(dx) + 2 * sqr(dx) + 0.5 * dx;
    u[3] = 0.5 * cube(dx) - 0.5 * sqr(dx);
    v[0] = -0.5 * cube(dy) + sqr(dy) - 0.5 * dy;
    v[1] = 1.5 * cube(dy) - 2.5 * sqr(dy) + 1;
    v[2] = -1.5 * cube(dy) + 2 * sqr(dy) + 0.5 * dy;
    v[3] = 0.5 * cube(dy) - 0.5 * sqr(dy);
    w[0] = -0.5 * cube(dz) + sqr(dz) - 0.5 * dz;
    w[1] = 1.5 * cube(dz) - 2.5 * sqr(dz) + 1;
    w[2] = -1.5 * cube(dz) + 2 * sqr(dz) + 0.5 * dz;
    w[3] = 0.5 * cube(dz) - 0.5 * sqr(dz);
    for (int k = 0; k < 4; k++) {  // :106-121
        const int zp = clampi(z + k - 1, 0, a.NZ - 1);
        q[k] = 0;
        for (int j = 0; j < 4; j++) {
            const int yp = clampi(y + j - 1, 0, a.NY - 1);
            r[j] = 0;
            for (int i = 0; i < 4; i++) {
                const int xp = clampi(x + i - 1, 0, a.NX - 1);
                double val;
                if (mode == 0) {
                    const long e = (long)xp * a.NY * a.NX + (long)yp * a.NX + zp;  // (*this)(zp, yp, xp)
                    if (e >= a.size()) {
                        *oob = 1;
                        return NAN;
                    }
                    val = a.d[e];
                } else {
                    val = a.at(xp, yp, zp);
                }
                r[j] += u[i] * val;
            }
            q[k] += v[j] * r[j];
        }
        vox += w[k] * q[k];
    }
    if (mode == 0) return vox < 0 ? 0.0 : vox;
    return vox;
}

struct Mac {  // MACGrid3D (MACGrid3D.h:12-23): u (nx+1, ny, nz), v (nx, ny+1, nz), w (nx, ny, nz+1)
    Arr u, v, w;
    int nx, ny, nz;
    double U(int i, int j, int k) const { return u.at(i, j, k); }
    double V(int i, int j, int k) const { return v.at(i, j, k); }
    double W(int i, int j, int k) const { return w.at(i, j, k); }
};

Mac make_mac(const double* u, const double* v, const double* w, int nx, int ny, int nz) {
    return Mac{Arr{u, nx + 1, ny, nz}, Arr{v, nx, ny + 1, nz}, Arr{w, nx, ny, nz + 1}, nx, ny, nz};
}

}  // namespace

extern "C" {

// MACGrid3D::velInterp (MACGrid3D.h:143-163) at n points (grid units, xyz
// interleaved); returns the number of samples whose taps left the array.
long o3_vel_interp(const double* u, const double* v, const double* w, int nx, int ny, int nz, const double* pos,
                   long n, double* out, int mode) {
    const Mac m = make_mac(u, v, w, nx, ny, nz);
    long bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (long p = 0; p < n; p++) {
        const double x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
        int oob = 0;
        out[3 * p] = tricubic(m.u, x, y - 0.5, z - 0.5, mode, &oob);      // velInterpU (:143-147)
        out[3 * p + 1] = tricubic(m.v, x - 0.5, y, z - 0.5, mode, &oob);  // velInterpV (:149-153)
        out[3 * p + 2] = tricubic(m.w, x - 0.5, y - 0.5, z, mode, &oob);  // velInterpW (:155-159)
        bad += oob;
    }
    return bad;
}

// Array3D::triCubic of one array at n points (for the sampler's own tests).
long o3_tricubic(const double* a, int NX, int NY, int NZ, const double* pos, long n, double* out, int mode) {
    const Arr arr{a, NX, NY, NZ};
    long bad = 0;
    for (long p = 0; p < n; p++) {
        int oob = 0;
        out[p] = tricubic(arr, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], mode, &oob);
        bad += oob;
    }
    return bad;
}

// MACGrid3D::velDiv (:165-167) at every cell, left to right as written.
void o3_vel_div(const double* u, const double* v, const double* w, int nx, int ny, int nz, double* out) {
    const Mac m = make_mac(u, v, w, nx, ny, nz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                out[((long)k * ny + j) * nx + i] =
                    m.U(i + 1, j, k) - m.U(i, j, k) + m.V(i, j + 1, k) - m.V(i, j, k) + m.W(i, j, k + 1) - m.W(i, j, k);
}

// MACGrid3D::vel (:73-78) at every cell: out[3 c + {0, 1, 2}].
void o3_vel_cell(const double* u, const double* v, const double* w, int nx, int ny, int nz, double* out) {
    const Mac m = make_mac(u, v, w, nx, ny, nz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double* o = out + 3 * (((long)k * ny + j) * nx + i);
                o[0] = 0.5 * (m.U(i + 1, j, k) + m.U(i, j, k));
                o[1] = 0.5 * (m.V(i, j + 1, k) + m.V(i, j, k));
                o[2] = 0.5 * (m.W(i, j, k + 1) + m.W(i, j, k));
            }
}

// MACGrid3D::velU / velV / velW (:80-141) at every face of component `which`
// (0 u, 1 v, 2 w), in that component's array order: out[3 f + {0, 1, 2}].
void o3_vel_face(const double* u, const double* v, const double* w, int nx, int ny, int nz, int which, double* out) {
    const Mac m = make_mac(u, v, w, nx, ny, nz);
    const int NX = nx + (which == 0), NY = ny + (which == 1), NZ = nz + (which == 2);
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                double* o = out + 3 * (((long)k * NY + j) * NX + i);
                if (which == 0) {  // velU (:80-99)
                    o[0] = m.U(i, j, k);
                    if (i > 0 && i < nx) {
                        o[1] = 0.25 * (m.V(i - 1, j, k) + m.V(i - 1, j + 1, k) + m.V(i, j, k) + m.V(i, j + 1, k));
                        o[2] = 0.25 * (m.W(i - 1, j, k) + m.W(i - 1, j, k + 1) + m.W(i, j, k) + m.W(i, j, k + 1));
                    } else if (i == 0) {
                        o[1] = 0.25 * (m.V(i, j, k) + m.V(i, j + 1, k));
                        o[2] = 0.25 * (m.W(i, j, k) + m.W(i, j, k + 1));
                    } else {
                        o[1] = 0.25 * (m.V(i - 1, j, k) + m.V(i - 1, j + 1, k));
                        o[2] = 0.25 * (m.W(i - 1, j, k) + m.W(i - 1, j, k + 1));
                    }
                } else if (which == 1) {  // velV (:101-120)
                    o[1] = m.V(i, j, k);
                    if (j > 0 && j < ny) {
                        o[0] = 0.25 * (m.U(i, j - 1, k) + m.U(i, j, k) + m.U(i + 1, j - 1, k) + m.U(i + 1, j, k));
                        o[2] = 0.25 * (m.W(i, j - 1, k) + m.W(i, j - 1, k + 1) + m.W(i, j, k) + m.W(i, j, k + 1));
                    } else if (j == 0) {
                        o[0] = 0.25 * (m.U(i, j, k) + m.U(i + 1, j, k));
                        o[2] = 0.25 * (m.W(i, j, k) + m.W(i, j, k + 1));
                    } else {
                        o[0] = 0.25 * (m.U(i, j - 1, k) + m.U(i + 1, j - 1, k));
                        o[2] = 0.25 * (m.W(i, j - 1, k) + m.W(i, j - 1, k + 1));
                    }
                } else {  // velW (:122-141)
                    o[2] = m.W(i, j, k);
                    if (k > 0 && k < nz) {
                        o[0] = 0.25 * (m.U(i, j, k - 1) + m.U(i + 1, j, k - 1) + m.U(i, j, k) + m.U(i + 1, j, k));
                        o[1] = 0.25 * (m.V(i, j, k - 1) + m.V(i, j + 1, k - 1) + m.V(i, j, k) + m.V(i, j + 1, k));
                    } else if (k == 0) {
                        o[0] = 0.25 * (m.U(i, j, k) + m.U(i + 1, j, k));
                        o[1] = 0.25 * (m.V(i, j, k) + m.V(i, j + 1, k));
                    } else {
                        o[0] = 0.25 * (m.U(i, j, k - 1) + m.U(i + 1, j, k - 1));
                        o[1] = 0.25 * (m.V(i, j, k - 1) + m.V(i, j + 1, k - 1));
                    }
                }
            }
}

// Array3D::innerProduct and infiniteNorm (Array3D.h:55-72), k, j, i order.
double o3_inner_product(const double* a, const double* b, long n) {
    double r = 0;
    for (long e = 0; e < n; e++) r += a[e] * b[e];
    return r;
}
double o3_infinite_norm(const double* a, long n) {
    double r = 0;
    for (long e = 0; e < n; e++)
        if (std::fabs(a[e]) > r) r = std::fabs(a[e]);
    return r;
}
// Array3D::setMultiplyAdd (:48-53): dst = a + b * c;  with b = 1, operator+= (:39-46)
void o3_multiply_add(double* dst, const double* a, double b, const double* c, long n) {
    for (long e = 0; e < n; e++) dst[e] = a[e] + b * c[e];
}

}  // extern "C"
