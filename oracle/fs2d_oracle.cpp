// fs2d_oracle.cpp -- CPU restatement of fluid-sim's 2D step (FluidSim2D::runFrame).
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the smoke
// check and bench.py's cpu_baseline leg load it as a checker.  The product
// library (fluid-sim_amd/csrc) never links, calls or falls back to it.
//
// Every function cites the reference lines it restates (paths relative to the
// reference repository root).  Arithmetic is FP64 throughout, written in the
// reference's evaluation order, including the summation orders of its AVX paths
// (the reference builds with USE_AVX_SIMD on, src/CMakeLists.txt:11-15).  The
// file is compiled with -ffp-contract=off so that no a*b+c is fused unless the
// reference fuses it explicitly (_mm256_fmadd_pd in Array2D.h:185).
//
// Extensions that have no reference counterpart (all off by default):
//   * tau / tol / max_iters of the PCG are parameters (reference: 0.999, 1e-12,
//     200; FluidSim2D.cpp:369,429,453);
//   * sl_inplace = 0 selects the out-of-place (Jacobi) semi-Lagrangian update
//     that a parallel GPU performs; 1 is the reference's in-place update (Q1);
//   * level_set = 0 is the "all-fluid" mode: cell flags stay as created;
//   * a passive cell-centred density advected with the same backtrace.
//
// OpenMP pragmas sit only on loops whose iterations are independent, so the
// results do not depend on the thread count.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <vector>
#include <deque>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum : uint8_t { EMPTY = 0, FLUID = 1, SOLID = 2 };  // include/FluidSim2D.h:44-46
enum { MODE_SL = 0, MODE_PIC = 1 };                    // include/FluidSim2D.h:40-42

// Stage ids follow FluidSim2D::StageType (include/FluidSim2D.h:93-97).
enum {
    ST_INIT = 0, ST_LEVELSET = 1, ST_P2G = 2, ST_SL = 3, ST_GRAVITY = 4,
    ST_SOLID = 5, ST_PROJECT = 6, ST_UPDATEVEL = 7, ST_G2P = 8, ST_ADVECT = 9
};

struct V2 { double x, y; };

// Dense row-major array, data[j*NX + i] (include/Array2D.h:41-44, 85-88).
template <class T>
struct Arr {
    int NX = 0, NY = 0;
    std::vector<T> d;
    void create(int nx, int ny) { NX = nx; NY = ny; d.assign((size_t)nx * ny, T(0)); }
    T& operator()(int i, int j) { return d[(size_t)j * NX + i]; }
    const T& operator()(int i, int j) const { return d[(size_t)j * NX + i]; }
    void reset() { std::fill(d.begin(), d.end(), T(0)); }
};
using ArrD = Arr<double>;

inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
inline double maxd(double a, double b) { return a > b ? a : b; }
inline double mind(double a, double b) { return a < b ? a : b; }

// Array2D<double>::triCubic, AVX path (include/Array2D.h:244-327).  Catmull-Rom
// weights (Keys, a = -1/2); taps clamped to the array; row dot products summed
// as (w0*a + w1*b) + (w2*c + w3*d) (the hadd pairs), then the four weighted rows
// as (f0 + f2) + (f1 + f3).
double triCubic(const ArrD& A, double px, double py) {
    int x = (int)px, y = (int)py;
    if (x < 0 || x >= A.NX || y < 0 || y >= A.NY) return 0;
    double dx = px - (double)x, dy = py - (double)y;
    double u[4], v[4];
    u[0] = -0.5 * (dx * dx * dx) + (dx * dx) - 0.5 * dx;
    u[1] = 1.5 * (dx * dx * dx) - 2.5 * (dx * dx) + 1;
    u[2] = -1.5 * (dx * dx * dx) + 2 * (dx * dx) + 0.5 * dx;
    u[3] = 0.5 * (dx * dx * dx) - 0.5 * (dx * dx);
    v[0] = -0.5 * (dy * dy * dy) + (dy * dy) - 0.5 * dy;
    v[1] = 1.5 * (dy * dy * dy) - 2.5 * (dy * dy) + 1;
    v[2] = -1.5 * (dy * dy * dy) + 2 * (dy * dy) + 0.5 * dy;
    v[3] = 0.5 * (dy * dy * dy) - 0.5 * (dy * dy);
    int xs[4], ys[4];
    for (int k = 0; k < 4; k++) {
        xs[k] = clampi(x - 1 + k, 0, A.NX - 1);
        ys[k] = clampi(y - 1 + k, 0, A.NY - 1);
    }
    double f[4];
    for (int r = 0; r < 4; r++) {
        const double* row = &A.d[(size_t)ys[r] * A.NX];
        double p0 = u[0] * row[xs[0]], p1 = u[1] * row[xs[1]];
        double p2 = u[2] * row[xs[2]], p3 = u[3] * row[xs[3]];
        double rs = (p0 + p1) + (p2 + p3);
        f[r] = rs * v[r];
    }
    return (f[0] + f[2]) + (f[1] + f[3]);
}

// Array2D<double>::linearDistribute (include/Array2D.h:361-376).
void linearDistribute(ArrD& A, double px, double py, double value) {
    int ui = (int)px, uj = (int)py;
    double ddx = px - ui, ddy = py - uj;
    int x1 = clampi(ui, 0, A.NX - 1), x2 = clampi(ui + 1, 0, A.NX - 1);
    int y1 = clampi(uj, 0, A.NY - 1), y2 = clampi(uj + 1, 0, A.NY - 1);
    A(x1, y1) += (1 - ddx) * (1 - ddy) * value;
    A(x2, y1) += ddx * (1 - ddy) * value;
    A(x1, y2) += (1 - ddx) * ddy * value;
    A(x2, y2) += ddx * ddy * value;
}

// Array2D<double>::linearExtract (include/Array2D.h:402-420).
double linearExtract(const ArrD& A, double px, double py) {
    int ui = (int)px, uj = (int)py;
    double ddx = px - ui, ddy = py - uj;
    double value = 0.0;
    int x1 = clampi(ui, 0, A.NX - 1), x2 = clampi(ui + 1, 0, A.NX - 1);
    int y1 = clampi(uj, 0, A.NY - 1), y2 = clampi(uj + 1, 0, A.NY - 1);
    value += A(x1, y1) * (1 - ddx) * (1 - ddy);
    value += A(x2, y1) * ddx * (1 - ddy);
    value += A(x1, y2) * (1 - ddx) * ddy;
    value += A(x2, y2) * ddx * ddy;
    return value;
}

// Array2D<double>::innerProduct, AVX path (include/Array2D.h:199-210): four
// partial sums over i mod 4 of the flat array, combined as (l0+l1)+(l2+l3).
double innerProduct(const ArrD& a, const ArrD& b) {
    double l[4] = {0, 0, 0, 0};
    size_t n = a.d.size();
    size_t n4 = n & ~(size_t)3;
    for (size_t i = 0; i < n4; i += 4)
        for (int k = 0; k < 4; k++) l[k] = l[k] + a.d[i + k] * b.d[i + k];
    for (size_t i = n4; i < n; i++) l[i - n4] = l[i - n4] + a.d[i] * b.d[i];
    return (l[0] + l[1]) + (l[2] + l[3]);
}

// Array2D<double>::infiniteNorm (include/Array2D.h:220-242).
double infiniteNorm(const ArrD& a) {
    double m = 0.0;
    for (double x : a.d) m = maxd(m, std::fabs(x));
    return m;
}

// Array2D<double>::extrapolate (include/Array2D.h:552-591): BFS from the known
// entries (mask 0).  Layer 1 = unknown entries with a known 4-neighbour, found in
// row-major order; each dequeued entry becomes the mean of the neighbours whose
// mask is smaller than its own and enqueues its still-unknown neighbours.
void extrapolate(ArrD& A, Arr<uint32_t>& mask) {
    const int NX = A.NX, NY = A.NY;
    std::deque<std::pair<int, int>> q;
    for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) {
            if (mask(i, j) != 0 &&
                ((i > 0 && mask(i - 1, j) == 0) || (i < NX - 1 && mask(i + 1, j) == 0) ||
                 (j > 0 && mask(i, j - 1) == 0) || (j < NY - 1 && mask(i, j + 1) == 0))) {
                mask(i, j) = 1;
                q.emplace_back(i, j);
            }
        }
    while (!q.empty()) {
        int x = q.front().first, y = q.front().second;
        q.pop_front();
        const int nx[4] = {x - 1, x + 1, x, x};
        const int ny[4] = {y, y, y - 1, y + 1};
        double sum = 0.0;
        int count = 0;
        for (int k = 0; k < 4; k++) {
            if (nx[k] < 0 || nx[k] >= NX || ny[k] < 0 || ny[k] >= NY) continue;
            if (mask(nx[k], ny[k]) < mask(x, y)) { sum += A(nx[k], ny[k]); count++; }
        }
        A(x, y) = count == 0 ? 0 : sum / count;
        for (int k = 0; k < 4; k++) {
            if (nx[k] < 0 || nx[k] >= NX || ny[k] < 0 || ny[k] >= NY) continue;
            if (mask(nx[k], ny[k]) == UINT32_MAX) {
                mask(nx[k], ny[k]) = mask(x, y) + 1;
                q.emplace_back(nx[k], ny[k]);
            }
        }
    }
}

// glibc's random() (TYPE_3 additive feedback, r[i] = r[i-3] + r[i-31]) seeded
// like srandom(seed); rand() returns the same stream.  A private copy keeps the
// oracle independent of the process-wide generator.
struct GlibcRand {
    int32_t tbl[31];
    int fptr, rptr;
    void seed(unsigned s) {
        if (s == 0) s = 1;
        tbl[0] = (int32_t)s;
        for (int i = 1; i < 31; i++) {
            int64_t hi = tbl[i - 1] / 127773, lo = tbl[i - 1] % 127773;
            int64_t word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            tbl[i] = (int32_t)word;
        }
        fptr = 3; rptr = 0;
        for (int i = 0; i < 310; i++) next();
    }
    uint32_t next() {
        uint32_t a = (uint32_t)tbl[fptr] + (uint32_t)tbl[rptr];
        tbl[fptr] = (int32_t)a;
        uint32_t result = a >> 1;
        fptr = (fptr + 1) % 31;
        rptr = (rptr + 1) % 31;
        return result;
    }
};

struct Sim {
    // FluidSim2D fields (include/FluidSim2D.h:65-114).
    int sizeX, sizeY, k, ppc;
    ArrD u, v, nu, nv;       // mac, newMac (MACGrid2D.h:12-35)
    ArrD p;
    Arr<uint8_t> cell;
    std::vector<V2> pos, vel;
    ArrD phi;                // waterLevelSet.phi
    V2 gravity;
    double rho, alpha, dt, dx, dr;
    int mode;
    double currentTime = 0.0;
    double waterVolume = 0, totalEnergy = 0, particleTotalEnergy = 0, origWaterVolume = 0;
    std::vector<double> volData, energyData, pEnergyData;
    // options (extensions; defaults are the reference's constants)
    double tau = 0.999, sigma = 0.25, tol = 1e-12;
    int maxIters = 200;
    int slInplace = 1, levelSet = 1, useDensity = 0;
    // Compile-time variants of the reference (src/FluidSim2D.cpp:15-16): with
    // particleCells the cells come from the particles instead of the level
    // set (USE_LEVEL_SET undefined); ghost = 0 drops the ghost-fluid pressure
    // terms (USE_GHOST_PRESSURE undefined).
    int particleCells = 0, ghost = 1;
    // Options of the 512^2 PIC/FLIP dam-break workload that the reference
    // does not have: particles advected by the midpoint rule (RK2) instead of
    // applyAdvection's RK3 (:588-603), and k^2 particles per FLUID cell drawn
    // uniformly over the whole cell instead of one per sub-cell (:52-64).
    int rk2 = 0;
    ArrD dens;
    // diagnostics of the last projection
    int lastIters = 0, hitMax = 0;
    double lastResidual = 0;
    double cflMax = 0;
    int nanFlag = 0;

    // ---- MACGrid2D::velInterpU/V/velInterp (include/MACGrid2D.h:80-98) ----
    double velInterpU(double px, double py) const {
        px /= dx; py /= dx;
        py -= 0.5;
        px = clampd(px, 1e-6, (double)(sizeX - 1) - 1e-6);
        py = clampd(py, 1e-6, (double)(sizeY - 1) - 1e-6);
        return triCubic(u, px, py);
    }
    double velInterpV(double px, double py) const {
        px /= dx; py /= dx;
        px -= 0.5;
        px = clampd(px, 1e-6, (double)(sizeX - 1) - 1e-6);
        py = clampd(py, 1e-6, (double)(sizeY - 1) - 1e-6);
        return triCubic(v, px, py);
    }
    V2 velInterp(V2 q) const { return V2{velInterpU(q.x, q.y), velInterpV(q.x, q.y)}; }
    // Same sampling on the passive density (cell centres): no reference counterpart.
    double densInterp(const ArrD& D, double px, double py) const {
        px /= dx; py /= dx;
        px -= 0.5; py -= 0.5;
        px = clampd(px, 1e-6, (double)(sizeX - 1) - 1e-6);
        py = clampd(py, 1e-6, (double)(sizeY - 1) - 1e-6);
        return triCubic(D, px, py);
    }

    // FluidSim2D::clampPos (src/FluidSim2D.cpp:645-651).
    V2 clampPos(V2 to) const {
        const double offset = 1e-3;
        return V2{clampd(to.x, (1.0 + offset) * dx, (sizeX - 1.0 - offset) * dx),
                  clampd(to.y, (1.0 + offset) * dx, (sizeY - 1.0 - offset) * dx)};
    }

    // Backward Ralston RK3 departure point used by applySemiLagrangianAdvection
    // (src/FluidSim2D.cpp:212-217): x - ((2/9)dt k1 + (3/9)dt k2 + (4/9)dt k3).
    V2 backtrace(V2 xp) const {
        V2 k1 = velInterp(xp);
        V2 k2 = velInterp(V2{xp.x - 0.5 * dt * k1.x, xp.y - 0.5 * dt * k1.y});
        V2 k3 = velInterp(V2{xp.x - 0.75 * dt * k2.x, xp.y - 0.75 * dt * k2.y});
        double sx = (2. / 9.) * dt * k1.x + (3. / 9.) * dt * k2.x;
        double sy = (2. / 9.) * dt * k1.y + (3. / 9.) * dt * k2.y;
        sx = sx + (4. / 9.) * dt * k3.x;
        sy = sy + (4. / 9.) * dt * k3.y;
        return clampPos(V2{xp.x - sx, xp.y - sy});
    }

    // ---- FluidSim2D::create (src/FluidSim2D.cpp:23-76) ----
    // uniform: each particle at ((i + U) dx, (j + U) dx), U = rand()/RAND_MAX,
    // x drawn before y, instead of the jittered sub-cell grid.
    void create(int nx, int ny, int kk, double dt_, double dx_, double rho_, double gx, double gy,
                int mode_, double alpha_, const uint8_t* init, unsigned seed, int uniform = 0) {
        sizeX = nx; sizeY = ny; k = kk; ppc = kk * kk;
        dt = dt_; dx = dx_; dr = 0.9 * dx_; rho = rho_;
        gravity = V2{gx, gy};
        mode = mode_; alpha = alpha_;
        u.create(nx + 1, ny); v.create(nx, ny + 1);
        nu.create(nx + 1, ny); nv.create(nx, ny + 1);
        p.create(nx, ny); cell.create(nx, ny); phi.create(nx, ny);
        dens.create(nx, ny);
        std::memcpy(cell.d.data(), init, (size_t)nx * ny);
        GlibcRand rng;
        rng.seed(seed);
        size_t fluidCount = 0;
        pos.clear();
        // randf (src/FluidSim2D.cpp:19-21): rand()/RAND_MAX * max, max a float.
        auto randf = [&](float mx) { return ((double)rng.next() / (double)2147483647) * mx; };
        float dist = kk > 0 ? 1.0f / kk : 0.0f;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                if (cell(i, j) == FLUID) {
                    fluidCount++;
                    for (int e = 0; e < ppc; e++) {
                        if (uniform) {
                            const double ux = randf(1.0f);
                            const double uy = randf(1.0f);
                            pos.push_back(V2{((double)i + ux) * dx, ((double)j + uy) * dx});
                            continue;
                        }
                        int x = e % kk, y = e / kk;
                        // braced initialiser: x jitter is drawn before y jitter
                        double jx = randf(dist);
                        double jy = randf(dist);
                        float fx = dist * x, fy = dist * y;
                        pos.push_back(V2{((double)i + fx + jx) * dx, ((double)j + fy + jy) * dx});
                    }
                }
        vel.assign(pos.size(), V2{0.0, 0.0});
        waterVolume = fluidCount * dx * dx;
        origWaterVolume = waterVolume;
        if (!levelSet) initStaticPhi();
    }

    // All-fluid mode keeps the created flags; phi is set so that a ghost-fluid
    // face (if any EMPTY cell exists) sits half-way between centres.
    void initStaticPhi() {
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++)
                phi(i, j) = cell(i, j) == EMPTY ? 0.5 * dx : -0.5 * dx;
    }

    // ---- LevelSet::constructFromParticles (src/FluidSim2D.cpp:752-794) ----
    void constructFromParticles() {
        Arr<int64_t> t;
        t.create(sizeX, sizeY);
        for (auto& x : phi.d) x = HUGE_VAL;
        for (auto& x : t.d) x = -1;
        for (size_t e = 0; e < pos.size(); e++) {
            V2 ps = pos[e];
            int x = (int)(ps.x / dx), y = (int)(ps.y / dx);
            double d = std::sqrt((ps.x - x * dx) * (ps.x - x * dx) + (ps.y - y * dx) * (ps.y - y * dx)) - dr;
            if (d < phi(x, y)) { phi(x, y) = d; t(x, y) = (int64_t)e; }
        }
        auto f = [&](int i, int j) {
            const int il[4] = {i - 1, i + 1, i, i};
            const int jl[4] = {j, j, j - 1, j + 1};
            for (int kk = 0; kk < 4; kk++) {
                if (il[kk] < 0 || il[kk] >= sizeX || jl[kk] < 0 || jl[kk] >= sizeY) continue;
                int64_t e = t(il[kk], jl[kk]);
                if (e != -1) {
                    V2 ps = pos[(size_t)e];
                    double d = std::sqrt((ps.x - i * dx) * (ps.x - i * dx) + (ps.y - j * dx) * (ps.y - j * dx)) - dr;
                    if (d < phi(i, j)) { phi(i, j) = d; t(i, j) = e; }
                }
            }
        };
        fastSweepIterate(f);
    }

    // LevelSet::fastSweepIterate (include/FluidSim2D.h:178-203): 4 repeats of the
    // four lexicographic directions.
    template <class F> void fastSweepIterate(F f) {
        for (int rep = 0; rep < 4; rep++) {
            for (int j = 0; j < sizeY; j++) for (int i = 0; i < sizeX; i++) f(i, j);
            for (int j = 0; j < sizeY; j++) for (int i = sizeX; i-- > 0;) f(i, j);
            for (int j = sizeY; j-- > 0;) for (int i = 0; i < sizeX; i++) f(i, j);
            for (int j = sizeY; j-- > 0;) for (int i = sizeX; i-- > 0;) f(i, j);
        }
    }

    // ---- LevelSet::redistance (src/FluidSim2D.cpp:796-920) ----
    void redistance() {
        Arr<uint8_t> isSurface;
        isSurface.create(sizeX, sizeY);
        ArrD oldPhi = phi;
        // Surface marks (:803-834).  The phi writes there store the values phi
        // already holds, so only the marks have an effect.
        for (int j = 0; j < sizeY - 1; j++)
            for (int i = 0; i < sizeX - 1; i++) {
                double phi0 = oldPhi(i, j), phi1 = oldPhi(i + 1, j), phi2 = oldPhi(i, j + 1);
                if (phi0 * phi1 < 0) { isSurface(i, j) = 1; isSurface(i + 1, j) = 1; }
                if (phi0 * phi2 < 0) { isSurface(i, j) = 1; isSurface(i, j + 1) = 1; }
            }
        // Interior fill (:836-842).
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++)
                if (!isSurface(i, j) && phi(i, j) < 0) phi(i, j) = -HUGE_VAL;
        // Eikonal Gauss-Seidel sweeps (:845-902).
        auto upd = [&](int i, int j, int ia, int jb) {
            if (phi(i, j) >= 0) return;
            double a = std::fabs(phi(ia, j)), b = std::fabs(phi(i, jb));
            double phi0 = mind(a, b);
            double phi1 = maxd(a, b);
            double d = phi0 + dx;
            if (d > phi1) d = 0.5 * (phi0 + phi1 + std::sqrt(2 * dx * dx - (phi1 - phi0) * (phi1 - phi0)));
            if (d < -phi(i, j)) phi(i, j) = -d;
        };
        for (int rep = 0; rep < 4; rep++) {
            for (int j = 1; j < sizeY; j++) for (int i = 1; i < sizeX; i++) upd(i, j, i - 1, j - 1);
            for (int j = 1; j < sizeY; j++) for (int i = sizeX - 1; i-- > 0;) upd(i, j, i + 1, j - 1);
            for (int j = sizeY - 1; j-- > 0;) for (int i = 1; i < sizeX; i++) upd(i, j, i - 1, j + 1);
            for (int j = sizeY - 1; j-- > 0;) for (int i = sizeX - 1; i-- > 0;) upd(i, j, i + 1, j + 1);
        }
        // Two smoothing passes phi = min(phi, mean of 4 neighbours) (:904-919).
        for (int rep = 0; rep < 2; rep++) {
            oldPhi = phi;
#pragma omp parallel for
            for (int j = 1; j < sizeY - 1; j++)
                for (int i = 1; i < sizeX - 1; i++) {
                    double avg = 0.25 * (oldPhi(i - 1, j) + oldPhi(i + 1, j) + oldPhi(i, j - 1) + oldPhi(i, j + 1));
                    if (avg < oldPhi(i, j)) phi(i, j) = avg;
                }
        }
    }

    // ---- FluidSim2D::createWaterLevelSet (src/FluidSim2D.cpp:653-732) ----
    void createWaterLevelSet() {
        if (levelSet && particleCells) {
            markCellsFromParticles();
        } else if (levelSet) {
            constructFromParticles();
            redistance();
            for (int j = 0; j < sizeY; j++)
                for (int i = 0; i < sizeX; i++)
                    if (cell(i, j) != SOLID) cell(i, j) = phi(i, j) < 0.0 ? FLUID : EMPTY;
        }
        computeDiagnostics();
    }

    // createWaterLevelSet without USE_LEVEL_SET (src/FluidSim2D.cpp:665-706):
    // FLUID cells become EMPTY; every particle marks the cell (int)x, (int)y of
    // its position *in metres* (the reference does not divide by dx); then, in
    // serial order (i outer, j inner, cells updated in place), an EMPTY cell
    // with 4 FLUID neighbours, or 3 FLUID and 1 SOLID, becomes FLUID.  Marks
    // outside the grid (possible only when the domain exceeds size metres) are
    // skipped here, where the reference would write out of bounds.
    void markCellsFromParticles() {
        for (auto& c : cell.d) if (c == FLUID) c = EMPTY;
        for (auto& ps : pos) {
            const int x = (int)ps.x, y = (int)ps.y;
            if (x < 0 || x >= sizeX || y < 0 || y >= sizeY) continue;
            if (cell(x, y) == EMPTY) cell(x, y) = FLUID;
        }
        for (int i = 1; i < sizeX - 1; i++)
            for (int j = 1; j < sizeY - 1; j++) {
                if (cell(i, j) != EMPTY) continue;
                int fluid = 0, solid = 0;
                const uint8_t nb[4] = {cell(i - 1, j), cell(i + 1, j), cell(i, j - 1), cell(i, j + 1)};
                for (uint8_t c : nb) {
                    if (c == FLUID) fluid++;
                    else if (c == SOLID) solid++;
                }
                if (fluid == 4 || (fluid == 3 && solid == 1)) cell(i, j) = FLUID;
            }
    }

    // Diagnostics (src/FluidSim2D.cpp:709-731).
    void computeDiagnostics() {
        waterVolume = 0; totalEnergy = 0; particleTotalEnergy = 0;
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++)
                if (cell(i, j) == FLUID) {
                    waterVolume += dx * dx;
                    V2 vv = velInterp(V2{i * dx, j * dx});
                    totalEnergy += 0.5 * (rho * dx * dx) * (vv.x * vv.x + vv.y * vv.y);
                    totalEnergy -= (rho * dx * dx) * (gravity.x * (i * dx) + gravity.y * (j * dx));
                }
        for (size_t e = 0; e < pos.size(); e++) {
            V2 ps = pos[e], vv = vel[e];
            particleTotalEnergy += 0.5 * (rho * dx * dx / (k * k)) * (vv.x * vv.x + vv.y * vv.y);
            particleTotalEnergy -= (rho * dx * dx / (k * k)) * (gravity.x * ps.x + gravity.y * ps.y);
        }
        volData.push_back(waterVolume);
        energyData.push_back(totalEnergy);
        pEnergyData.push_back(particleTotalEnergy);
    }

    // ---- FluidSim2D::transferVelocityToGrid (src/FluidSim2D.cpp:144-204) ----
    void transferVelocityToGrid() {
        u.reset(); v.reset();
        ArrD udiv, vdiv;
        udiv.create(sizeX + 1, sizeY); vdiv.create(sizeX, sizeY + 1);
        for (size_t e = 0; e < pos.size(); e++) {
            V2 xp = pos[e], vp = vel[e];
            double ux = xp.x / dx, uy = xp.y / dx - 0.5;
            linearDistribute(u, ux, uy, vp.x);
            linearDistribute(udiv, ux, uy, 1.0);
            double vx = xp.x / dx - 0.5, vy = xp.y / dx;
            linearDistribute(v, vx, vy, vp.y);
            linearDistribute(vdiv, vx, vy, 1.0);
        }
        for (size_t n = 0; n < u.d.size(); n++) if (udiv.d[n] > 0) u.d[n] /= udiv.d[n];
        for (size_t n = 0; n < v.d.size(); n++) if (vdiv.d[n] > 0) v.d[n] /= vdiv.d[n];
        Arr<uint32_t> uf, vf;
        uf.create(sizeX + 1, sizeY); vf.create(sizeX, sizeY + 1);
        for (size_t n = 0; n < u.d.size(); n++) uf.d[n] = u.d[n] == 0.0 ? UINT32_MAX : 0;
        for (size_t n = 0; n < v.d.size(); n++) vf.d[n] = v.d[n] == 0.0 ? UINT32_MAX : 0;
        extrapolate(u, uf);
        extrapolate(v, vf);
    }

    // ---- FluidSim2D::applySemiLagrangianAdvection (src/FluidSim2D.cpp:206-235) ----
    // sl_inplace = 1 reproduces the reference's in-place update: the u loop
    // overwrites mac.u while later faces sample it, and the v loop sees the new u.
    // sl_inplace = 0 samples the field as it was when the stage began.
    void applySemiLagrangianAdvection() {
        if (useDensity) advectDensity();
        if (slInplace) {
            for (int j = 0; j < sizeY; j++)
                for (int i = 0; i < sizeX + 1; i++) {
                    V2 xn = backtrace(V2{(double)i * dx, ((double)j + 0.5) * dx});
                    u(i, j) = velInterpU(xn.x, xn.y);
                }
            for (int j = 0; j < sizeY + 1; j++)
                for (int i = 0; i < sizeX; i++) {
                    V2 xn = backtrace(V2{((double)i + 0.5) * dx, (double)j * dx});
                    v(i, j) = velInterpV(xn.x, xn.y);
                }
        } else {
            ArrD un = u, vn = v;
#pragma omp parallel for
            for (int j = 0; j < sizeY; j++)
                for (int i = 0; i < sizeX + 1; i++) {
                    V2 xn = backtrace(V2{(double)i * dx, ((double)j + 0.5) * dx});
                    un(i, j) = velInterpU(xn.x, xn.y);
                }
#pragma omp parallel for
            for (int j = 0; j < sizeY + 1; j++)
                for (int i = 0; i < sizeX; i++) {
                    V2 xn = backtrace(V2{((double)i + 0.5) * dx, (double)j * dx});
                    vn(i, j) = velInterpV(xn.x, xn.y);
                }
            u = un; v = vn;
        }
    }

    // Passive density (no reference counterpart): cell centres traced back with
    // the same RK3 departure point, sampled by the same Catmull-Rom kernel, using
    // the velocity field as it was when the stage began.
    void advectDensity() {
        ArrD dn = dens;
#pragma omp parallel for
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                V2 xn = backtrace(V2{((double)i + 0.5) * dx, ((double)j + 0.5) * dx});
                dn(i, j) = densInterp(dens, xn.x, xn.y);
            }
        dens = dn;
    }

    // ---- FluidSim2D::applyGravity (src/FluidSim2D.cpp:237-250) ----
    void applyGravity() {
        for (auto& x : u.d) x += dt * gravity.x;
        for (auto& x : v.d) x += dt * gravity.y;
    }

    // ---- FluidSim2D::applyProjection (src/FluidSim2D.cpp:252-467) ----
    ArrD Adiag, Ax, Ay, rhs, precon;
    void assemble() {
        Adiag.create(sizeX, sizeY); Ax.create(sizeX, sizeY); Ay.create(sizeX, sizeY);
        rhs.create(sizeX, sizeY);
        double scaleA = dt / (rho * dx * dx);
        // ghost-fluid matrix (:262-303)
#pragma omp parallel for
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                if (cell(i, j) != FLUID) continue;
                if (!ghost) {
                    // without USE_GHOST_PRESSURE (:304-331): only the +x and +y
                    // EMPTY neighbours add to the diagonal
                    if (cell(i - 1, j) == FLUID) Adiag(i, j) += scaleA;
                    if (cell(i + 1, j) == FLUID) { Adiag(i, j) += scaleA; Ax(i, j) = -scaleA; }
                    else if (cell(i + 1, j) == EMPTY) Adiag(i, j) += scaleA;
                    if (cell(i, j - 1) == FLUID) Adiag(i, j) += scaleA;
                    if (cell(i, j + 1) == FLUID) { Adiag(i, j) += scaleA; Ay(i, j) = -scaleA; }
                    else if (cell(i, j + 1) == EMPTY) Adiag(i, j) += scaleA;
                    continue;
                }
                if (cell(i - 1, j) == FLUID) Adiag(i, j) += scaleA;
                else if (cell(i - 1, j) == EMPTY) Adiag(i, j) -= scaleA * maxd(phi(i - 1, j) / phi(i, j), -1e3);
                if (cell(i + 1, j) == FLUID) { Adiag(i, j) += scaleA; Ax(i, j) = -scaleA; }
                else if (cell(i + 1, j) == EMPTY) Adiag(i, j) += scaleA * (1 - maxd(phi(i + 1, j) / phi(i, j), -1e3));
                if (cell(i, j - 1) == FLUID) Adiag(i, j) += scaleA;
                else if (cell(i, j - 1) == EMPTY) Adiag(i, j) -= scaleA * maxd(phi(i, j - 1) / phi(i, j), -1e3);
                if (cell(i, j + 1) == FLUID) { Adiag(i, j) += scaleA; Ay(i, j) = -scaleA; }
                else if (cell(i, j + 1) == EMPTY) Adiag(i, j) += scaleA * (1 - maxd(phi(i, j + 1) / phi(i, j), -1e3));
            }
        // rhs = -div u with solid-face corrections (:338-362)
        double scale = 1.0 / dx;
#pragma omp parallel for
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                if (cell(i, j) != FLUID) continue;
                rhs(i, j) = -scale * (u(i + 1, j) - u(i, j) + v(i, j + 1) - v(i, j));
                if (cell(i - 1, j) == SOLID) rhs(i, j) -= scale * (u(i, j) - 0);
                if (cell(i + 1, j) == SOLID) rhs(i, j) += scale * (u(i + 1, j) - 0);
                if (cell(i, j - 1) == SOLID) rhs(i, j) -= scale * (v(i, j) - 0);
                if (cell(i, j + 1) == SOLID) rhs(i, j) += scale * (v(i, j + 1) - 0);
            }
    }

    // MIC(0) factor (:364-388); tau and sigma are parameters here.
    void factor() {
        precon.create(sizeX, sizeY);
        for (int j = 1; j < sizeY; j++)
            for (int i = 1; i < sizeX; i++) {
                if (cell(i, j) != FLUID) continue;
                double a = Ax(i - 1, j) * precon(i - 1, j);
                double b = Ay(i, j - 1) * precon(i, j - 1);
                // SQUARE(x) is (x)*(x) without outer parentheses, so the tau
                // terms associate as ((Ax*Ay)*P)*P.
                double e = Adiag(i, j) - a * a - b * b -
                           tau * (Ax(i - 1, j) * Ay(i - 1, j) * precon(i - 1, j) * precon(i - 1, j) +
                                  Ay(i, j - 1) * Ax(i, j - 1) * precon(i, j - 1) * precon(i, j - 1));
                if (e < sigma * Adiag(i, j)) e = Adiag(i, j);
                precon(i, j) = 1 / std::sqrt(e);
            }
    }

    // applyPreconditioner (:397-421): z = (L L^T)^-1 r with a zeroed q.
    void applyPreconditioner(const ArrD& r, ArrD& z) const {
        ArrD q;
        q.create(sizeX, sizeY);
        for (int j = 1; j < sizeY; j++)
            for (int i = 1; i < sizeX; i++) {
                if (cell(i, j) != FLUID) continue;
                double t = r(i, j) - Ax(i - 1, j) * precon(i - 1, j) * q(i - 1, j) -
                           Ay(i, j - 1) * precon(i, j - 1) * q(i, j - 1);
                q(i, j) = t * precon(i, j);
            }
        for (int j = sizeY - 1; j-- > 0;)
            for (int i = sizeX - 1; i-- > 0;) {
                if (cell(i, j) != FLUID) continue;
                double t = q(i, j) - Ax(i, j) * precon(i, j) * z(i + 1, j) - Ay(i, j) * precon(i, j) * z(i, j + 1);
                z(i, j) = t * precon(i, j);
            }
    }

    void applyA(const ArrD& s, ArrD& z) const {
#pragma omp parallel for
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++)
                if (cell(i, j) == FLUID)
                    z(i, j) = Adiag(i, j) * s(i, j) + Ax(i - 1, j) * s(i - 1, j) + Ax(i, j) * s(i + 1, j) +
                              Ay(i, j - 1) * s(i, j - 1) + Ay(i, j) * s(i, j + 1);
    }

    // PCG (:423-466).  fma() is the reference's _mm256_fmadd_pd (Array2D.h:185).
    void applyProjection() {
        assemble();
        factor();
        pcgSolve(rhs, p);
    }

    // The PCG loop of applyProjection (:423-466) on the assembled and factored
    // system: p = A^-1 b, starting from p = 0.
    void pcgSolve(const ArrD& rhs, ArrD& p) {
        ArrD r, z, s;
        r.create(sizeX, sizeY); z.create(sizeX, sizeY); s.create(sizeX, sizeY);
        p.create(sizeX, sizeY);
        r = rhs;
        applyPreconditioner(r, z);
        s = z;
        double sig = innerProduct(z, r);
        int iter = 0;
        lastResidual = infiniteNorm(r);
        while (iter < maxIters) {
            applyA(s, z);
            double rhsNorm = infiniteNorm(rhs);
            if (rhsNorm <= 1e-12) break;
            double alpha_ = sig / innerProduct(z, s);
            for (size_t n = 0; n < p.d.size(); n++) p.d[n] = std::fma(alpha_, s.d[n], p.d[n]);
            for (size_t n = 0; n < r.d.size(); n++) r.d[n] = std::fma(-alpha_, z.d[n], r.d[n]);
            lastResidual = infiniteNorm(r);
            if (lastResidual <= tol * rhsNorm) break;
            applyPreconditioner(r, z);
            double sigNew = innerProduct(z, r);
            double beta = sigNew / sig;
            for (size_t n = 0; n < s.d.size(); n++) s.d[n] = std::fma(beta, s.d[n], z.d[n]);
            sig = sigNew;
            iter++;
        }
        lastIters = iter;
        hitMax = iter == maxIters;
    }

    // ---- FluidSim2D::updateVelocity (src/FluidSim2D.cpp:469-550) ----
    void updateVelocity() {
        Arr<uint32_t> du, dv;
        du.create(sizeX + 1, sizeY); dv.create(sizeX, sizeY + 1);
        nu = u; nv = v;
        double scale = dt / (rho * dx);
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                if ((i > 0 && cell(i - 1, j) == FLUID) || cell(i, j) == FLUID) {
                    if ((i == 0 || cell(i - 1, j) == SOLID) || cell(i, j) == SOLID) nu(i, j) = 0;
                    else if (!ghost) nu(i, j) -= scale * (p(i, j) - p(i - 1, j));  // i > 0 here
                    else if (i > 0 && cell(i - 1, j) == EMPTY)
                        nu(i, j) -= scale * (1 - maxd(phi(i - 1, j) / phi(i, j), -1e3)) * p(i, j);
                    else if (cell(i, j) == EMPTY)
                        nu(i, j) -= scale * (maxd(phi(i, j) / phi(i - 1, j), -1e3) - 1) * p(i - 1, j);
                    else if (i > 0) nu(i, j) -= scale * (p(i, j) - p(i - 1, j));
                    else nu(i, j) = 0;
                } else du(i, j) = UINT32_MAX;
                if ((j > 0 && cell(i, j - 1) == FLUID) || cell(i, j) == FLUID) {
                    if ((j == 0 || cell(i, j - 1) == SOLID) || cell(i, j) == SOLID) nv(i, j) = 0;
                    else if (!ghost) nv(i, j) -= scale * (p(i, j) - p(i, j - 1));  // j > 0 here
                    else if (j > 0 && cell(i, j - 1) == EMPTY)
                        nv(i, j) -= scale * (1 - maxd(phi(i, j - 1) / phi(i, j), -1e3)) * p(i, j);
                    else if (cell(i, j) == EMPTY)
                        nv(i, j) -= scale * (maxd(phi(i, j) / phi(i, j - 1), -1e3) - 1) * p(i, j - 1);
                    else if (j > 0) nv(i, j) -= scale * (p(i, j) - p(i, j - 1));
                    else nv(i, j) = 0;
                } else dv(i, j) = UINT32_MAX;
            }
        extrapolate(nu, du);
        extrapolate(nv, dv);
        if (mode == MODE_SL) { u = nu; v = nv; }
    }

    // ---- FluidSim2D::updateParticleVelocities (src/FluidSim2D.cpp:552-568) ----
    void updateParticleVelocities() {
        ArrD ud = nu, vd = nv;
        for (size_t n = 0; n < ud.d.size(); n++) ud.d[n] = nu.d[n] - u.d[n];
        for (size_t n = 0; n < vd.d.size(); n++) vd.d[n] = nv.d[n] - v.d[n];
#pragma omp parallel for
        for (long e = 0; e < (long)pos.size(); e++) {
            double ux = pos[e].x / dx, uy = pos[e].y / dx - 0.5;
            double vx = pos[e].x / dx - 0.5, vy = pos[e].y / dx;
            V2 pic{linearExtract(nu, ux, uy), linearExtract(nv, vx, vy)};
            V2 flip{vel[e].x + linearExtract(ud, ux, uy), vel[e].y + linearExtract(vd, vx, vy)};
            vel[e] = V2{alpha * pic.x + (1 - alpha) * flip.x, alpha * pic.y + (1 - alpha) * flip.y};
        }
        u = nu; v = nv;
    }

    // ---- FluidSim2D::applyAdvection (src/FluidSim2D.cpp:570-604) ----
    void applyAdvection() {
        double cmax = 0.0;
        for (size_t e = 0; e < vel.size(); e++) {
            double c = (vel[e].x + vel[e].y) * dt / dx;
            if (c > cmax) cmax = c;
        }
        cflMax = cmax;
        int nan = 0;
#pragma omp parallel for reduction(| : nan)
        for (long e = 0; e < (long)pos.size(); e++) {
            V2 ps = pos[e];
            V2 k1 = velInterp(ps);
            V2 k2 = velInterp(V2{ps.x + 0.5 * dt * k1.x, ps.y + 0.5 * dt * k1.y});
            V2 nx;
            if (rk2) {
                // midpoint rule: the velocity at the half step, for the whole step
                nx = V2{ps.x + dt * k2.x, ps.y + dt * k2.y};
            } else {
                V2 k3 = velInterp(V2{ps.x + 0.75 * dt * k2.x, ps.y + 0.75 * dt * k2.y});
                nx = V2{ps.x + (2. / 9.) * dt * k1.x, ps.y + (2. / 9.) * dt * k1.y};
                nx = V2{nx.x + (3. / 9.) * dt * k2.x, nx.y + (3. / 9.) * dt * k2.y};
                nx = V2{nx.x + (4. / 9.) * dt * k3.x, nx.y + (4. / 9.) * dt * k3.y};
            }
            nx = clampPos(nx);
            if (std::isnan(ps.x) || std::isnan(ps.y)) nan = 1;
            pos[e] = nx;
        }
        nanFlag = nan;
    }

    // ======================================================================
    // Adjoint of one step of the all-fluid semi-Lagrangian mode (no reference
    // counterpart).  The step is: out-of-place SL advection of u, v and the
    // density, applyGravity, applyProjection, updateVelocity with its
    // extrapolation, on static cell flags without EMPTY cells (so A is the
    // symmetric Laplacian) and without particles.  adjointStep() maps the
    // adjoints of the step's outputs to the adjoints of its inputs at the
    // current state, which it leaves unchanged.  FP64 throughout; the solve
    // uses the same PCG as the forward step.
    // ======================================================================

    // Catmull-Rom weights of triCubic and their derivatives in the offset.
    static void crWeights(double f, double w[4], double dw[4]) {
        w[0] = -0.5 * (f * f * f) + (f * f) - 0.5 * f;
        w[1] = 1.5 * (f * f * f) - 2.5 * (f * f) + 1;
        w[2] = -1.5 * (f * f * f) + 2 * (f * f) + 0.5 * f;
        w[3] = 0.5 * (f * f * f) - 0.5 * (f * f);
        dw[0] = -1.5 * (f * f) + 2 * f - 0.5;
        dw[1] = 4.5 * (f * f) - 5 * f;
        dw[2] = -4.5 * (f * f) + 4 * f + 0.5;
        dw[3] = 1.5 * (f * f) - f;
    }

    // Transpose of triCubic(A, px, py) applied to g: G(tap) += g * weight.
    // Adds g * d triCubic / d(px, py) to (bx, by).
    static void triCubicAdj(const ArrD& A, ArrD& G, double px, double py, double g, double& bx, double& by) {
        int x = (int)px, y = (int)py;
        if (x < 0 || x >= A.NX || y < 0 || y >= A.NY) return;
        double wu[4], du[4], wv[4], dv[4];
        crWeights(px - (double)x, wu, du);
        crWeights(py - (double)y, wv, dv);
        double sx = 0, sy = 0;
        for (int r = 0; r < 4; r++) {
            const int yr = clampi(y - 1 + r, 0, A.NY - 1);
            double row = 0, drow = 0;
            for (int c = 0; c < 4; c++) {
                const int xc = clampi(x - 1 + c, 0, A.NX - 1);
                const double a = A(xc, yr);
                row += wu[c] * a;
                drow += du[c] * a;
                G(xc, yr) += g * wu[c] * wv[r];
            }
            sx += wv[r] * drow;
            sy += dv[r] * row;
        }
        bx += g * sx;
        by += g * sy;
    }

    // d clampd(x, lo, hi) / dx (one-sided 1 on the bounds, where clampd returns x).
    static double clampGrad(double x, double lo, double hi) { return (x < lo || x > hi) ? 0.0 : 1.0; }

    // Transposes of velInterpU / velInterpV / densInterp: scatter g into the
    // field's adjoint and add g * d(sample)/d(position) to (bx, by).
    void interpAdj(int which, const ArrD& A, ArrD& G, double px, double py, double g, double& bx,
                   double& by) const {
        double gx = px / dx, gy = py / dx;
        if (which == 0) gy -= 0.5;
        else if (which == 1) gx -= 0.5;
        else { gx -= 0.5; gy -= 0.5; }
        const double hx = (double)(sizeX - 1) - 1e-6, hy = (double)(sizeY - 1) - 1e-6;
        double ax = 0, ay = 0;
        triCubicAdj(A, G, clampd(gx, 1e-6, hx), clampd(gy, 1e-6, hy), g, ax, ay);
        bx += clampGrad(gx, 1e-6, hx) * ax / dx;
        by += clampGrad(gy, 1e-6, hy) * ay / dx;
    }

    // Transpose of backtrace(xp) at the current velocity: Xb is the adjoint of
    // the departure point; the velocity samples of the three RK stages scatter
    // into (ub, vb).
    void backtraceAdj(V2 xp, V2 Xb, ArrD& ub, ArrD& vb) const {
        V2 k1 = velInterp(xp);
        V2 x2{xp.x - 0.5 * dt * k1.x, xp.y - 0.5 * dt * k1.y};
        V2 k2 = velInterp(x2);
        V2 x3{xp.x - 0.75 * dt * k2.x, xp.y - 0.75 * dt * k2.y};
        V2 k3 = velInterp(x3);
        double sx = (2. / 9.) * dt * k1.x + (3. / 9.) * dt * k2.x;
        double sy = (2. / 9.) * dt * k1.y + (3. / 9.) * dt * k2.y;
        sx = sx + (4. / 9.) * dt * k3.x;
        sy = sy + (4. / 9.) * dt * k3.y;
        const double off = 1e-3;
        Xb.x *= clampGrad(xp.x - sx, (1.0 + off) * dx, (sizeX - 1.0 - off) * dx);
        Xb.y *= clampGrad(xp.y - sy, (1.0 + off) * dx, (sizeY - 1.0 - off) * dx);
        V2 k1b{-(2. / 9.) * dt * Xb.x, -(2. / 9.) * dt * Xb.y};
        V2 k2b{-(3. / 9.) * dt * Xb.x, -(3. / 9.) * dt * Xb.y};
        V2 k3b{-(4. / 9.) * dt * Xb.x, -(4. / 9.) * dt * Xb.y};
        V2 x3b{0, 0}, x2b{0, 0}, x1b{0, 0};
        interpAdj(0, u, ub, x3.x, x3.y, k3b.x, x3b.x, x3b.y);
        interpAdj(1, v, vb, x3.x, x3.y, k3b.y, x3b.x, x3b.y);
        k2b.x += -0.75 * dt * x3b.x;
        k2b.y += -0.75 * dt * x3b.y;
        interpAdj(0, u, ub, x2.x, x2.y, k2b.x, x2b.x, x2b.y);
        interpAdj(1, v, vb, x2.x, x2.y, k2b.y, x2b.x, x2b.y);
        k1b.x += -0.5 * dt * x2b.x;
        k1b.y += -0.5 * dt * x2b.y;
        interpAdj(0, u, ub, xp.x, xp.y, k1b.x, x1b.x, x1b.y);
        interpAdj(1, v, vb, xp.x, xp.y, k1b.y, x1b.x, x1b.y);
    }

    // Transpose of the out-of-place SL stage (u, v and, with useDensity, the
    // density): (ub1, vb1, db1) are the adjoints of its outputs.
    void slAdjoint(const ArrD& ub1, const ArrD& vb1, const ArrD& db1, ArrD& ub, ArrD& vb, ArrD& db) const {
        ub.create(sizeX + 1, sizeY); vb.create(sizeX, sizeY + 1); db.create(sizeX, sizeY);
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX + 1; i++) {
                const double g = ub1(i, j);
                if (g == 0.0) continue;
                V2 xp{(double)i * dx, ((double)j + 0.5) * dx};
                V2 X = backtrace(xp), Xb{0, 0};
                interpAdj(0, u, ub, X.x, X.y, g, Xb.x, Xb.y);
                backtraceAdj(xp, Xb, ub, vb);
            }
        for (int j = 0; j < sizeY + 1; j++)
            for (int i = 0; i < sizeX; i++) {
                const double g = vb1(i, j);
                if (g == 0.0) continue;
                V2 xp{((double)i + 0.5) * dx, (double)j * dx};
                V2 X = backtrace(xp), Xb{0, 0};
                interpAdj(1, v, vb, X.x, X.y, g, Xb.x, Xb.y);
                backtraceAdj(xp, Xb, ub, vb);
            }
        if (!useDensity) return;
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                const double g = db1(i, j);
                if (g == 0.0) continue;
                V2 xp{((double)i + 0.5) * dx, ((double)j + 0.5) * dx};
                V2 X = backtrace(xp), Xb{0, 0};
                interpAdj(2, dens, db, X.x, X.y, g, Xb.x, Xb.y);
                backtraceAdj(xp, Xb, ub, vb);
            }
    }

    // Transpose of extrapolate(A, mask): the BFS's dequeue order reversed; an
    // extrapolated entry hands its adjoint in equal parts to the neighbours it
    // averaged and is zeroed, since extrapolation overwrote its input value.
    static void extrapolateAdj(ArrD& G, Arr<uint32_t> mask) {
        const int NX = G.NX, NY = G.NY;
        std::vector<std::pair<int, int>> order;
        std::deque<std::pair<int, int>> q;
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++)
                if (mask(i, j) != 0 &&
                    ((i > 0 && mask(i - 1, j) == 0) || (i < NX - 1 && mask(i + 1, j) == 0) ||
                     (j > 0 && mask(i, j - 1) == 0) || (j < NY - 1 && mask(i, j + 1) == 0))) {
                    mask(i, j) = 1;
                    q.emplace_back(i, j);
                }
        while (!q.empty()) {
            auto f = q.front();
            q.pop_front();
            order.push_back(f);
            const int xs[4] = {f.first - 1, f.first + 1, f.first, f.first};
            const int ys[4] = {f.second, f.second, f.second - 1, f.second + 1};
            for (int k = 0; k < 4; k++) {
                if (xs[k] < 0 || xs[k] >= NX || ys[k] < 0 || ys[k] >= NY) continue;
                if (mask(xs[k], ys[k]) == UINT32_MAX) {
                    mask(xs[k], ys[k]) = mask(f.first, f.second) + 1;
                    q.emplace_back(xs[k], ys[k]);
                }
            }
        }
        for (size_t n = order.size(); n-- > 0;) {
            const int x = order[n].first, y = order[n].second;
            const int xs[4] = {x - 1, x + 1, x, x};
            const int ys[4] = {y, y, y - 1, y + 1};
            int cnt = 0;
            for (int k = 0; k < 4; k++)
                if (xs[k] >= 0 && xs[k] < NX && ys[k] >= 0 && ys[k] < NY && mask(xs[k], ys[k]) < mask(x, y)) cnt++;
            const double g = G(x, y);
            G(x, y) = 0;
            if (cnt == 0) continue;
            for (int k = 0; k < 4; k++)
                if (xs[k] >= 0 && xs[k] < NX && ys[k] >= 0 && ys[k] < NY && mask(xs[k], ys[k]) < mask(x, y))
                    G(xs[k], ys[k]) += g / cnt;
        }
        // entries the BFS never reached keep mask ~0 and their (unchanged) value
    }

    // Transpose of applyProjection + updateVelocity (with extrapolation) for
    // static flags without EMPTY cells: (ub, vb) enter as the adjoints of the
    // updated velocity and leave as the adjoints of the velocity the
    // projection started from.
    void projectAdjoint(ArrD& ub, ArrD& vb) {
        Arr<uint32_t> du, dv;
        du.create(sizeX + 1, sizeY); dv.create(sizeX, sizeY + 1);
        auto fluidAdjU = [&](int i, int j) { return (i > 0 && cell(i - 1, j) == FLUID) || cell(i, j) == FLUID; };
        auto fluidAdjV = [&](int i, int j) { return (j > 0 && cell(i, j - 1) == FLUID) || cell(i, j) == FLUID; };
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                if (!fluidAdjU(i, j)) du(i, j) = UINT32_MAX;
                if (!fluidAdjV(i, j)) dv(i, j) = UINT32_MAX;
            }
        extrapolateAdj(ub, du);
        extrapolateAdj(vb, dv);
        // updateVelocity: nu = u - scale (p(i) - p(i-1)) between two FLUID
        // cells, 0 next to SOLID; u(sizeX, j) and v(i, sizeY) pass through.
        const double scale = dt / (rho * dx);
        ArrD pb;
        pb.create(sizeX, sizeY);
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                if (fluidAdjU(i, j)) {
                    if ((i == 0 || cell(i - 1, j) == SOLID) || cell(i, j) == SOLID) ub(i, j) = 0;
                    else { pb(i, j) -= scale * ub(i, j); pb(i - 1, j) += scale * ub(i, j); }
                }
                if (fluidAdjV(i, j)) {
                    if ((j == 0 || cell(i, j - 1) == SOLID) || cell(i, j) == SOLID) vb(i, j) = 0;
                    else { pb(i, j) -= scale * vb(i, j); pb(i, j - 1) += scale * vb(i, j); }
                }
            }
        // p = A^-1 rhs with A symmetric: the adjoint solve uses A itself.
        assemble();
        factor();
        ArrD lam;
        pcgSolve(pb, lam);
        // rhs = -(1/dx) div u, with the faces next to SOLID cancelled (assemble()).
        const double sr = 1.0 / dx;
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) {
                if (cell(i, j) != FLUID) continue;
                const double l = lam(i, j);
                if (cell(i - 1, j) != SOLID) ub(i, j) += sr * l;
                if (cell(i + 1, j) != SOLID) ub(i + 1, j) -= sr * l;
                if (cell(i, j - 1) != SOLID) vb(i, j) += sr * l;
                if (cell(i, j + 1) != SOLID) vb(i, j + 1) -= sr * l;
            }
    }

    // True when the adjoint's assumptions hold: SL mode, static flags, no
    // EMPTY cell, no particles, out-of-place advection.
    bool adjointSupported() const {
        if (mode != MODE_SL || levelSet || slInplace || !pos.empty()) return false;
        for (auto c : cell.d) if (c == EMPTY) return false;
        return true;
    }

    // Adjoint of runFrame() at the current state (left unchanged; the solver
    // overwrites the assembled system and lastIters).  Inputs are the
    // adjoints of u, v, density after the step; outputs those before it.
    void adjointStep(const ArrD& ubOut, const ArrD& vbOut, const ArrD& dbOut, ArrD& ubIn, ArrD& vbIn,
                     ArrD& dbIn) {
        ArrD ub = ubOut, vb = vbOut;
        projectAdjoint(ub, vb);
        // applyGravity adds a constant: identity on the adjoint.
        slAdjoint(ub, vb, dbOut, ubIn, vbIn, dbIn);
    }

    void runStage(int st) {
        switch (st) {
            case ST_LEVELSET: createWaterLevelSet(); break;
            case ST_P2G: transferVelocityToGrid(); break;
            case ST_SL: applySemiLagrangianAdvection(); break;
            case ST_GRAVITY: applyGravity(); break;
            case ST_SOLID: break;  // createSolidLevelSet is a no-op (:734-736)
            case ST_PROJECT: applyProjection(); break;
            case ST_UPDATEVEL: updateVelocity(); break;
            case ST_G2P: updateParticleVelocities(); break;
            case ST_ADVECT: applyAdvection(); break;
            default: break;
        }
    }

    // ---- FluidSim2D::runFrame (src/FluidSim2D.cpp:94-138) ----
    void runFrame() {
        if (mode == MODE_SL) {
            const int st[] = {ST_LEVELSET, ST_SL, ST_GRAVITY, ST_SOLID, ST_PROJECT, ST_UPDATEVEL, ST_ADVECT};
            for (int s : st) runStage(s);
        } else {
            const int st[] = {ST_LEVELSET, ST_P2G, ST_GRAVITY, ST_SOLID, ST_PROJECT, ST_UPDATEVEL, ST_G2P, ST_ADVECT};
            for (int s : st) runStage(s);
        }
        currentTime += dt;
    }
};

}  // namespace

// ----------------------------------------------------------------------------
// C interface used by the Python tests (oracle/oracle.py).
// Field ids: 0 u, 1 v, 2 newu, 3 newv, 4 p, 5 phi, 6 density (double arrays);
//            7 cell (uint8); 8 particle positions, 9 particle velocities (x,y pairs);
//            10 Adiag, 11 Ax, 12 Ay, 13 rhs, 14 precon (after a projection).
// ----------------------------------------------------------------------------
extern "C" {

struct orc_stats {
    double water_volume, total_energy, particle_total_energy, current_time;
    double last_residual, cfl_max, orig_water_volume;
    int last_iters, hit_max, nan_flag, n_particles;
};

void* orc_create(int nx, int ny, int k, double dt, double dx, double rho, double gx, double gy,
                 int mode, double alpha, const uint8_t* init, unsigned seed, int level_set, int uniform_seeding) {
    Sim* s = new Sim();
    s->levelSet = level_set;
    s->create(nx, ny, k, dt, dx, rho, gx, gy, mode, alpha, init, seed, uniform_seeding);
    return s;
}

void orc_destroy(void* h) { delete (Sim*)h; }

void orc_set_options(void* h, double tau, double sigma, double tol, int max_iters, int sl_inplace,
                     int use_density) {
    Sim* s = (Sim*)h;
    s->tau = tau; s->sigma = sigma; s->tol = tol; s->maxIters = max_iters;
    s->slInplace = sl_inplace; s->useDensity = use_density;
}

// The reference's compile-time variants (src/FluidSim2D.cpp:15-16).
void orc_set_variants(void* h, int particle_cells, int ghost, int rk2) {
    Sim* s = (Sim*)h;
    s->particleCells = particle_cells;
    s->ghost = ghost;
    s->rk2 = rk2;
}

void orc_set_params(void* h, double dt, double gx, double gy, double alpha) {
    Sim* s = (Sim*)h;
    s->dt = dt; s->gravity = V2{gx, gy}; s->alpha = alpha;
}

void orc_step(void* h) { ((Sim*)h)->runFrame(); }
void orc_run_stage(void* h, int stage) { ((Sim*)h)->runStage(stage); }

static ArrD* fieldPtr(Sim* s, int f) {
    switch (f) {
        case 0: return &s->u; case 1: return &s->v; case 2: return &s->nu; case 3: return &s->nv;
        case 4: return &s->p; case 5: return &s->phi; case 6: return &s->dens;
        case 10: return &s->Adiag; case 11: return &s->Ax; case 12: return &s->Ay;
        case 13: return &s->rhs; case 14: return &s->precon;
        default: return nullptr;
    }
}

long orc_field_size(void* h, int f) {
    Sim* s = (Sim*)h;
    if (f == 7) return (long)s->cell.d.size();
    if (f == 8 || f == 9) return (long)s->pos.size() * 2;
    ArrD* a = fieldPtr(s, f);
    return a ? (long)a->d.size() : -1;
}

int orc_get_field(void* h, int f, void* out, long n) {
    Sim* s = (Sim*)h;
    if (n != orc_field_size(h, f)) return -1;
    if (f == 7) { std::memcpy(out, s->cell.d.data(), (size_t)n); return 0; }
    if (f == 8 || f == 9) {
        auto& src = f == 8 ? s->pos : s->vel;
        std::memcpy(out, src.data(), sizeof(double) * (size_t)n);
        return 0;
    }
    std::memcpy(out, fieldPtr(s, f)->d.data(), sizeof(double) * (size_t)n);
    return 0;
}

int orc_set_field(void* h, int f, const void* in, long n) {
    Sim* s = (Sim*)h;
    if (f == 8 || f == 9) {
        auto& dst = f == 8 ? s->pos : s->vel;
        dst.resize((size_t)n / 2);
        std::memcpy(dst.data(), in, sizeof(double) * (size_t)n);
        if (f == 8 && s->vel.size() != s->pos.size()) s->vel.resize(s->pos.size(), V2{0, 0});
        return 0;
    }
    if (n != orc_field_size(h, f)) return -1;
    if (f == 7) { std::memcpy(s->cell.d.data(), in, (size_t)n); return 0; }
    std::memcpy(fieldPtr(s, f)->d.data(), in, sizeof(double) * (size_t)n);
    return 0;
}

void orc_get_stats(void* h, orc_stats* st) {
    Sim* s = (Sim*)h;
    st->water_volume = s->waterVolume; st->total_energy = s->totalEnergy;
    st->particle_total_energy = s->particleTotalEnergy; st->current_time = s->currentTime;
    st->last_residual = s->lastResidual; st->cfl_max = s->cflMax;
    st->orig_water_volume = s->origWaterVolume;
    st->last_iters = s->lastIters; st->hit_max = s->hitMax; st->nan_flag = s->nanFlag;
    st->n_particles = (int)s->pos.size();
}

// Apply the MIC(0) preconditioner of the last assembled/factored system to r.
void orc_apply_precon(void* h, const double* r, double* z) {
    Sim* s = (Sim*)h;
    ArrD rr, zz;
    rr.create(s->sizeX, s->sizeY); zz.create(s->sizeX, s->sizeY);
    std::memcpy(rr.d.data(), r, sizeof(double) * rr.d.size());
    s->applyPreconditioner(rr, zz);
    std::memcpy(z, zz.d.data(), sizeof(double) * zz.d.size());
}

// Level-set sub-steps: 1 = constructFromParticles only, 2 = + redistance.
void orc_level_set_parts(void* h, int which) {
    Sim* s = (Sim*)h;
    s->constructFromParticles();
    if (which >= 2) s->redistance();
}

// Assemble and factor only (the first part of applyProjection).
void orc_assemble_factor(void* h) { Sim* s = (Sim*)h; s->assemble(); s->factor(); }

double orc_tricubic(const double* data, int nx, int ny, double x, double y) {
    ArrD a;
    a.create(nx, ny);
    std::memcpy(a.d.data(), data, sizeof(double) * a.d.size());
    return triCubic(a, x, y);
}

void orc_extrapolate(double* data, uint32_t* mask, int nx, int ny) {
    ArrD a; Arr<uint32_t> m;
    a.create(nx, ny); m.create(nx, ny);
    std::memcpy(a.d.data(), data, sizeof(double) * a.d.size());
    std::memcpy(m.d.data(), mask, sizeof(uint32_t) * m.d.size());
    extrapolate(a, m);
    std::memcpy(data, a.d.data(), sizeof(double) * a.d.size());
    std::memcpy(mask, m.d.data(), sizeof(uint32_t) * m.d.size());
}

// Adjoint of one step at the current state (Sim::adjointStep): u/v/density
// adjoints after the step in, before the step out (row-major, field shapes).
// Returns -1 when the state is outside the adjoint's scope.
int orc_adjoint_step(void* h, const double* ub_out, const double* vb_out, const double* db_out,
                     double* ub_in, double* vb_in, double* db_in) {
    Sim* s = (Sim*)h;
    if (!s->adjointSupported()) return -1;
    ArrD uo, vo, dO, ui, vi, di;
    uo.create(s->sizeX + 1, s->sizeY); vo.create(s->sizeX, s->sizeY + 1); dO.create(s->sizeX, s->sizeY);
    std::memcpy(uo.d.data(), ub_out, sizeof(double) * uo.d.size());
    std::memcpy(vo.d.data(), vb_out, sizeof(double) * vo.d.size());
    std::memcpy(dO.d.data(), db_out, sizeof(double) * dO.d.size());
    s->adjointStep(uo, vo, dO, ui, vi, di);
    std::memcpy(ub_in, ui.d.data(), sizeof(double) * ui.d.size());
    std::memcpy(vb_in, vi.d.data(), sizeof(double) * vi.d.size());
    std::memcpy(db_in, di.d.data(), sizeof(double) * di.d.size());
    return 0;
}

void orc_rand_stream(unsigned seed, uint32_t* out, int n) {
    GlibcRand g;
    g.seed(seed);
    for (int i = 0; i < n; i++) out[i] = g.next();
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
