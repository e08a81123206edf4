// FluidSim2D.hpp -- the reference's C++ FluidSim2D API over the gfx950 library.
//
// Header-only.  Code written against the reference's include/FluidSim2D.h,
// Array2D.h and MACGrid2D.h compiles against this header instead and runs the
// step on the GPU through the C ABI of include/fs2d.h (link -lfs2d):
//
//   FluidSim2DConfig config = {...};                 // include/FluidSim2D.h:48-63
//   FluidSim2D sim = FluidSim2D::create(config);     // src/FluidSim2D.cpp:23
//   sim.update();                                    // demo/App.cpp:213
//   sim.mac.u(i, j) += du;                           // demo/FluidRenderer2D.cpp:305
//   sim.gravity = vec2d{gx, gy};                     // demo/FluidRenderer2D.cpp:322
//   sim.saveStats();  sim.free();                    // demo/App.cpp:172-173
//
// Public fields are host mirrors of device state:
//   * an element access through operator() / operator[] downloads the field
//     the first time after a stage ran (lazy download);
//   * a non-const access marks the field dirty; dirty fields are uploaded
//     before the next stage runs;
//   * raw `.data` pointers bypass both: call pull() before reading through
//     them and markDirty() after writing through them;
//   * the scalars gravity, picFlipAlpha and dt are pushed before each stage
//     when they changed; currentTime and the energy/volume statistics are
//     refreshed after each stage.
// Copies of a FluidSim2D alias the same simulation, as in the reference
// (value type with manual create/free).  Errors throw std::runtime_error
// (the reference calls exit()); warnings (PCG cap, CFL) go to stderr, as the
// reference logs them.
//
// Define FS2D_HAVE_VEC2D before including this header when another header
// (the reference's altmath) already defines `vec2d` with members x, y.
#ifndef FS2D_FLUIDSIM2D_HPP
#define FS2D_FLUIDSIM2D_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fs2d.h"

#ifndef FS2D_HAVE_VEC2D
struct vec2d {
    double x, y;
};
inline vec2d operator+(vec2d a, vec2d b) { return {a.x + b.x, a.y + b.y}; }
inline vec2d operator-(vec2d a, vec2d b) { return {a.x - b.x, a.y - b.y}; }
inline vec2d operator*(vec2d a, double s) { return {a.x * s, a.y * s}; }
inline vec2d operator/(vec2d a, double s) { return {a.x / s, a.y / s}; }
#endif

// include/FluidSim2D.h:40-46
enum FluidSimMode { FS_SEMILAGRANGIAN, FS_PICFLIP };
enum FluidCellType : uint8_t { FS_EMPTY, FS_FLUID, FS_SOLID };

// include/FluidSim2D.h:48-63
struct FluidSim2DConfig {
    int sizeX;
    int sizeY;
    int particlesPerCellSqrt;
    double dt;
    double dx;
    double rho;
    double gravityX;
    double gravityY;
    FluidSimMode mode;
    double picFlipAlpha;
    FluidCellType* initialValues;
};

namespace fs2d_detail {

inline void check(fs2d_sim* h, fs2d_status st, const char* what) {
    if (st == FS2D_OK) return;
    std::string msg = std::string(what) + ": " + fs2d_status_string(st);
    if (h && fs2d_last_error(h) && *fs2d_last_error(h)) msg += std::string(" (") + fs2d_last_error(h) + ")";
    if (st > 0) {  // warnings: logged, as the reference logs them (FluidSim2D.cpp:464-466, 582-585)
        std::fprintf(stderr, "[fs2d] warning: %s\n", msg.c_str());
        return;
    }
    throw std::runtime_error(msg);
}

// One host mirror of one device field.
struct Mirror {
    fs2d_sim* h = nullptr;
    int field = 0, dtype = FS2D_F64;
    size_t count = 0;  // elements of dtype
    void* host = nullptr;
    bool valid = false, dirty = false;
    void read() {
        if (!valid) {
            check(h, fs2d_get_field(h, field, dtype, host, count), "fs2d_get_field");
            valid = true;
        }
    }
    void write() {
        read();
        dirty = true;
    }
    void push() {
        if (dirty) {
            check(h, fs2d_set_field(h, field, dtype, host, count), "fs2d_set_field");
            dirty = false;
        }
    }
};

}  // namespace fs2d_detail

// include/Array2D.h:21-101: row-major, element (i, j) at data[j*NX + i].
template <typename T>
struct Array2D {
    T* data = nullptr;
    int NX = 0;
    int NY = 0;
    fs2d_detail::Mirror* mirror = nullptr;

    T& operator()(int i, int j) {
        if (mirror) mirror->write();
        return data[(size_t)j * NX + i];
    }
    const T& operator()(int i, int j) const {
        if (mirror) mirror->read();
        return data[(size_t)j * NX + i];
    }
    bool isInBounds(int i, int j) const { return i >= 0 && i < NX && j >= 0 && j < NY; }

    // Catmull-Rom (Keys, a = -1/2) bicubic sample at grid coordinates
    // (Array2D.h:244-359): taps clamped to the array; 0 outside it.
    double triCubic(vec2d p) const {
        if (mirror) mirror->read();
        const int i = (int)p.x, j = (int)p.y;
        if (i < 0 || i >= NX || j < 0 || j >= NY) return 0.0;
        const double fx = p.x - i, fy = p.y - j;
        double wx[4], wy[4];
        keys(fx, wx);
        keys(fy, wy);
        double rows[4];
        for (int r = 0; r < 4; r++) {
            const int jj = std::min(std::max(j - 1 + r, 0), NY - 1);
            double acc = 0.0;
            for (int c = 0; c < 4; c++) {
                const int ii = std::min(std::max(i - 1 + c, 0), NX - 1);
                acc += wx[c] * (double)data[(size_t)jj * NX + ii];
            }
            rows[r] = acc;
        }
        return (wy[0] * rows[0] + wy[2] * rows[2]) + (wy[1] * rows[1] + wy[3] * rows[3]);
    }

   private:
    static void keys(double t, double* w) {
        const double t2 = t * t, t3 = t2 * t;
        w[0] = -0.5 * t3 + t2 - 0.5 * t;
        w[1] = 1.5 * t3 - 2.5 * t2 + 1.0;
        w[2] = -1.5 * t3 + 2.0 * t2 + 0.5 * t;
        w[3] = 0.5 * t3 - 0.5 * t2;
    }
};

// The reference's Vec<T> (altlib) as used for particles: data + size.
template <typename T>
struct Vec {
    T* data = nullptr;
    size_t size = 0;
    fs2d_detail::Mirror* mirror = nullptr;
    T& operator[](size_t k) {
        if (mirror) mirror->write();
        return data[k];
    }
    const T& operator[](size_t k) const {
        if (mirror) mirror->read();
        return data[k];
    }
};

// include/MACGrid2D.h:12-102: u is (NX+1) x NY, v is NX x (NY+1).
struct MACGrid2D {
    Array2D<double> u;
    Array2D<double> v;
    size_t NX = 0;
    size_t NY = 0;
    double dx = 1;

    vec2d velU(size_t i, size_t j) const {
        const Array2D<double>& V = v;
        double vy;
        if (i > 0 && i < NX)
            vy = 0.25 * (V(i - 1, j) + V(i - 1, j + 1) + V(i, j) + V(i, j + 1));
        else if (i == 0)
            vy = 0.25 * (V(i, j) + V(i, j + 1));
        else
            vy = 0.25 * (V(i - 1, j) + V(i - 1, j + 1));
        return vec2d{u(i, j), vy};
    }
    vec2d velV(size_t i, size_t j) const {
        const Array2D<double>& U = u;
        double ux;
        if (j > 0 && j < NY)
            ux = 0.25 * (U(i, j - 1) + U(i + 1, j - 1) + U(i, j) + U(i + 1, j));
        else if (j == 0)
            ux = 0.25 * (U(i, j) + U(i + 1, j));
        else
            ux = 0.25 * (U(i, j - 1) + U(i + 1, j - 1));
        return vec2d{ux, v(i, j)};
    }
    vec2d vel(size_t i, size_t j) const {
        const Array2D<double>&U = u, &V = v;
        return vec2d{0.5 * (U(i + 1, j) + U(i, j)), 0.5 * (V(i, j + 1) + V(i, j))};
    }
    double velInterpU(vec2d pos) const {
        double x = pos.x / dx, y = pos.y / dx - 0.5;
        x = std::min(std::max(x, 1e-6), (double)NX - 1 - 1e-6);
        y = std::min(std::max(y, 1e-6), (double)NY - 1 - 1e-6);
        return u.triCubic(vec2d{x, y});
    }
    double velInterpV(vec2d pos) const {
        double x = pos.x / dx - 0.5, y = pos.y / dx;
        x = std::min(std::max(x, 1e-6), (double)NX - 1 - 1e-6);
        y = std::min(std::max(y, 1e-6), (double)NY - 1 - 1e-6);
        return v.triCubic(vec2d{x, y});
    }
    vec2d velInterp(vec2d pos) const { return vec2d{velInterpU(pos), velInterpV(pos)}; }
    double velDiv(size_t i, size_t j) const {
        const Array2D<double>&U = u, &V = v;
        return (U(i + 1, j) - U(i, j) + V(i, j + 1) - V(i, j)) / dx;
    }
};

// include/FluidSim2D.h:15-38 (the level set's data; its methods run on the GPU
// inside createWaterLevelSet).
struct LevelSet {
    Array2D<double> phi;
    int sizeX = 0;
    int sizeY = 0;
    double dx = 0;
};

// include/PerformanceCounter.h / src/PerformanceCounter.cpp: per-stage host
// stopwatch, 30-frame rolling means, one row per frame in perf.csv.
struct PerformanceCounter {
    using Clock = std::chrono::high_resolution_clock;
    static constexpr int SampleCount = 30;
    Clock::time_point currentTime{};
    int currentStage = 0;
    int currentFrame = 0;
    std::vector<std::vector<float>> samples;
    std::vector<float> average;
    float avgTimePerFrame = 0.0f;
    bool sampleFinished = false;
    std::vector<float> averageStore;

    static PerformanceCounter create(int numStages) {
        PerformanceCounter c;
        c.samples.assign(numStages, std::vector<float>(SampleCount, 0.0f));
        c.average.assign(numStages, 0.0f);
        return c;
    }
    void beginStage() { currentTime = Clock::now(); }
    void endStage() {
        const auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - currentTime).count();
        if (currentStage < (int)samples.size()) samples[currentStage][currentFrame] = (float)ns * 1e-6f;
        currentStage++;
    }
    void endFrame() {
        if (sampleFinished)
            for (size_t s = 0; s < samples.size(); s++) {
                float a = 0.0f;
                for (int k = 0; k < SampleCount; k++) a += samples[s][k];
                average[s] = a / SampleCount;
            }
        currentStage = 0;
        if (currentFrame == SampleCount - 1) sampleFinished = true;
        currentFrame = (currentFrame + 1) % SampleCount;
        avgTimePerFrame = 0.0f;
        for (float a : average) {
            averageStore.push_back(a);
            avgTimePerFrame += a;
        }
    }
    void saveToFile(const char* filename) const {
        FILE* f = std::fopen(filename, "w+");
        if (!f) throw std::runtime_error(std::string("cannot open ") + filename);
        const size_t ns = average.size();
        for (size_t i = SampleCount; ns && i < averageStore.size() / ns; i++) {
            for (size_t j = 0; j < ns; j++) std::fprintf(f, "%s,", std::to_string(averageStore[ns * i + j]).c_str());
            std::fputc('\n', f);
        }
        std::fclose(f);
    }
};

// include/FluidSim2D.h:65-176
struct FluidSim2D {
    int sizeX = 0;
    int sizeY = 0;
    int particlesPerCellSqrt = 0;
    int particlesPerCell = 0;

    MACGrid2D mac;
    MACGrid2D newMac;
    Array2D<double> p;
    Array2D<FluidCellType> cell;
    Vec<vec2d> particles;
    Vec<vec2d> particleVels;

    LevelSet waterLevelSet;

    vec2d gravity = {0, -9.81};
    vec2d origGravity = {0, -9.81};
    double rho = 997.0;
    double picFlipAlpha = 1.0;
    bool rendered = false;
    double currentTime = 0.0;
    double dt = 0.0001;
    double dx = 0.001;
    double dr = 0.0009;

    enum class StageType {
        Init, CreateWaterLevelSet, TransferVelocityToGrid,
        ApplySemiLagrangianAdvection, ApplyGravity, CreateSolidLevelSet, ApplyProjection,
        UpdateVelocity, UpdateParticleVelocities, ApplyAdvection
    };
    int numStages = 0;
    StageType stage = StageType::Init;
    FluidSimMode mode = FS_PICFLIP;

    PerformanceCounter perfCounter;
    // Per-stage host stopwatch as in the reference (each stage then waits for
    // the GPU).  Off: runFrame() queues the whole step with one call.
    bool timeStages = true;

    double origWaterVolume = 0.0;
    double waterVolume = 0.0;
    std::vector<double> waterVolumeData;
    double totalEnergy = 0.0;
    std::vector<double> totalEnergyData;
    double particleTotalEnergy = 0.0;
    std::vector<double> particleTotalEnergyData;

    // src/FluidSim2D.cpp:23-76.  Solver constants default to the reference's
    // (tau 0.999, sigma 0.25, tol 1e-12, 200 iterations): setSolver() changes them.
    static FluidSim2D create(const FluidSim2DConfig& c) {
        fs2d_config cfg;
        fs2d_detail::check(nullptr, fs2d_config_init(&cfg), "fs2d_config_init");
        cfg.size_x = c.sizeX;
        cfg.size_y = c.sizeY;
        cfg.particles_per_cell_sqrt = c.particlesPerCellSqrt;
        cfg.dt = c.dt;
        cfg.dx = c.dx;
        cfg.rho = c.rho;
        cfg.gravity_x = c.gravityX;
        cfg.gravity_y = c.gravityY;
        cfg.mode = c.mode == FS_SEMILAGRANGIAN ? FS2D_SEMILAGRANGIAN : FS2D_PICFLIP;
        cfg.pic_flip_alpha = c.picFlipAlpha;
        cfg.initial_values = reinterpret_cast<const uint8_t*>(c.initialValues);
        fs2d_sim* h = nullptr;
        fs2d_detail::check(nullptr, fs2d_create(&cfg, &h), "fs2d_create");

        FluidSim2D s;
        s.st_ = std::make_shared<State>();
        s.st_->h = h;
        s.sizeX = c.sizeX;
        s.sizeY = c.sizeY;
        s.particlesPerCellSqrt = c.particlesPerCellSqrt;
        s.particlesPerCell = c.particlesPerCellSqrt * c.particlesPerCellSqrt;
        s.dt = c.dt;
        s.dx = c.dx;
        s.dr = 0.9 * c.dx;  // FluidSim2D.cpp:32
        s.rho = c.rho;
        s.gravity = s.origGravity = vec2d{c.gravityX, c.gravityY};
        s.picFlipAlpha = c.picFlipAlpha;
        s.mode = c.mode;
        s.numStages = c.mode == FS_SEMILAGRANGIAN ? 7 : 8;
        s.perfCounter = PerformanceCounter::create(s.numStages);
        s.mac.NX = s.newMac.NX = (size_t)c.sizeX;
        s.mac.NY = s.newMac.NY = (size_t)c.sizeY;
        s.mac.dx = s.newMac.dx = c.dx;
        s.waterLevelSet.sizeX = c.sizeX;
        s.waterLevelSet.sizeY = c.sizeY;
        s.waterLevelSet.dx = c.dx;
        s.bind(s.mac.u, FS2D_FIELD_U, c.sizeX + 1, c.sizeY);
        s.bind(s.mac.v, FS2D_FIELD_V, c.sizeX, c.sizeY + 1);
        s.bind(s.newMac.u, FS2D_FIELD_NEWU, c.sizeX + 1, c.sizeY);
        s.bind(s.newMac.v, FS2D_FIELD_NEWV, c.sizeX, c.sizeY + 1);
        s.bind(s.p, FS2D_FIELD_P, c.sizeX, c.sizeY);
        s.bind(s.waterLevelSet.phi, FS2D_FIELD_PHI, c.sizeX, c.sizeY);
        s.bind(s.cell, FS2D_FIELD_CELL, c.sizeX, c.sizeY);
        s.bindParticles();
        s.pushed_ = Pushed{s.gravity.x, s.gravity.y, s.picFlipAlpha, s.dt};
        s.refreshStats();
        s.origWaterVolume = s.waterVolume;
        return s;
    }

    // src/FluidSim2D.cpp:78-92
    void free() {
        if (st_ && st_->h) {
            fs2d_destroy(st_->h);
            st_->h = nullptr;
        }
        st_.reset();
    }

    // src/FluidSim2D.cpp:94-138
    void runFrame() {
        if (!timeStages) {
            beforeStage();
            fs2d_detail::check(h(), fs2d_step(h()), "fs2d_step");
            afterStage();
            recordStats();  // the step's level set computed them
            stage = StageType::ApplyAdvection;
        } else if (mode == FS_SEMILAGRANGIAN) {
            timed(StageType::CreateWaterLevelSet);
            timed(StageType::ApplySemiLagrangianAdvection);
            timed(StageType::ApplyGravity);
            timed(StageType::CreateSolidLevelSet);
            timed(StageType::ApplyProjection);
            timed(StageType::UpdateVelocity);
            timed(StageType::ApplyAdvection);
        } else {
            timed(StageType::CreateWaterLevelSet);
            timed(StageType::TransferVelocityToGrid);
            timed(StageType::ApplyGravity);
            timed(StageType::CreateSolidLevelSet);
            timed(StageType::ApplyProjection);
            timed(StageType::UpdateVelocity);
            timed(StageType::UpdateParticleVelocities);
            timed(StageType::ApplyAdvection);
        }
        rendered = false;
        currentTime += dt;
        perfCounter.endFrame();
    }
    // src/FluidSim2D.cpp:140-142
    void update() { runFrame(); }

    // The stage methods (include/FluidSim2D.h:124-140).
    void createWaterLevelSet() { runStage(StageType::CreateWaterLevelSet); }
    void applySemiLagrangianAdvection() { runStage(StageType::ApplySemiLagrangianAdvection); }
    void transferVelocityToGrid() { runStage(StageType::TransferVelocityToGrid); }
    void applyGravity() { runStage(StageType::ApplyGravity); }
    void createSolidLevelSet() { runStage(StageType::CreateSolidLevelSet); }
    void applyProjection() { runStage(StageType::ApplyProjection); }
    void updateVelocity() { runStage(StageType::UpdateVelocity); }
    void updateParticleVelocities() { runStage(StageType::UpdateParticleVelocities); }
    void applyAdvection() { runStage(StageType::ApplyAdvection); }

    // src/FluidSim2D.cpp:607-638
    double avgPressure() { return avgP(0); }
    double avgPressureInFluid() { return avgP(1); }
    double maxVelocity() {
        beforeStage();
        double m = 0;
        fs2d_detail::check(h(), fs2d_max_velocity(h(), &m), "fs2d_max_velocity");
        return m;
    }
    vec2d getGridCenter() const { return vec2d{sizeX * dx / 2, sizeY * dx / 2}; }
    // src/FluidSim2D.cpp:645-651
    vec2d clampPos(vec2d to) const {
        const double off = 1e-3;
        return vec2d{std::min(std::max(to.x, (1.0 + off) * dx), (sizeX - 1.0 - off) * dx),
                     std::min(std::max(to.y, (1.0 + off) * dx), (sizeY - 1.0 - off) * dx)};
    }
    // src/FluidSim2D.cpp:738-750: perf.csv and conservation.csv in the working directory.
    void saveStats() const { saveStats("perf.csv", "conservation.csv"); }
    void saveStats(const char* perfFile, const char* conservationFile) const {
        perfCounter.saveToFile(perfFile);
        FILE* f = std::fopen(conservationFile, "w+");
        if (!f) throw std::runtime_error(std::string("cannot open ") + conservationFile);
        std::fputs("Total Volume, Total Energy, Total Energy (Particle) \n", f);
        for (size_t k = 0; k < waterVolumeData.size(); k++)
            std::fprintf(f, "%f, %f, %f\n", waterVolumeData[k], totalEnergyData[k], particleTotalEnergyData[k]);
        std::fclose(f);
    }
    template <typename Fun>
    void iterate(Fun f) {
        for (int j = 0; j < sizeY; j++)
            for (int i = 0; i < sizeX; i++) f(i, j);
    }
    const char* printStage() const {
        switch (stage) {
            case StageType::Init: return "Init";
            case StageType::CreateWaterLevelSet: return "CreateWaterLevelSet";
            case StageType::TransferVelocityToGrid: return "TransferVelocityToGrid";
            case StageType::ApplySemiLagrangianAdvection: return "ApplySemiLagrangianAdvection";
            case StageType::ApplyGravity: return "ApplyGravity";
            case StageType::CreateSolidLevelSet: return "CreateSolidLevelSet";
            case StageType::ApplyProjection: return "ApplyProjection";
            case StageType::UpdateVelocity: return "UpdateVelocity";
            case StageType::UpdateParticleVelocities: return "UpdateParticleVelocities";
            case StageType::ApplyAdvection: return "ApplyAdvection";
        }
        return nullptr;
    }

    // --- no reference counterpart ------------------------------------------
    // Solver constants the reference hard-codes (FluidSim2D.cpp:369-370, 429, 453).
    void setSolver(double tau, double sigma, double tol, int maxIters) {
        fs2d_detail::check(h(), fs2d_set_solver(h(), tau, sigma, tol, maxIters), "fs2d_set_solver");
    }
    // Download every mirror now (before reading through raw .data pointers).
    void pull() {
        for (auto& m : st_->mirrors) m->read();
    }
    // Upload a mirror written through its raw .data pointer before the next stage.
    template <typename T>
    void markDirty(Array2D<T>& a) {
        if (a.mirror) a.mirror->write();
    }
    void markDirty(Vec<vec2d>& a) {
        if (a.mirror) a.mirror->write();
    }
    fs2d_sim* handle() const { return st_ ? st_->h : nullptr; }

   private:
    struct State {
        fs2d_sim* h = nullptr;
        std::vector<std::unique_ptr<fs2d_detail::Mirror>> mirrors;
        std::vector<std::vector<double>> f64;
        std::vector<std::vector<uint8_t>> u8;
        ~State() {
            if (h) fs2d_destroy(h);
        }
    };
    struct Pushed {
        double gx, gy, alpha, dt;
    };
    std::shared_ptr<State> st_;
    Pushed pushed_{};

    fs2d_sim* h() const {
        if (!st_ || !st_->h) throw std::runtime_error("FluidSim2D used after free()");
        return st_->h;
    }
    fs2d_detail::Mirror* newMirror(int field, int dtype, size_t count, void* host) {
        auto m = std::make_unique<fs2d_detail::Mirror>();
        m->h = st_->h;
        m->field = field;
        m->dtype = dtype;
        m->count = count;
        m->host = host;
        st_->mirrors.push_back(std::move(m));
        return st_->mirrors.back().get();
    }
    void bind(Array2D<double>& a, int field, int nx, int ny) {
        st_->f64.emplace_back((size_t)nx * ny, 0.0);
        a.data = st_->f64.back().data();
        a.NX = nx;
        a.NY = ny;
        a.mirror = newMirror(field, FS2D_F64, (size_t)nx * ny, a.data);
    }
    void bind(Array2D<FluidCellType>& a, int field, int nx, int ny) {
        st_->u8.emplace_back((size_t)nx * ny, 0);
        a.data = reinterpret_cast<FluidCellType*>(st_->u8.back().data());
        a.NX = nx;
        a.NY = ny;
        a.mirror = newMirror(field, FS2D_U8, (size_t)nx * ny, a.data);
    }
    void bindParticles() {
        const size_t n = fs2d_field_count(st_->h, FS2D_FIELD_PARTICLE_POS) / 2;
        static_assert(sizeof(vec2d) == 2 * sizeof(double), "vec2d must be two doubles");
        for (Vec<vec2d>* v : {&particles, &particleVels}) {
            st_->f64.emplace_back(2 * n, 0.0);
            v->data = reinterpret_cast<vec2d*>(st_->f64.back().data());
            v->size = n;
            v->mirror = newMirror(v == &particles ? FS2D_FIELD_PARTICLE_POS : FS2D_FIELD_PARTICLE_VEL, FS2D_F64,
                                  2 * n, v->data);
        }
    }
    void beforeStage() {
        fs2d_sim* H = h();
        if (gravity.x != pushed_.gx || gravity.y != pushed_.gy)
            fs2d_detail::check(H, fs2d_set_gravity(H, gravity.x, gravity.y), "fs2d_set_gravity");
        if (picFlipAlpha != pushed_.alpha)
            fs2d_detail::check(H, fs2d_set_pic_flip_alpha(H, picFlipAlpha), "fs2d_set_pic_flip_alpha");
        if (dt != pushed_.dt) fs2d_detail::check(H, fs2d_set_dt(H, dt), "fs2d_set_dt");
        pushed_ = Pushed{gravity.x, gravity.y, picFlipAlpha, dt};
        for (auto& m : st_->mirrors) m->push();
    }
    void afterStage() {
        for (auto& m : st_->mirrors) m->valid = false;
        refreshStats();
    }
    void refreshStats() {
        fs2d_stats s;
        fs2d_detail::check(h(), fs2d_get_stats(h(), &s), "fs2d_get_stats");
        waterVolume = s.water_volume;
        totalEnergy = s.total_energy;
        particleTotalEnergy = s.particle_total_energy;
    }
    // The reference appends the statistics inside createWaterLevelSet
    // (FluidSim2D.cpp:709-731), once per frame.
    void recordStats() {
        waterVolumeData.push_back(waterVolume);
        totalEnergyData.push_back(totalEnergy);
        particleTotalEnergyData.push_back(particleTotalEnergy);
    }
    static int stageId(StageType t) { return (int)t; }
    void runStage(StageType t) {
        beforeStage();
        stage = t;
        fs2d_detail::check(h(), fs2d_run_stage(h(), stageId(t)), "fs2d_run_stage");
        afterStage();
        if (t == StageType::CreateWaterLevelSet) recordStats();
    }
    void timed(StageType t) {
        beforeStage();
        stage = t;
        perfCounter.beginStage();
        fs2d_detail::check(h(), fs2d_run_stage(h(), stageId(t)), "fs2d_run_stage");
        fs2d_detail::check(h(), fs2d_synchronize(h()), "fs2d_synchronize");
        perfCounter.endStage();
        afterStage();
        if (t == StageType::CreateWaterLevelSet) recordStats();
    }
    double avgP(int fluidOnly) {
        beforeStage();
        double a = 0;
        fs2d_detail::check(h(), fs2d_avg_pressure(h(), fluidOnly, &a), "fs2d_avg_pressure");
        return a;
    }
};

#endif  // FS2D_FLUIDSIM2D_HPP
