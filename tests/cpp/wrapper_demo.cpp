// Demo-style driver of fluid-sim_amd/cpp/FluidSim2D.hpp, written the way the
// reference's demo uses FluidSim2D (demo/App.cpp:133-173,
// demo/FluidRenderer2D.cpp:305-322, 382-388): create, update() per frame,
// host writes to mac.u/v, gravity and picFlipAlpha between frames, reads of
// p, cell, particles and mac.velInterp, saveStats, free.
//
// usage: wrapper_demo <mode 0=SL 1=PIC> <n> <frames> <time_stages 0/1> <outdir>
// Writes <outdir>/frame<k>.bin: u, v, p (f64), cell (u8), particle x,y (f64),
// and <outdir>/scalars.txt with per-frame diagnostics.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../fluid-sim_amd/cpp/FluidSim2D.hpp"

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s mode n frames time_stages outdir\n", argv[0]);
        return 2;
    }
    const int mode = std::atoi(argv[1]), n = std::atoi(argv[2]), frames = std::atoi(argv[3]);
    const bool timeStages = std::atoi(argv[4]) != 0;
    const std::string out = argv[5];

    // dam break: SOLID border, FLUID block in the lower-left 40% x 60%
    std::vector<FluidCellType> cells((size_t)n * n, FS_EMPTY);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            FluidCellType c = FS_EMPTY;
            if (i == 0 || j == 0 || i == n - 1 || j == n - 1)
                c = FS_SOLID;
            else if (i < (int)(0.4 * n) && j < (int)(0.6 * n))
                c = FS_FLUID;
            cells[(size_t)j * n + i] = c;
        }
    FluidSim2DConfig config;
    config.sizeX = n;
    config.sizeY = n;
    config.particlesPerCellSqrt = 2;
    config.dt = 0.005;
    config.dx = 1.0 / n;
    config.rho = 997.0;
    config.gravityX = 0.0;
    config.gravityY = -9.81;
    config.mode = mode == 0 ? FS_SEMILAGRANGIAN : FS_PICFLIP;
    config.picFlipAlpha = 0.05;
    config.initialValues = cells.data();

    FluidSim2D sim = FluidSim2D::create(config);
    sim.timeStages = timeStages;
    sim.setSolver(0.97, 0.25, 1e-6, 500);

    FILE* sc = std::fopen((out + "/scalars.txt").c_str(), "w");
    for (int f = 0; f < frames; f++) {
        if (f == 2) {
            // FluidRenderer2D.cpp:305-308: mouse drag adds velocity to faces
            for (int j = n / 4; j < n / 2; j++) {
                sim.mac.u(n / 5, j) += 0.5;
                sim.mac.v(n / 5, j) -= 0.25;
            }
        }
        if (f == 3) {
            sim.gravity = vec2d{1.0, -12.0};  // FluidRenderer2D.cpp:322, 383-384
            sim.picFlipAlpha = 0.2;           // FluidRenderer2D.cpp:388
        }
        sim.update();

        const std::string fn = out + "/frame" + std::to_string(f) + ".bin";
        FILE* fo = std::fopen(fn.c_str(), "wb");
        const Array2D<double>& U = sim.mac.u;
        (void)U(0, 0);  // const access: download without marking dirty
        const Array2D<double>& V = sim.mac.v;
        (void)V(0, 0);
        const Array2D<double>& P = sim.p;
        (void)P(0, 0);
        const Array2D<FluidCellType>& C = sim.cell;
        (void)C(0, 0);
        const Vec<vec2d>& X = sim.particles;
        if (X.size) (void)X[0];
        std::fwrite(U.data, sizeof(double), (size_t)U.NX * U.NY, fo);
        std::fwrite(V.data, sizeof(double), (size_t)V.NX * V.NY, fo);
        std::fwrite(P.data, sizeof(double), (size_t)P.NX * P.NY, fo);
        std::fwrite(C.data, 1, (size_t)C.NX * C.NY, fo);
        std::fwrite(X.data, sizeof(vec2d), X.size, fo);
        std::fclose(fo);

        const vec2d c = sim.getGridCenter();
        const vec2d vi = sim.mac.velInterp(c);
        std::fprintf(sc, "%d %s %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %zu\n", f, sim.printStage(),
                     sim.currentTime, sim.waterVolume, sim.totalEnergy, sim.particleTotalEnergy, sim.avgPressure(),
                     sim.avgPressureInFluid(), vi.x, vi.y, sim.particles.size);
    }
    std::fclose(sc);
    sim.saveStats((out + "/perf.csv").c_str(), (out + "/conservation.csv").c_str());
    sim.free();
    std::printf("wrapper_demo ok\n");
    return 0;
}
