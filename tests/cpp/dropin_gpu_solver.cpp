// Drives the GPU solver drop-in (include/physgrad/mpm_solver.h) the way a
// caller of the reference's MPMSolver<float> would: config, addParticles,
// initialize, step (or run).
//   in.bin: int32 nx, n, steps, use_run, jplane; float32 dt, x[3n], v[3n],
//           mass[n]; int32 type[n] (MaterialType).
//   jplane >= 0 adds DIRICHLET_Y to the node plane y = jplane (setBoundaryCondition).
//   use_run: config.simulation_time = steps dt, output every 2 dt to argv[3],
//            then run(); otherwise initialize() and steps x step().
//   out.bin: float32 x[3n], v[3n], F[9n], stress[6n], grid mass[G]; uint32 flags[G].
#include <cstdio>
#include <string>
#include <vector>

#include "physgrad/mpm_solver.h"

using namespace physgrad::mpm;

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[5];
    float dt;
    if (std::fread(hdr, 4, 5, f) != 5 || std::fread(&dt, 4, 1, f) != 1) return 2;
    const int nx = hdr[0], n = hdr[1], steps = hdr[2], use_run = hdr[3], jplane = hdr[4];
    std::vector<float> x(3 * n), v(3 * n), m(n);
    std::vector<int> type(n);
    if (std::fread(x.data(), 4, 3 * n, f) != (size_t)(3 * n) || std::fread(v.data(), 4, 3 * n, f) != (size_t)(3 * n) ||
        std::fread(m.data(), 4, n, f) != (size_t)n || std::fread(type.data(), 4, n, f) != (size_t)n)
        return 2;
    std::fclose(f);
    if (use_run && argc != 4) return 2;

    MPMSolverConfig cfg;
    cfg.grid_dimensions = {nx, nx, nx};
    cfg.grid_spacing = {1.0f / nx, 1.0f / nx, 1.0f / nx};
    cfg.time_step = dt;
    cfg.max_particles = n;
    cfg.enable_profiling = !use_run;
    if (use_run) {
        cfg.simulation_time = steps * dt;
        cfg.output_frequency = 2 * dt;
        cfg.output_directory = argv[3];
    }
    MPMSolver<float> solver(cfg);
    bool threw = false;
    try {
        solver.step();
    } catch (const std::runtime_error&) {
        threw = true;   // step before initialize
    }
    // particles grouped by material type, as a caller adds them
    for (int t = 0; t < 6; ++t) {
        std::vector<T3<float>> P, V;
        std::vector<float> M;
        for (int i = 0; i < n; ++i)
            if (type[i] == t) {
                P.push_back({x[3 * i], x[3 * i + 1], x[3 * i + 2]});
                V.push_back({v[3 * i], v[3 * i + 1], v[3 * i + 2]});
                M.push_back(m[i]);
            }
        if (!P.empty()) solver.addParticles(P, V, M, static_cast<MaterialType>(t));
    }
    if (jplane >= 0)
        for (int k = 0; k < nx; ++k)
            for (int i = 0; i < nx; ++i) {
                const size_t id = solver.getGrid().getLinearIndex(i, jplane, k);
                solver.setBoundaryCondition(
                    id, solver.getGrid().getBoundaryCondition(id) | static_cast<uint32_t>(BoundaryType::DIRICHLET_Y));
            }
    if (use_run) {
        solver.run();
    } else {
        solver.initialize();
        for (int s = 0; s < steps; ++s) solver.step();
    }
    const auto& parts = solver.getParticles();
    const auto& grid = solver.getGrid();
    const auto& met = solver.getMetrics();
    std::printf("steps %d time %.9g threw %d metrics_steps %d avg_step_ms %.4g\n", solver.getCurrentStep(),
                solver.getCurrentTime(), threw ? 1 : 0, met.total_steps, met.average_step_time * 1e3);
    std::vector<float> ox(3 * n), ov(3 * n), oF(9 * n), os(6 * n);
    for (int i = 0; i < n; ++i) {
        parts.getPosition(i, ox[3 * i], ox[3 * i + 1], ox[3 * i + 2]);
        parts.getVelocity(i, ov[3 * i], ov[3 * i + 1], ov[3 * i + 2]);
        parts.getDeformationGradient(i, &oF[9 * i]);
        parts.getStress(i, &os[6 * i]);
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(ox.data(), 4, ox.size(), o);
    std::fwrite(ov.data(), 4, ov.size(), o);
    std::fwrite(oF.data(), 4, oF.size(), o);
    std::fwrite(os.data(), 4, os.size(), o);
    std::fwrite(grid.getMassData(), 4, grid.getTotalNodes(), o);
    std::fwrite(grid.getBoundaryConditionData(), 4, grid.getTotalNodes(), o);
    std::fclose(o);
    return 0;
}
