// Uses the C++ drop-in of the reference CPU solver exactly as reference code
// would (physgrad::mpm::MPMSolver<float,64>): reads a scene from a raw binary
// file, runs `steps` steps, writes particle positions/velocities and the grid.
//
// input : int32 nx ny nz order steps n; float32 cell[3] origin[3] dt gravity[3];
//         float32 x[3n] v[3n] m[n]
// output: float32 x[3n] v[3n] grid_mass[G] grid_mom[3G] grid_vel[3G]
#include <cstdio>
#include <vector>

#include "physgrad/material_point_method.h"

using physgrad::ConceptVector3D;
using Solver = physgrad::mpm::MPMSolver<float, 64>;

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[6];
    float geo[10];
    if (std::fread(hdr, 4, 6, f) != 6 || std::fread(geo, 4, 10, f) != 10) return 2;
    const int n = hdr[5];
    std::vector<float> x(3 * n), v(3 * n), m(n);
    if (std::fread(x.data(), 4, x.size(), f) != x.size() || std::fread(v.data(), 4, v.size(), f) != v.size() ||
        std::fread(m.data(), 4, m.size(), f) != m.size())
        return 2;
    std::fclose(f);

    Solver::SimulationParams prm;
    prm.timestep = geo[6];
    prm.gravity = ConceptVector3D<float>(geo[7], geo[8], geo[9]);
    prm.interpolation_order = hdr[3];
    Solver s(int3(hdr[0], hdr[1], hdr[2]), ConceptVector3D<float>(geo[0], geo[1], geo[2]),
             ConceptVector3D<float>(geo[3], geo[4], geo[5]), prm, n);
    std::vector<ConceptVector3D<float>> P(n), V(n);
    for (int i = 0; i < n; ++i) {
        P[i] = ConceptVector3D<float>(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        V[i] = ConceptVector3D<float>(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
    s.addParticles(P, V, m, std::vector<float>(n, 1.0f));
    for (int k = 0; k < hdr[4]; ++k) s.step();

    const auto& parts = s.getParticles();
    const auto& grid = s.getGrid();
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    for (int i = 0; i < n; ++i) {
        auto p = parts.getPosition(i);
        std::fwrite(&p[0], 4, 3, o);
    }
    for (int i = 0; i < n; ++i) {
        auto u = parts.getVelocity(i);
        std::fwrite(&u[0], 4, 3, o);
    }
    std::fwrite(grid.mass.data(), 4, grid.total_nodes, o);
    for (size_t i = 0; i < grid.total_nodes; ++i) std::fwrite(&grid.momentum[i][0], 4, 3, o);
    for (size_t i = 0; i < grid.total_nodes; ++i) std::fwrite(&grid.velocity[i][0], 4, 3, o);
    std::fclose(o);
    std::printf("steps %zu time %.6g\n", s.getStepCount(), (double)s.getCurrentTime());
    return 0;
}
