// Driver that instantiates the reference's own CPU MPM solver, spatial hash and
// SPH-MPM coupling system from their headers where they lie
// (src/material_point_method.h, src/sparse_data_structures.h,
// src/sph_mpm_coupling.h) and exposes them through extern "C" for the
// golden-vector script and the CPU baseline.  Built by oracle/Makefile into
// oracle/_ref/ only.  TEST INFRASTRUCTURE ONLY.
#include <cstring>
#include <vector>
#include <array>
#include <chrono>
#include "material_point_method.h"
#include "sparse_data_structures.h"
#include "sph_mpm_coupling.h"

using physgrad::ConceptVector3D;

namespace {
void copy_out(const physgrad::mpm::MPMSolver<float, 64>& s, int n, float* x, float* v) {
    const auto& P = s.getParticles();
    for (int i = 0; i < n; ++i) {
        auto p = P.getPosition(i);
        auto u = P.getVelocity(i);
        for (int a = 0; a < 3; ++a) { x[3 * i + a] = p[a]; v[3 * i + a] = u[a]; }
    }
}
}  // namespace

extern "C" {

// Reference MPMSolver<float,64>: ctor (material_point_method.h:450-452),
// addParticles (:454-477), step (:479-500), getParticles/getGrid (:503-504).
int ref_mpm_run(int nx, int ny, int nz, float dx, float dy, float dz, float ox, float oy, float oz,
                float dt, float gx, float gy, float gz, int order, int n, float* x, float* v,
                const float* m, int nsteps, float* grid_mass, float* grid_mom, float* grid_vel,
                double* seconds) {
    using Solver = physgrad::mpm::MPMSolver<float, 64>;
    Solver::SimulationParams prm;
    prm.timestep = dt;
    prm.gravity = ConceptVector3D<float>(gx, gy, gz);
    prm.interpolation_order = order;
    int3 dims(nx, ny, nz);
    Solver s(dims, ConceptVector3D<float>(dx, dy, dz), ConceptVector3D<float>(ox, oy, oz), prm);
    std::vector<ConceptVector3D<float>> P(n), V(n);
    std::vector<float> M(m, m + n), Vol(n, 1.0f);
    for (int i = 0; i < n; ++i) {
        P[i] = ConceptVector3D<float>(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        V[i] = ConceptVector3D<float>(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
    s.addParticles(P, V, M, Vol);
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nsteps; ++k) s.step();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    copy_out(s, n, x, v);
    const auto& G = s.getGrid();
    size_t tot = (size_t)nx * ny * nz;
    for (size_t i = 0; i < tot; ++i) {
        if (grid_mass) grid_mass[i] = G.mass[i];
        for (int a = 0; a < 3; ++a) {
            if (grid_mom) grid_mom[3 * i + a] = G.momentum[i][a];
            if (grid_vel) grid_vel[3 * i + a] = G.velocity[i][a];
        }
    }
    return 0;
}

// Reference SpatialHashGrid<float,uint32_t>::buildHashTable
// (sparse_data_structures.h:56-82 ctor, :155-190 build, :284-287 accessors).
// cell ids are recovered from the sorted index list and the per-cell counts.
int ref_spatial_hash(const float* pos, unsigned n, float cell_size, const float* dmin, const float* dmax,
                     unsigned* cell_ids, unsigned* particle_indices, unsigned* cell_start,
                     unsigned* cell_count, unsigned* total_cells_out) {
    std::array<float, 3> lo{dmin[0], dmin[1], dmin[2]}, hi{dmax[0], dmax[1], dmax[2]};
    physgrad::sparse::SpatialHashGrid<float, uint32_t> h(cell_size, lo, hi, n);
    std::vector<float> p(pos, pos + 3 * (size_t)n);
    h.buildHashTable(p, n);
    unsigned tot = h.getTotalCells();
    if (total_cells_out) *total_cells_out = tot;
    const auto& idx = h.getParticleIndices();
    const auto& st = h.getCellStartIndices();
    const auto& ct = h.getCellParticleCounts();
    for (unsigned i = 0; i < n; ++i) particle_indices[i] = idx[i];
    for (unsigned c = 0; c < tot; ++c) {
        cell_start[c] = st[c];
        cell_count[c] = ct[c];
        for (unsigned k = 0; k < ct[c]; ++k) cell_ids[idx[st[c] + k]] = c;
    }
    return 0;
}

// Reference SpatialHashGrid::getNeighbors (sparse_data_structures.h:193-228)
// for nq queries after buildHashTable; counts [nq], out (capacity cap) the
// lists back to back.  Returns the total, or -1 when it exceeds cap.
long long ref_hash_neighbors(const float* pos, unsigned n, float cell_size, const float* dmin, const float* dmax,
                             const float* queries, unsigned nq, float radius, unsigned* counts, unsigned* out,
                             unsigned long long cap) {
    std::array<float, 3> lo{dmin[0], dmin[1], dmin[2]}, hi{dmax[0], dmax[1], dmax[2]};
    physgrad::sparse::SpatialHashGrid<float, uint32_t> h(cell_size, lo, hi, n);
    std::vector<float> p(pos, pos + 3 * (size_t)n);
    h.buildHashTable(p, n);
    unsigned long long k = 0;
    for (unsigned q = 0; q < nq; ++q) {
        auto nb = h.getNeighbors(queries[3 * q], queries[3 * q + 1], queries[3 * q + 2], radius);
        counts[q] = (unsigned)nb.size();
        for (auto i : nb) {
            if (k >= cap) return -1;
            out[k++] = i;
        }
    }
    return (long long)k;
}

// Reference SPHMPMCouplingSystem<float> (sph_mpm_coupling.h:413-579):
// findCouplingPairs then computeCouplingForces on the given SPH and MPM
// particles.  Outputs per pair (capacity cap): sph id, mpm id, penetration,
// solid and fluid force [3]; the SPH particles' accumulated force [3 ns].
// Returns the pair count, or -1 when it exceeds cap.
long long ref_coupling(const float* sph_x, const float* sph_v, unsigned ns, const float* mpm_x, const float* mpm_v,
                       unsigned nm, float radius, float stiffness, float damping, float friction, int two_way,
                       unsigned* pair_sph, unsigned* pair_mpm, float* penetration, float* solid_force,
                       float* fluid_force, float* sph_force, unsigned long long cap) {
    using namespace physgrad::coupling::sph_mpm;
    CouplingConfig<float> cfg;
    cfg.coupling_radius = radius;
    cfg.coupling_stiffness = stiffness;
    cfg.coupling_damping = damping;
    cfg.friction_coefficient = friction;
    cfg.enable_two_way_coupling = two_way != 0;
    SPHMPMCouplingSystem<float> sys(cfg);
    sys.initializeSPHSystem(ns);
    sys.initializeMPMSystem(int3(4, 4, 4), ConceptVector3D<float>(1.f, 1.f, 1.f), ConceptVector3D<float>(0.f, 0.f, 0.f),
                            nm);
    auto& sph = sys.getSPHSystem().getParticles();
    for (unsigned s = 0; s < ns; ++s) {
        sph[s].position = ConceptVector3D<float>(sph_x[3 * s], sph_x[3 * s + 1], sph_x[3 * s + 2]);
        sph[s].velocity = ConceptVector3D<float>(sph_v[3 * s], sph_v[3 * s + 1], sph_v[3 * s + 2]);
    }
    auto& mp = sys.getMPMParticles();
    for (unsigned m = 0; m < nm; ++m) {
        mp.setPosition(m, ConceptVector3D<float>(mpm_x[3 * m], mpm_x[3 * m + 1], mpm_x[3 * m + 2]));
        mp.setVelocity(m, ConceptVector3D<float>(mpm_v[3 * m], mpm_v[3 * m + 1], mpm_v[3 * m + 2]));
    }
    sys.findCouplingPairs();
    sys.computeCouplingForces();
    const auto& F = sys.getCouplingForces();
    if (F.size() > cap) return -1;
    for (size_t k = 0; k < F.size(); ++k) {
        pair_sph[k] = (unsigned)F[k].fluid_particle_id;
        pair_mpm[k] = (unsigned)F[k].solid_particle_id;
        penetration[k] = F[k].penetration_depth;
        for (int a = 0; a < 3; ++a) {
            solid_force[3 * k + a] = F[k].solid_force[a];
            fluid_force[3 * k + a] = F[k].fluid_force[a];
        }
    }
    for (unsigned s = 0; s < ns; ++s)
        for (int a = 0; a < 3; ++a) sph_force[3 * s + a] = sph[s].force[a];
    return (long long)F.size();
}

// Reference SparseMatrix<float,uint32_t> (sparse_data_structures.h:304-443):
// addElement in the given order, finalize, multiply; row pointers and y out.
int ref_sparse_matrix(unsigned nrows, unsigned ncols, unsigned nnz, const unsigned* rows, const unsigned* cols,
                      const float* vals, const float* x, float* y, unsigned* row_ptr) {
    physgrad::sparse::SparseMatrix<float, uint32_t> A(nrows, ncols, nnz);
    for (unsigned k = 0; k < nnz; ++k) A.addElement(rows[k], cols[k], vals[k]);
    A.finalize();
    std::vector<float> xv(x, x + ncols);
    std::vector<float> yv = A.multiply(xv);
    for (unsigned r = 0; r < nrows; ++r) y[r] = yv[r];
    const auto& rp = A.getRowPointers();
    for (unsigned r = 0; r <= nrows; ++r) row_ptr[r] = rp[r];
    return 0;
}

}  // extern "C"
