// Drop-in for the reference's GPU MPM solver header (src/mpm_solver.h) and the
// data-structure header it includes (src/mpm_data_structures.h).
//
// Keeps the names, namespaces and members a caller touches: int3, vec3<T>,
// float3, MaterialType, MaterialParameters, createElasticMaterial /
// createFluidMaterial, ParticleAoSoA<T,256> (xyz-interleaved within a chunk,
// mpm_data_structures.h:106-358), MPMGrid<T> (flat arrays, node index
// k*nx*ny + j*nx + i, :365-527), MaterialDatabase (:539-612), BoundaryType /
// BoundaryCondition (:617-634), MPMSolverConfig (mpm_solver.h:29-66),
// MPMPerformanceMetrics (:72-101) and MPMSolver<T> with initialize,
// addParticles, step, run, getParticles, getGrid, getMetrics, getCurrentTime,
// getCurrentStep (:107-328).
//
// The step runs on the GPU through the C-ABI (include/physgrad_mpm.h) as one
// MLS-MPM context (tiled implementation, hipGraph replay of each step unless
// enable_profiling asks for per-kernel times):
// APIC transfers, F <- (I + dt C) F, per-particle materials as the reference's
// fusedG2P2GKernel selects them (mpm_kernels.cu:407-421): FLUID particles the
// Newtonian fluid (:131-149), every other type the neo-Hookean model
// (:97-128), each with its database entry's Young's modulus, Poisson ratio,
// viscosity and density (volume = mass / density, mpm_solver.h:490-493), the
// stress entering P2G as the Kirchhoff stress tau = J sigma; grid
// v = p/m + dt g and the node's BoundaryType flags (mpm_kernels.cu:536-539,
// plus STICKY and SLIP, which the reference declares and never applies).
// Differences from the reference, each because the reference's version
// cannot run as written:
//   * The reference builds ParticleAoSoA with max_particles entries, so its
//     addParticles always throws; here the container starts empty and
//     max_particles is the capacity.
//   * Material lookup: the reference indexes the database by the enum value
//     (ELASTIC -> entry 0 = water, nu = 0.5, a singular elastic model); here a
//     particle of type t uses the first database entry whose type is t (or,
//     failing that, entry t).
//   * Boundaries: the reference marks the y = 0 node plane DIRICHLET_Y but
//     never copies the flags to the device (d_grid_boundary_conditions_ is
//     left unset); here the flags of grid_ are uploaded and applied.  The
//     plane is the whole y = 0 plane (the reference's loop runs its z index
//     up to dims.y).  Every node within config.boundary_nodes of a wall is
//     also sticky (flag STICKY), so particles cannot leave the grid.
//   * getGrid(): momentum is m v of the post-update velocity and forces are
//     zero (the fused kernels never store them); getParticles() refreshes x, v,
//     F and the Cauchy stress (neo-Hookean from F, or the fluid's
//     2 viscosity sym(C)).
//   * outputData() (print-only in the reference, :501-504) writes a legacy
//     binary VTK file of the particles to output_directory, or calls the
//     callback given to setOutputCallback.
//   * Informational printing happens only with enable_debug_output.
//
// Link with libphysgrad_mpm.so.  T must be float.  Do not include together
// with physgrad/material_point_method.h: the reference's two headers define
// different ParticleAoSoA / MPMGrid in the same namespace, and so do these.
#ifndef PHYSGRAD_DROPIN_MPM_SOLVER_H
#define PHYSGRAD_DROPIN_MPM_SOLVER_H

#ifdef PHYSGRAD_DROPIN_MATERIAL_POINT_METHOD_H
#error "physgrad/mpm_solver.h and physgrad/material_point_method.h define the same types; include one of them"
#endif

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../physgrad_mpm.h"

struct int3 {
    int x, y, z;
    int3() = default;
    int3(int x_, int y_, int z_) : x(x_), y(y_), z(z_) {}
};

template <typename T>
struct vec3 {
    T x, y, z;
    vec3() = default;
    vec3(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
};

using float3 = vec3<float>;
using double3 = vec3<double>;
template <typename T>
using T3 = vec3<T>;

namespace physgrad {
namespace mpm {

enum class MaterialType : uint32_t { ELASTIC = 0, ELASTOPLASTIC = 1, FLUID = 2, SAND = 3, SNOW = 4, JELLY = 5 };

struct MaterialParameters {
    float density = 1000.0f;
    float youngs_modulus = 1e6f;
    float poisson_ratio = 0.3f;
    float yield_stress = 1e5f;
    float hardening_coefficient = 0.0f;
    float viscosity = 0.001f;
    float friction_angle = 30.0f;
    float cohesion = 1000.0f;
    float critical_compression = 2.5e-2f;
    float critical_stretch = 7.5e-3f;
    MaterialType type = MaterialType::ELASTIC;
};

template <typename T>
MaterialParameters createElasticMaterial(T density, T youngs_modulus, T poisson_ratio) {
    MaterialParameters m;
    m.density = density;
    m.youngs_modulus = youngs_modulus;
    m.poisson_ratio = poisson_ratio;
    m.type = MaterialType::ELASTIC;
    return m;
}

template <typename T>
MaterialParameters createFluidMaterial(T density, T viscosity, T gamma = 7.0f) {
    MaterialParameters m;
    m.density = density;
    m.viscosity = viscosity;
    m.youngs_modulus = gamma;
    m.type = MaterialType::FLUID;
    return m;
}

// Chunked particle storage; within a chunk, vectors are xyz-interleaved
// (x of particle l at [3l], F row-major at [9l .. 9l+8], stress xx yy zz xy xz yz
// at [6l .. 6l+5]).
template <typename T, size_t ChunkSize = 256>
class ParticleAoSoA {
public:
    static constexpr size_t chunk_size = ChunkSize;
    using scalar_type = T;

    ParticleAoSoA() = default;
    explicit ParticleAoSoA(size_t n) { resize(n); }

    void resize(size_t n) {
        num_particles_ = n;
        num_chunks_ = (n + chunk_size - 1) / chunk_size;
        position_chunks_.resize(num_chunks_);
        velocity_chunks_.resize(num_chunks_);
        mass_chunks_.resize(num_chunks_);
        volume_chunks_.resize(num_chunks_);
        deformation_gradient_chunks_.resize(num_chunks_);
        stress_chunks_.resize(num_chunks_);
        plastic_deformation_chunks_.resize(num_chunks_);
        material_type_chunks_.resize(num_chunks_);
        material_id_chunks_.resize(num_chunks_);
        active_chunks_.resize(num_chunks_);
    }
    size_t size() const { return num_particles_; }
    size_t num_chunks() const { return num_chunks_; }

    void setPosition(size_t p, T x, T y, T z) { set3(position_chunks_, p, x, y, z); }
    void getPosition(size_t p, T& x, T& y, T& z) const { get3(position_chunks_, p, x, y, z); }
    void setVelocity(size_t p, T x, T y, T z) { set3(velocity_chunks_, p, x, y, z); }
    void getVelocity(size_t p, T& x, T& y, T& z) const { get3(velocity_chunks_, p, x, y, z); }
    void setMass(size_t p, T m) { mass_chunks_[p / chunk_size][p % chunk_size] = m; }
    T getMass(size_t p) const { return mass_chunks_[p / chunk_size][p % chunk_size]; }
    void setVolume(size_t p, T v) { volume_chunks_[p / chunk_size][p % chunk_size] = v; }
    T getVolume(size_t p) const { return volume_chunks_[p / chunk_size][p % chunk_size]; }
    void setDeformationGradient(size_t p, const T F[9]) {
        T* c = &deformation_gradient_chunks_[p / chunk_size][9 * (p % chunk_size)];
        for (int i = 0; i < 9; ++i) c[i] = F[i];
    }
    void getDeformationGradient(size_t p, T F[9]) const {
        const T* c = &deformation_gradient_chunks_[p / chunk_size][9 * (p % chunk_size)];
        for (int i = 0; i < 9; ++i) F[i] = c[i];
    }
    void setStress(size_t p, const T s[6]) {
        T* c = &stress_chunks_[p / chunk_size][6 * (p % chunk_size)];
        for (int i = 0; i < 6; ++i) c[i] = s[i];
    }
    void getStress(size_t p, T s[6]) const {
        const T* c = &stress_chunks_[p / chunk_size][6 * (p % chunk_size)];
        for (int i = 0; i < 6; ++i) s[i] = c[i];
    }
    void setMaterialType(size_t p, MaterialType t) { material_type_chunks_[p / chunk_size][p % chunk_size] = t; }
    MaterialType getMaterialType(size_t p) const { return material_type_chunks_[p / chunk_size][p % chunk_size]; }
    void setActive(size_t p, bool a) { active_chunks_[p / chunk_size][p % chunk_size] = a ? 1u : 0u; }
    bool isActive(size_t p) const { return active_chunks_[p / chunk_size][p % chunk_size] != 0u; }

    void setPositions(const std::vector<std::array<T, 3>>& v) {
        for (size_t i = 0; i < std::min(v.size(), size()); ++i) setPosition(i, v[i][0], v[i][1], v[i][2]);
    }
    void setVelocities(const std::vector<std::array<T, 3>>& v) {
        for (size_t i = 0; i < std::min(v.size(), size()); ++i) setVelocity(i, v[i][0], v[i][1], v[i][2]);
    }
    void setMasses(const std::vector<T>& m) {
        for (size_t i = 0; i < std::min(m.size(), size()); ++i) setMass(i, m[i]);
    }
    std::vector<std::array<T, 3>> getPositions() const {
        std::vector<std::array<T, 3>> r(size());
        for (size_t i = 0; i < size(); ++i) getPosition(i, r[i][0], r[i][1], r[i][2]);
        return r;
    }
    std::vector<std::array<T, 3>> getVelocities() const {
        std::vector<std::array<T, 3>> r(size());
        for (size_t i = 0; i < size(); ++i) getVelocity(i, r[i][0], r[i][1], r[i][2]);
        return r;
    }
    std::vector<T> getMasses() const {
        std::vector<T> r(size());
        for (size_t i = 0; i < size(); ++i) r[i] = getMass(i);
        return r;
    }

    T* getPositionData() { return reinterpret_cast<T*>(position_chunks_.data()); }
    T* getVelocityData() { return reinterpret_cast<T*>(velocity_chunks_.data()); }
    T* getMassData() { return reinterpret_cast<T*>(mass_chunks_.data()); }
    T* getVolumeData() { return reinterpret_cast<T*>(volume_chunks_.data()); }
    T* getDeformationGradientData() { return reinterpret_cast<T*>(deformation_gradient_chunks_.data()); }
    T* getStressData() { return reinterpret_cast<T*>(stress_chunks_.data()); }
    MaterialType* getMaterialTypeData() { return reinterpret_cast<MaterialType*>(material_type_chunks_.data()); }
    uint32_t* getActiveData() { return reinterpret_cast<uint32_t*>(active_chunks_.data()); }
    const T* getPositionData() const { return reinterpret_cast<const T*>(position_chunks_.data()); }
    const T* getVelocityData() const { return reinterpret_cast<const T*>(velocity_chunks_.data()); }
    const T* getMassData() const { return reinterpret_cast<const T*>(mass_chunks_.data()); }
    const T* getVolumeData() const { return reinterpret_cast<const T*>(volume_chunks_.data()); }
    const T* getDeformationGradientData() const {
        return reinterpret_cast<const T*>(deformation_gradient_chunks_.data());
    }
    const T* getStressData() const { return reinterpret_cast<const T*>(stress_chunks_.data()); }
    const MaterialType* getMaterialTypeData() const {
        return reinterpret_cast<const MaterialType*>(material_type_chunks_.data());
    }
    const uint32_t* getActiveData() const { return reinterpret_cast<const uint32_t*>(active_chunks_.data()); }

    size_t getMemoryFootprint() const {
        return num_chunks_ * (sizeof(position_chunks_[0]) + sizeof(velocity_chunks_[0]) + sizeof(mass_chunks_[0]) +
                              sizeof(volume_chunks_[0]) + sizeof(deformation_gradient_chunks_[0]) +
                              sizeof(stress_chunks_[0]) + sizeof(plastic_deformation_chunks_[0]) +
                              sizeof(material_type_chunks_[0]) + sizeof(material_id_chunks_[0]) +
                              sizeof(active_chunks_[0]));
    }

private:
    using V3 = std::vector<std::array<T, chunk_size * 3>>;
    static void set3(V3& v, size_t p, T x, T y, T z) {
        T* c = &v[p / chunk_size][3 * (p % chunk_size)];
        c[0] = x; c[1] = y; c[2] = z;
    }
    static void get3(const V3& v, size_t p, T& x, T& y, T& z) {
        const T* c = &v[p / chunk_size][3 * (p % chunk_size)];
        x = c[0]; y = c[1]; z = c[2];
    }
    V3 position_chunks_, velocity_chunks_;
    std::vector<std::array<T, chunk_size>> mass_chunks_, volume_chunks_;
    std::vector<std::array<T, chunk_size * 9>> deformation_gradient_chunks_;
    std::vector<std::array<T, chunk_size * 6>> stress_chunks_;
    std::vector<std::array<T, chunk_size * 9>> plastic_deformation_chunks_;
    std::vector<std::array<MaterialType, chunk_size>> material_type_chunks_;
    std::vector<std::array<uint32_t, chunk_size>> material_id_chunks_;
    std::vector<std::array<uint32_t, chunk_size>> active_chunks_;
    size_t num_particles_ = 0;
    size_t num_chunks_ = 0;
};

template <typename T>
class MPMGrid {
public:
    using scalar_type = T;

    MPMGrid() = default;
    MPMGrid(int3 dimensions, T3<T> spacing, T3<T> origin = {T{0}, T{0}, T{0}})
        : dimensions_(dimensions), grid_spacing_(spacing), origin_(origin) {
        total_nodes_ = static_cast<size_t>(dimensions_.x) * dimensions_.y * dimensions_.z;
        masses_.assign(total_nodes_, T{0});
        velocities_.assign(3 * total_nodes_, T{0});
        forces_.assign(3 * total_nodes_, T{0});
        momentum_.assign(3 * total_nodes_, T{0});
        boundary_conditions_.assign(total_nodes_, 0u);
    }
    MPMGrid(const std::array<size_t, 3>& res, const std::array<T, 3>& domain)
        : MPMGrid(int3{static_cast<int>(res[0]), static_cast<int>(res[1]), static_cast<int>(res[2])},
                  T3<T>{domain[0] / res[0], domain[1] / res[1], domain[2] / res[2]}, T3<T>{T{0}, T{0}, T{0}}) {}

    void clearGrid() {
        std::fill(masses_.begin(), masses_.end(), T{0});
        std::fill(velocities_.begin(), velocities_.end(), T{0});
        std::fill(forces_.begin(), forces_.end(), T{0});
        std::fill(momentum_.begin(), momentum_.end(), T{0});
    }
    std::array<T, 3> getGridSpacing() const { return {{grid_spacing_.x, grid_spacing_.y, grid_spacing_.z}}; }
    int3 getDimensions() const { return dimensions_; }
    T3<T> getSpacing() const { return grid_spacing_; }
    T3<T> getOrigin() const { return origin_; }
    size_t getTotalNodes() const { return total_nodes_; }

    size_t getLinearIndex(int i, int j, int k) const {
        return static_cast<size_t>(k) * dimensions_.x * dimensions_.y + static_cast<size_t>(j) * dimensions_.x +
               static_cast<size_t>(i);
    }
    int3 getGridIndex(size_t l) const {
        const size_t plane = static_cast<size_t>(dimensions_.x) * dimensions_.y;
        const int k = static_cast<int>(l / plane);
        const int r = static_cast<int>(l % plane);
        return {r % dimensions_.x, r / dimensions_.x, k};
    }
    T3<T> getWorldPosition(int i, int j, int k) const {
        return {origin_.x + i * grid_spacing_.x, origin_.y + j * grid_spacing_.y, origin_.z + k * grid_spacing_.z};
    }
    int3 getGridPosition(T x, T y, T z) const {
        return {static_cast<int>(std::floor((x - origin_.x) / grid_spacing_.x)),
                static_cast<int>(std::floor((y - origin_.y) / grid_spacing_.y)),
                static_cast<int>(std::floor((z - origin_.z) / grid_spacing_.z))};
    }

    void setMass(size_t n, T m) { masses_[n] = m; }
    T getMass(size_t n) const { return masses_[n]; }
    void setVelocity(size_t n, T x, T y, T z) { velocities_[3 * n] = x; velocities_[3 * n + 1] = y; velocities_[3 * n + 2] = z; }
    void getVelocity(size_t n, T& x, T& y, T& z) const {
        x = velocities_[3 * n]; y = velocities_[3 * n + 1]; z = velocities_[3 * n + 2];
    }
    void addVelocity(size_t n, T x, T y, T z) {
        velocities_[3 * n] += x; velocities_[3 * n + 1] += y; velocities_[3 * n + 2] += z;
    }
    void setForce(size_t n, T x, T y, T z) { forces_[3 * n] = x; forces_[3 * n + 1] = y; forces_[3 * n + 2] = z; }
    void addForce(size_t n, T x, T y, T z) { forces_[3 * n] += x; forces_[3 * n + 1] += y; forces_[3 * n + 2] += z; }
    void setBoundaryCondition(size_t n, uint32_t f) { boundary_conditions_[n] = f; }
    uint32_t getBoundaryCondition(size_t n) const { return boundary_conditions_[n]; }

    T* getMassData() { return masses_.data(); }
    T* getVelocityData() { return velocities_.data(); }
    T* getForceData() { return forces_.data(); }
    T* getMomentumData() { return momentum_.data(); }
    uint32_t* getBoundaryConditionData() { return boundary_conditions_.data(); }
    const T* getMassData() const { return masses_.data(); }
    const T* getVelocityData() const { return velocities_.data(); }
    const T* getForceData() const { return forces_.data(); }
    const T* getMomentumData() const { return momentum_.data(); }
    const uint32_t* getBoundaryConditionData() const { return boundary_conditions_.data(); }

    size_t getMemoryFootprint() const {
        return (masses_.size() + velocities_.size() + forces_.size() + momentum_.size()) * sizeof(T) +
               boundary_conditions_.size() * sizeof(uint32_t);
    }

private:
    int3 dimensions_{0, 0, 0};
    T3<T> grid_spacing_{T{0}, T{0}, T{0}};
    T3<T> origin_{T{0}, T{0}, T{0}};
    std::vector<T> masses_, velocities_, forces_, momentum_;
    std::vector<uint32_t> boundary_conditions_;
    size_t total_nodes_ = 0;
};

class MaterialDatabase {
public:
    MaterialDatabase() {
        addMaterial(createWaterMaterial());
        addMaterial(createJelloMaterial());
        addMaterial(createSandMaterial());
        addMaterial(createSnowMaterial());
    }
    uint32_t addMaterial(const MaterialParameters& m) {
        materials_.push_back(m);
        return static_cast<uint32_t>(materials_.size() - 1);
    }
    const MaterialParameters& getMaterial(uint32_t id) const { return materials_.at(id); }
    MaterialParameters& getMaterial(uint32_t id) { return materials_.at(id); }
    size_t getNumMaterials() const { return materials_.size(); }

    static MaterialParameters createWaterMaterial() {
        MaterialParameters m;
        m.type = MaterialType::FLUID;
        m.density = 1000.0f;
        m.viscosity = 0.001f;
        m.youngs_modulus = 2.2e9f;
        m.poisson_ratio = 0.5f;
        return m;
    }
    static MaterialParameters createJelloMaterial() {
        MaterialParameters m;
        m.type = MaterialType::ELASTIC;
        m.density = 1100.0f;
        m.youngs_modulus = 1e4f;
        m.poisson_ratio = 0.3f;
        return m;
    }
    static MaterialParameters createSandMaterial() {
        MaterialParameters m;
        m.type = MaterialType::SAND;
        m.density = 1600.0f;
        m.youngs_modulus = 3.5e7f;
        m.poisson_ratio = 0.3f;
        m.friction_angle = 30.0f;
        m.cohesion = 0.0f;
        return m;
    }
    static MaterialParameters createSnowMaterial() {
        MaterialParameters m;
        m.type = MaterialType::SNOW;
        m.density = 400.0f;
        m.youngs_modulus = 1.4e5f;
        m.poisson_ratio = 0.2f;
        m.critical_compression = 2.5e-2f;
        m.critical_stretch = 7.5e-3f;
        m.hardening_coefficient = 10.0f;
        return m;
    }

    // The entry a particle of type t uses: the first entry of that type, else entry t.
    uint32_t idForType(MaterialType t) const {
        for (size_t i = 0; i < materials_.size(); ++i)
            if (materials_[i].type == t) return static_cast<uint32_t>(i);
        return static_cast<uint32_t>(t);
    }

private:
    std::vector<MaterialParameters> materials_;
};

enum class BoundaryType : uint32_t {
    FREE = 0,
    DIRICHLET_X = 1,
    DIRICHLET_Y = 2,
    DIRICHLET_Z = 4,
    NEUMANN_X = 8,
    NEUMANN_Y = 16,
    NEUMANN_Z = 32,
    STICKY = 64,
    SLIP = 128
};

template <typename T>
struct BoundaryCondition {
    BoundaryType type;
    T3<T> value;
    T3<T> normal;
};

struct MPMSolverConfig {
    int3 grid_dimensions = {64, 64, 64};
    float3 grid_spacing = {0.01f, 0.01f, 0.01f};
    float3 domain_origin = {0.0f, 0.0f, 0.0f};

    float time_step = 1e-4f;
    float simulation_time = 1.0f;
    float cfl_factor = 0.4f;

    int max_particles = 1000000;
    float particle_radius = 0.005f;
    bool enable_gravity = true;
    float3 gravity = {0.0f, -9.81f, 0.0f};

    bool enable_multi_material = true;
    int max_materials = 8;

    int cuda_block_size = 256;
    bool enable_g2p2g_fusion = true;
    bool enable_adaptive_timestep = false;
    float min_timestep = 1e-6f;
    float max_timestep = 1e-3f;

    bool enable_output = true;
    float output_frequency = 0.01f;
    std::string output_directory = "./mpm_output";

    bool enable_profiling = false;
    bool enable_debug_output = false;
    bool verify_conservation = false;

    // not in the reference: thickness of the sticky wall band in nodes
    int boundary_nodes = 3;
};

struct MPMPerformanceMetrics {
    double total_simulation_time = 0.0;
    double average_step_time = 0.0;
    double g2p_time = 0.0;
    double p2g_time = 0.0;
    double grid_update_time = 0.0;
    double particle_update_time = 0.0;

    size_t total_memory_usage = 0;
    size_t particle_memory = 0;
    size_t grid_memory = 0;

    int total_steps = 0;
    int rejected_steps = 0;
    float average_particles_per_second = 0.0f;

    void reset() { *this = MPMPerformanceMetrics{}; }
    void print() const {
        std::cout << "MPM Performance Metrics:\n"
                  << "  Total simulation time: " << total_simulation_time << " s\n"
                  << "  Average step time: " << average_step_time * 1000 << " ms\n"
                  << "  Particles per second: " << average_particles_per_second / 1e6 << " M\n"
                  << "  Total memory usage: " << total_memory_usage / (1024 * 1024) << " MB\n";
        if (total_simulation_time > 0)
            std::cout << "  Grid update time: " << grid_update_time / total_simulation_time * 100 << "%\n"
                      << "  G2P2G time: " << (g2p_time + p2g_time) / total_simulation_time * 100 << "%\n";
    }
};

template <typename T = float>
class MPMSolver {
    static_assert(std::is_same<T, float>::value, "the GPU MPM path computes in float");

public:
    using scalar_type = T;
    using particle_container = ParticleAoSoA<T>;
    using grid_container = MPMGrid<T>;

    explicit MPMSolver(const MPMSolverConfig& config = MPMSolverConfig{})
        : config_(config), grid_(config.grid_dimensions, config.grid_spacing, config.domain_origin) {
        setupBoundaryConditions();
    }
    ~MPMSolver() { cleanup(); }
    MPMSolver(const MPMSolver&) = delete;
    MPMSolver& operator=(const MPMSolver&) = delete;

    // Creates the device context (reference :167-187).
    void initialize() {
        if (ctx_) return;
        const float3 h = config_.grid_spacing;
        if (!(h.x == h.y && h.y == h.z)) throw std::runtime_error("MPMSolver: the MLS-MPM path needs cubic cells");
        pg_mpm_params p;
        pg_mpm_default_params(&p);
        p.mode = PG_MODE_MLS_MPM;
        p.impl = PG_IMPL_TILED;
        p.nx = config_.grid_dimensions.x;
        p.ny = config_.grid_dimensions.y;
        p.nz = config_.grid_dimensions.z;
        p.dx = p.dy = p.dz = h.x;
        p.ox = config_.domain_origin.x;
        p.oy = config_.domain_origin.y;
        p.oz = config_.domain_origin.z;
        p.dt = config_.time_step;
        const float3 g = config_.enable_gravity ? config_.gravity : float3{0.f, 0.f, 0.f};
        p.gx = g.x; p.gy = g.y; p.gz = g.z;
        p.material = PG_MAT_NEO_HOOKEAN;
        p.boundary_thickness = config_.boundary_nodes;
        p.max_particles = config_.max_particles;
        check(pg_mpm_create(&p, &ctx_));
        // per-kernel timing needs eager launches: a replayed graph records no events
        if (config_.enable_profiling) check(pg_mpm_set_profiling(ctx_, 1));
        else check(pg_mpm_set_graph(ctx_, 1));
        upload();
        const size_t n = particles_.size(), G = grid_.getTotalNodes();
        metrics_.particle_memory = n * (24 * sizeof(T) + 2 * sizeof(uint32_t));
        metrics_.grid_memory = G * 8 * sizeof(T);
        metrics_.total_memory_usage = metrics_.particle_memory + metrics_.grid_memory;
        if (config_.enable_debug_output)
            std::cout << "MPM Solver initialized: grid " << p.nx << "x" << p.ny << "x" << p.nz << ", "
                      << n << " particles, dx " << h.x << ", dt " << config_.time_step << "\n";
    }

    // Appends particles (reference :190-225): F = I, stress 0, volume = mass / density.
    void addParticles(const std::vector<T3<T>>& positions, const std::vector<T3<T>>& velocities,
                      const std::vector<T>& masses, MaterialType material_type = MaterialType::ELASTIC) {
        const size_t add = positions.size(), old = particles_.size();
        if (velocities.size() != add || masses.size() != add)
            throw std::runtime_error("addParticles: positions, velocities and masses differ in length");
        if (old + add > static_cast<size_t>(config_.max_particles))
            throw std::runtime_error("Exceeded maximum particle count");
        syncParticles();
        particles_.resize(old + add);
        const MaterialParameters& mat = material_db_.getMaterial(material_db_.idForType(material_type));
        const T I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Z[6] = {0, 0, 0, 0, 0, 0};
        for (size_t i = 0; i < add; ++i) {
            const size_t q = old + i;
            particles_.setPosition(q, positions[i].x, positions[i].y, positions[i].z);
            particles_.setVelocity(q, velocities[i].x, velocities[i].y, velocities[i].z);
            particles_.setMass(q, masses[i]);
            particles_.setVolume(q, masses[i] / mat.density);
            particles_.setMaterialType(q, material_type);
            particles_.setActive(q, true);
            particles_.setDeformationGradient(q, I);
            particles_.setStress(q, Z);
        }
        C_.resize(9 * (old + add), T{0});
        dirty_ = true;
        if (ctx_) {   // re-create: material constants and capacity may have changed
            pg_mpm_destroy(ctx_);
            ctx_ = nullptr;
            initialize();
        }
    }

    void step() {
        if (!ctx_) throw std::runtime_error("Solver not initialized");
        const auto t0 = std::chrono::high_resolution_clock::now();
        if (dirty_) upload();
        if (bc_dirty_) uploadBoundary();
        check(pg_mpm_step(ctx_, 1, nullptr));
        host_fresh_ = false;
        grid_fresh_ = false;
        current_time_ += config_.time_step;
        current_step_++;
        const auto t1 = std::chrono::high_resolution_clock::now();
        metrics_.total_steps++;
        host_time_ += std::chrono::duration<double>(t1 - t0).count();
    }

    // Steps until simulation_time (reference :288-320); the output hook runs
    // each time the clock crosses a multiple of output_frequency.
    void run() {
        initialize();
        const auto t0 = std::chrono::high_resolution_clock::now();
        while (current_time_ < config_.simulation_time) {
            step();
            if (config_.enable_debug_output && current_step_ % 100 == 0)
                std::cout << "Progress: " << 100.f * current_time_ / config_.simulation_time << "%\n";
            if (config_.enable_output && config_.output_frequency > 0.f &&
                static_cast<int>(current_time_ / config_.output_frequency) >
                    static_cast<int>((current_time_ - config_.time_step) / config_.output_frequency))
                outputData();
        }
        syncParticles();   // blocking read-back: the run is complete
        metrics_.total_simulation_time =
            std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        if (config_.enable_profiling) getMetrics().print();
    }

    const particle_container& getParticles() const {
        const_cast<MPMSolver*>(this)->syncParticles();
        return particles_;
    }
    const grid_container& getGrid() const {
        const_cast<MPMSolver*>(this)->syncGrid();
        return grid_;
    }
    // Kernel times come from HIP events on the launch stream when
    // enable_profiling is set; otherwise average_step_time is host time per
    // step() call (launch cost, the step is asynchronous).
    const MPMPerformanceMetrics& getMetrics() const {
        MPMPerformanceMetrics& m = const_cast<MPMSolver*>(this)->metrics_;
        if (m.total_steps > 0) {
            if (ctx_ && config_.enable_profiling) {
                double p2g = 0, fused = 0, g2p = 0, grid = 0, bin = 0;
                long long k = 0;
                pg_mpm_kernel_time(ctx_, 0, &p2g, &k);
                pg_mpm_kernel_time(ctx_, 1, &fused, &k);
                pg_mpm_kernel_time(ctx_, 2, &g2p, &k);
                pg_mpm_kernel_time(ctx_, 3, &grid, &k);
                pg_mpm_kernel_time(ctx_, 4, &bin, &k);
                m.p2g_time = (p2g + 0.5 * fused) / 1e3;
                m.g2p_time = (g2p + 0.5 * fused) / 1e3;
                m.grid_update_time = grid / 1e3;
                m.particle_update_time = bin / 1e3;
                m.average_step_time = (p2g + fused + g2p + grid + bin) / 1e3 / m.total_steps;
            } else {
                m.average_step_time = host_time_ / m.total_steps;
            }
            if (m.average_step_time > 0) m.average_particles_per_second = particles_.size() / m.average_step_time;
        }
        return m;
    }
    float getCurrentTime() const { return current_time_; }
    int getCurrentStep() const { return current_step_; }

    // not in the reference: replaces the VTK output of outputData
    void setOutputCallback(std::function<void(const MPMSolver&)> cb) { output_cb_ = std::move(cb); }
    // not in the reference: the material database (changes apply from the
    // next addParticles or initialize)
    MaterialDatabase& getMaterialDatabase() { return material_db_; }
    // not in the reference: BoundaryType bits of a grid node (index
    // getLinearIndex(i, j, k)); applied from the next step()
    void setBoundaryCondition(size_t node, uint32_t flags) {
        if (node >= grid_.getTotalNodes()) throw std::runtime_error("setBoundaryCondition: node out of range");
        grid_.setBoundaryCondition(node, flags);
        bc_dirty_ = true;
    }
    // Writes the particles as legacy binary VTK polydata: points, VECTORS
    // velocity, SCALARS mass and J = det F.
    void writeVTK(const std::string& path) const {
        const particle_container& P = getParticles();
        const size_t n = P.size();
        std::FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) throw std::runtime_error("writeVTK: cannot open " + path);
        std::string hdr = "# vtk DataFile Version 3.0\nphysgrad_amd MPMSolver t=" + std::to_string(current_time_) +
                          "\nBINARY\nDATASET POLYDATA\nPOINTS " + std::to_string(n) + " float\n";
        std::fwrite(hdr.data(), 1, hdr.size(), f);
        std::vector<uint32_t> buf;
        auto put = [&](const std::vector<uint32_t>& b) { std::fwrite(b.data(), 4, b.size(), f); };
        auto be = [](float v) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            return __builtin_bswap32(u);
        };
        buf.resize(3 * n);
        for (size_t i = 0; i < n; ++i) {
            T x, y, z;
            P.getPosition(i, x, y, z);
            buf[3 * i] = be(x); buf[3 * i + 1] = be(y); buf[3 * i + 2] = be(z);
        }
        put(buf);
        hdr = "\nVERTICES " + std::to_string(n) + " " + std::to_string(2 * n) + "\n";
        std::fwrite(hdr.data(), 1, hdr.size(), f);
        buf.resize(2 * n);
        for (size_t i = 0; i < n; ++i) {
            buf[2 * i] = __builtin_bswap32(1u);
            buf[2 * i + 1] = __builtin_bswap32(static_cast<uint32_t>(i));
        }
        put(buf);
        hdr = "\nPOINT_DATA " + std::to_string(n) + "\nVECTORS velocity float\n";
        std::fwrite(hdr.data(), 1, hdr.size(), f);
        buf.resize(3 * n);
        for (size_t i = 0; i < n; ++i) {
            T x, y, z;
            P.getVelocity(i, x, y, z);
            buf[3 * i] = be(x); buf[3 * i + 1] = be(y); buf[3 * i + 2] = be(z);
        }
        put(buf);
        hdr = "\nSCALARS mass float 1\nLOOKUP_TABLE default\n";
        std::fwrite(hdr.data(), 1, hdr.size(), f);
        buf.resize(n);
        for (size_t i = 0; i < n; ++i) buf[i] = be(P.getMass(i));
        put(buf);
        hdr = "\nSCALARS J float 1\nLOOKUP_TABLE default\n";
        std::fwrite(hdr.data(), 1, hdr.size(), f);
        for (size_t i = 0; i < n; ++i) {
            T F[9];
            P.getDeformationGradient(i, F);
            buf[i] = be(F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                        F[2] * (F[3] * F[7] - F[4] * F[6]));
        }
        put(buf);
        std::fputc('\n', f);
        if (std::fclose(f) != 0) throw std::runtime_error("writeVTK: write to " + path + " failed");
    }

private:
    static void fail(const pg_mpm_ctx* c) {
        throw std::runtime_error(std::string("physgrad_mpm: ") + pg_mpm_last_error(c));
    }
    void check(int rc) const {
        if (rc != PG_OK) fail(ctx_);
    }

    // Material table of the database entries the particles use, and each
    // particle's index into it.
    void resolveMaterials() {
        const size_t n = particles_.size();
        std::vector<int> slot(material_db_.getNumMaterials(), -1);
        table_.clear();
        ids_.assign(n, 0);
        for (size_t i = 0; i < n; ++i) {
            const uint32_t id = material_db_.idForType(particles_.getMaterialType(i));
            if (id >= slot.size())
                throw std::runtime_error("MPMSolver: no material database entry for particle type " +
                                         std::to_string(static_cast<uint32_t>(particles_.getMaterialType(i))));
            if (slot[id] < 0) {
                if (table_.size() == 256) throw std::runtime_error("MPMSolver: more than 256 materials in use");
                const MaterialParameters& m = material_db_.getMaterial(id);
                pg_material_desc d;
                const bool fluid = m.type == MaterialType::FLUID;
                d.model = fluid ? PG_MAT_FLUID : PG_MAT_NEO_HOOKEAN;
                d.youngs_modulus = fluid ? 0.f : m.youngs_modulus;
                d.poisson_ratio = fluid ? 0.f : m.poisson_ratio;
                d.density = m.density;
                d.viscosity = fluid ? m.viscosity : 0.f;
                slot[id] = static_cast<int>(table_.size());
                table_.push_back(d);
            }
            ids_[i] = static_cast<uint8_t>(slot[id]);
        }
    }

    // The reference's floor (DIRICHLET_Y on the y = 0 plane, mpm_solver.h:
    // 472-484) plus the sticky band of config.boundary_nodes nodes; set at
    // construction, so flags a caller sets afterwards are kept.
    void setupBoundaryConditions() {
        const int3 d = config_.grid_dimensions;
        const int b = config_.boundary_nodes;
        for (int k = 0; k < d.z; ++k)
            for (int j = 0; j < d.y; ++j)
                for (int i = 0; i < d.x; ++i) {
                    const bool s = i < b || j < b || k < b || i >= d.x - b || j >= d.y - b || k >= d.z - b;
                    uint32_t f = s ? static_cast<uint32_t>(BoundaryType::STICKY) : 0u;
                    if (j == 0) f |= static_cast<uint32_t>(BoundaryType::DIRICHLET_Y);
                    grid_.setBoundaryCondition(grid_.getLinearIndex(i, j, k), f);
                }
        bc_dirty_ = true;
    }

    void uploadBoundary() {
        check(pg_mpm_set_boundary(ctx_, grid_.getBoundaryConditionData(), PG_HOST, nullptr));
        bc_dirty_ = false;
    }

    void upload() {
        const size_t n = particles_.size();
        std::vector<T> x(3 * n), v(3 * n), F(9 * n), m(n);
        for (size_t i = 0; i < n; ++i) {
            particles_.getPosition(i, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
            particles_.getVelocity(i, v[3 * i], v[3 * i + 1], v[3 * i + 2]);
            particles_.getDeformationGradient(i, &F[9 * i]);
            m[i] = particles_.getMass(i);
        }
        C_.resize(9 * n, T{0});
        resolveMaterials();
        check(pg_mpm_set_particles(ctx_, static_cast<int>(n), x.data(), v.data(), C_.data(), F.data(), m.data(),
                                   nullptr, PG_HOST, nullptr));
        if (n) check(pg_mpm_set_materials(ctx_, table_.data(), static_cast<int>(table_.size()), ids_.data(), PG_HOST,
                                          nullptr));
        uploadBoundary();
        dirty_ = false;
        host_fresh_ = true;
    }

    void syncParticles() {
        if (!ctx_ || host_fresh_) return;
        const size_t n = particles_.size();
        std::vector<T> x(3 * n), v(3 * n), F(9 * n);
        C_.resize(9 * n);
        check(pg_mpm_get_particles(ctx_, x.data(), v.data(), C_.data(), F.data(), PG_HOST, nullptr));
        for (size_t i = 0; i < n; ++i) {
            particles_.setPosition(i, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
            particles_.setVelocity(i, v[3 * i], v[3 * i + 1], v[3 * i + 2]);
            particles_.setDeformationGradient(i, &F[9 * i]);
            const pg_material_desc& d = table_[ids_[i]];
            T s[6];
            if (d.model == PG_MAT_FLUID) fluidStress(&C_[9 * i], d.viscosity, s);
            else cauchy(&F[9 * i], d.youngs_modulus, d.poisson_ratio, s);
            particles_.setStress(i, s);
        }
        host_fresh_ = true;
    }

    // Neo-Hookean Cauchy stress sigma = mu/J (F F^T - I) + lambda ln(J)/J I
    // (mpm_kernels.cu:97-128), stored xx yy zz xy xz yz.
    static void cauchy(const T* F, float E, float nu, T* s) {
        const double mu = E / (2.0 * (1.0 + nu)), lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        const double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                         F[2] * (F[3] * F[7] - F[4] * F[6]);
        auto B = [&](int r, int c) {
            return (double)F[3 * r] * F[3 * c] + (double)F[3 * r + 1] * F[3 * c + 1] + (double)F[3 * r + 2] * F[3 * c + 2];
        };
        const double iso = J > 0 ? lam * std::log(J) / J : 0.0, a = J > 0 ? mu / J : 0.0;
        s[0] = static_cast<T>(a * (B(0, 0) - 1.0) + iso);
        s[1] = static_cast<T>(a * (B(1, 1) - 1.0) + iso);
        s[2] = static_cast<T>(a * (B(2, 2) - 1.0) + iso);
        s[3] = static_cast<T>(a * B(0, 1));
        s[4] = static_cast<T>(a * B(0, 2));
        s[5] = static_cast<T>(a * B(1, 2));
    }

    // Newtonian fluid stress 2 viscosity sym(grad v) with grad v the APIC C
    // (mpm_kernels.cu:131-149), stored xx yy zz xy xz yz.
    static void fluidStress(const T* C, float visc, T* s) {
        s[0] = 2.f * visc * C[0];
        s[1] = 2.f * visc * C[4];
        s[2] = 2.f * visc * C[8];
        s[3] = visc * (C[1] + C[3]);
        s[4] = visc * (C[2] + C[6]);
        s[5] = visc * (C[5] + C[7]);
    }

    void syncGrid() {
        if (!ctx_ || grid_fresh_) return;
        const size_t G = grid_.getTotalNodes();
        check(pg_mpm_get_grid(ctx_, grid_.getMassData(), nullptr, grid_.getVelocityData(), PG_HOST, nullptr));
        T* p = grid_.getMomentumData();
        const T* v = grid_.getVelocityData();
        const T* m = grid_.getMassData();
        for (size_t i = 0; i < G; ++i)
            for (int c = 0; c < 3; ++c) p[3 * i + c] = m[i] * v[3 * i + c];
        std::fill(grid_.getForceData(), grid_.getForceData() + 3 * G, T{0});
        grid_fresh_ = true;
    }

    void outputData() {
        if (config_.enable_debug_output) std::cout << "Writing output at t=" << current_time_ << "\n";
        if (output_cb_) {
            output_cb_(*this);
            return;
        }
        std::filesystem::create_directories(config_.output_directory);
        char name[32];
        std::snprintf(name, sizeof(name), "particles_%06d.vtk", current_step_);
        writeVTK((std::filesystem::path(config_.output_directory) / name).string());
    }

    void cleanup() {
        if (ctx_) pg_mpm_destroy(ctx_);
        ctx_ = nullptr;
    }

    MPMSolverConfig config_;
    MaterialDatabase material_db_;
    MPMPerformanceMetrics metrics_;
    particle_container particles_;
    grid_container grid_;
    pg_mpm_ctx* ctx_ = nullptr;
    std::vector<T> C_;
    std::vector<pg_material_desc> table_;
    std::vector<uint8_t> ids_;
    std::function<void(const MPMSolver&)> output_cb_;
    bool dirty_ = true, host_fresh_ = true, grid_fresh_ = true, bc_dirty_ = true;
    double host_time_ = 0.0;
    float current_time_ = 0.0f;
    int current_step_ = 0;
};

}  // namespace mpm
}  // namespace physgrad

#endif
