// Drop-in for the reference CPU MPM solver header (src/material_point_method.h).
//
// Keeps the reference's names, namespaces and signatures for the types a
// caller touches: ConceptVector3D<T>, int3, ParticleAoSoA<T,ChunkSize>
// (chunked, component-major within a chunk, :45-177), MPMGrid<T> (node index
// k*nx*ny + j*nx + i, :186-263), MPMShapeFunctions<T> (:272-324) and
// MPMSolver<T,ChunkSize> with SimulationParams, addParticles, step,
// getParticles, getGrid, getCurrentTime, getStepCount (:430-506).  The step
// runs on the GPU through the C-ABI (include/physgrad_mpm.h, ref-compat mode:
// same PIC semantics and float operations as the reference); getParticles()
// and getGrid() refresh host mirrors in the reference layouts on demand.
//
// Link with libphysgrad_mpm.so.  T must be float (the device path is FP32).
#ifndef PHYSGRAD_DROPIN_MATERIAL_POINT_METHOD_H
#define PHYSGRAD_DROPIN_MATERIAL_POINT_METHOD_H

#include <array>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../physgrad_mpm.h"

#ifndef PHYSGRAD_COMMON_TYPES_H
#define PHYSGRAD_COMMON_TYPES_H
struct int3 {
    int x, y, z;
    int3() : x(0), y(0), z(0) {}
    int3(int x_, int y_, int z_) : x(x_), y(y_), z(z_) {}
};
namespace physgrad {
template <typename T>
class ConceptVector3D {
public:
    using value_type = T;
    T x, y, z;
    constexpr ConceptVector3D() : x(T{0}), y(T{0}), z(T{0}) {}
    constexpr ConceptVector3D(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
    constexpr T& operator[](size_t i) { return (&x)[i]; }
    constexpr const T& operator[](size_t i) const { return (&x)[i]; }
    static constexpr size_t size() { return 3; }
    ConceptVector3D operator+(const ConceptVector3D& o) const { return {x + o.x, y + o.y, z + o.z}; }
    ConceptVector3D operator-(const ConceptVector3D& o) const { return {x - o.x, y - o.y, z - o.z}; }
    ConceptVector3D operator*(T s) const { return {x * s, y * s, z * s}; }
};
}  // namespace physgrad
#endif

namespace physgrad {
namespace mpm {

template <typename T, size_t ChunkSize = 64>
struct ParticleAoSoA {
    static constexpr size_t chunk_size = ChunkSize;
    using value_type = T;
    std::vector<std::array<T, chunk_size * 3>> position_chunks;
    std::vector<std::array<T, chunk_size * 3>> velocity_chunks;
    std::vector<std::array<T, chunk_size>> mass_chunks;
    std::vector<std::array<T, chunk_size>> volume_chunks;
    std::vector<std::array<T, chunk_size * 9>> deformation_gradient_chunks;
    std::vector<std::array<T, chunk_size * 6>> stress_chunks;
    std::vector<std::array<T, chunk_size>> young_modulus_chunks;
    std::vector<std::array<T, chunk_size>> poisson_ratio_chunks;
    std::vector<std::array<T, chunk_size>> density_chunks;
    size_t num_particles = 0;
    size_t num_chunks = 0;

    size_t size() const { return num_particles; }
    void resize(size_t n) {
        num_particles = n;
        num_chunks = (n + chunk_size - 1) / chunk_size;
        position_chunks.resize(num_chunks);
        velocity_chunks.resize(num_chunks);
        mass_chunks.resize(num_chunks);
        volume_chunks.resize(num_chunks);
        deformation_gradient_chunks.resize(num_chunks);
        stress_chunks.resize(num_chunks);
        young_modulus_chunks.resize(num_chunks);
        poisson_ratio_chunks.resize(num_chunks);
        density_chunks.resize(num_chunks);
    }
    ConceptVector3D<T> getPosition(size_t p) const {
        const auto& c = position_chunks[p / chunk_size];
        const size_t l = p % chunk_size;
        return {c[l], c[l + chunk_size], c[l + 2 * chunk_size]};
    }
    void setPosition(size_t p, const ConceptVector3D<T>& v) {
        auto& c = position_chunks[p / chunk_size];
        const size_t l = p % chunk_size;
        c[l] = v[0]; c[l + chunk_size] = v[1]; c[l + 2 * chunk_size] = v[2];
    }
    ConceptVector3D<T> getVelocity(size_t p) const {
        const auto& c = velocity_chunks[p / chunk_size];
        const size_t l = p % chunk_size;
        return {c[l], c[l + chunk_size], c[l + 2 * chunk_size]};
    }
    void setVelocity(size_t p, const ConceptVector3D<T>& v) {
        auto& c = velocity_chunks[p / chunk_size];
        const size_t l = p % chunk_size;
        c[l] = v[0]; c[l + chunk_size] = v[1]; c[l + 2 * chunk_size] = v[2];
    }
    T getMass(size_t p) const { return mass_chunks[p / chunk_size][p % chunk_size]; }
    void setMass(size_t p, T m) { mass_chunks[p / chunk_size][p % chunk_size] = m; }
    void getDeformationGradient(size_t p, T* F) const {
        const auto& c = deformation_gradient_chunks[p / chunk_size];
        for (int i = 0; i < 9; ++i) F[i] = c[p % chunk_size + i * chunk_size];
    }
    void setDeformationGradient(size_t p, const T* F) {
        auto& c = deformation_gradient_chunks[p / chunk_size];
        for (int i = 0; i < 9; ++i) c[p % chunk_size + i * chunk_size] = F[i];
    }
};

template <typename T>
struct MPMGrid {
    std::vector<T> mass;
    std::vector<ConceptVector3D<T>> velocity;
    std::vector<ConceptVector3D<T>> momentum;
    std::vector<ConceptVector3D<T>> force;
    int3 dimensions;
    int3 dims;
    ConceptVector3D<T> cell_size;
    ConceptVector3D<T> origin;
    size_t total_nodes;

    MPMGrid(const int3& d, const ConceptVector3D<T>& cs, const ConceptVector3D<T>& o)
        : dimensions(d), dims(d), cell_size(cs), origin(o) {
        total_nodes = static_cast<size_t>(d.x) * d.y * d.z;
        mass.assign(total_nodes, T{0});
        velocity.assign(total_nodes, ConceptVector3D<T>{});
        momentum.assign(total_nodes, ConceptVector3D<T>{});
        force.assign(total_nodes, ConceptVector3D<T>{});
    }
    size_t getNodeIndex(int i, int j, int k) const {
        return static_cast<size_t>(k) * dimensions.x * dimensions.y + static_cast<size_t>(j) * dimensions.x + i;
    }
    ConceptVector3D<T> getNodePosition(int i, int j, int k) const {
        return {origin[0] + i * cell_size[0], origin[1] + j * cell_size[1], origin[2] + k * cell_size[2]};
    }
};

template <typename T>
class MPMShapeFunctions {
public:
    static T linear(T x) {
        T a = x < T{0} ? -x : x;
        return a >= T{1} ? T{0} : T{1} - a;
    }
    static T quadratic(T x) {
        T a = x < T{0} ? -x : x;
        if (a >= T{1.5}) return T{0};
        if (a < T{0.5}) return T{0.75} - x * x;
        return T{0.5} * (T{1.5} - a) * (T{1.5} - a);
    }
    static T cubic(T x) {
        T a = x < T{0} ? -x : x;
        if (a >= T{2}) return T{0};
        if (a < T{1}) return static_cast<T>(2.0 / 3.0) - a * a + T{0.5} * a * a * a;
        T t = T{2} - a;
        return static_cast<T>(1.0 / 6.0) * t * t * t;
    }
};

template <typename T, size_t ChunkSize = 64>
class MPMSolver {
    static_assert(std::is_same<T, float>::value, "the GPU MPM path computes in float");

public:
    using ParticleData = ParticleAoSoA<T, ChunkSize>;
    using Grid = MPMGrid<T>;
    using ShapeFunc = MPMShapeFunctions<T>;

    struct SimulationParams {
        T timestep = T{1e-4};
        ConceptVector3D<T> gravity = {T{0}, T{-9.81}, T{0}};
        T cfl_factor = T{0.4};
        bool use_affine_pic = true;
        bool use_flip = true;
        T flip_ratio = T{0.95};
        int interpolation_order = 2;
    };

    MPMSolver(const int3& grid_dims, const ConceptVector3D<T>& cell_size, const ConceptVector3D<T>& origin,
              const SimulationParams& params = {}, int max_particles = 1 << 22)
        : grid_(grid_dims, cell_size, origin), params_(params) {
        pg_mpm_params p;
        pg_mpm_default_params(&p);
        p.mode = PG_MODE_REF_COMPAT;
        p.nx = grid_dims.x; p.ny = grid_dims.y; p.nz = grid_dims.z;
        p.dx = cell_size[0]; p.dy = cell_size[1]; p.dz = cell_size[2];
        p.ox = origin[0]; p.oy = origin[1]; p.oz = origin[2];
        p.dt = params.timestep;
        p.gx = params.gravity[0]; p.gy = params.gravity[1]; p.gz = params.gravity[2];
        p.interpolation_order = params.interpolation_order;
        p.max_particles = max_particles;
        check(pg_mpm_create(&p, &ctx_), nullptr);
    }
    ~MPMSolver() { pg_mpm_destroy(ctx_); }
    MPMSolver(const MPMSolver&) = delete;
    MPMSolver& operator=(const MPMSolver&) = delete;

    // addParticles (reference :454-477): appends; F starts at I; material
    // arguments are accepted and, as in the reference step, unused.
    void addParticles(const std::vector<ConceptVector3D<T>>& positions,
                      const std::vector<ConceptVector3D<T>>& velocities, const std::vector<T>& masses,
                      const std::vector<T>& volumes, T young_modulus = T{1e6}, T poisson_ratio = T{0.3},
                      T density = T{1000}) {
        (void)volumes; (void)young_modulus; (void)poisson_ratio; (void)density;
        syncParticles();
        const size_t old = x_.size() / 3, add = positions.size();
        for (size_t i = 0; i < add; ++i)
            for (int a = 0; a < 3; ++a) {
                x_.push_back(positions[i][a]);
                v_.push_back(velocities[i][a]);
            }
        m_.insert(m_.end(), masses.begin(), masses.begin() + add);
        particles_.resize(old + add);
        for (size_t i = old; i < old + add; ++i) {
            T F[9] = {T{1}, T{0}, T{0}, T{0}, T{1}, T{0}, T{0}, T{0}, T{1}};
            particles_.setDeformationGradient(i, F);
            particles_.setMass(i, m_[i]);
        }
        check(pg_mpm_set_particles(ctx_, (int)(old + add), x_.data(), v_.data(), nullptr, nullptr, m_.data(),
                                   nullptr, PG_HOST, nullptr), ctx_);
        host_fresh_ = false;
    }

    void step() {
        check(pg_mpm_step(ctx_, 1, nullptr), ctx_);
        current_time_ += params_.timestep;
        step_count_++;
        host_fresh_ = false;
    }

    const ParticleData& getParticles() const {
        const_cast<MPMSolver*>(this)->syncParticles();
        return particles_;
    }
    const Grid& getGrid() const {
        const_cast<MPMSolver*>(this)->syncGrid();
        return grid_;
    }
    T getCurrentTime() const { return current_time_; }
    size_t getStepCount() const { return step_count_; }

private:
    static void check(int rc, const pg_mpm_ctx* c) {
        if (rc != PG_OK) throw std::runtime_error(std::string("physgrad_mpm: ") + pg_mpm_last_error(c));
    }
    void syncParticles() {
        if (host_fresh_ || x_.empty()) return;
        check(pg_mpm_get_particles(ctx_, x_.data(), v_.data(), nullptr, nullptr, PG_HOST, nullptr), ctx_);
        for (size_t i = 0; i < x_.size() / 3; ++i) {
            particles_.setPosition(i, {x_[3 * i], x_[3 * i + 1], x_[3 * i + 2]});
            particles_.setVelocity(i, {v_[3 * i], v_[3 * i + 1], v_[3 * i + 2]});
        }
        host_fresh_ = true;
    }
    void syncGrid() {
        const size_t G = grid_.total_nodes;
        std::vector<T> m(G), p(3 * G), v(3 * G);
        check(pg_mpm_get_grid(ctx_, m.data(), p.data(), v.data(), PG_HOST, nullptr), ctx_);
        for (size_t i = 0; i < G; ++i) {
            grid_.mass[i] = m[i];
            grid_.momentum[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
            grid_.velocity[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        }
    }

    pg_mpm_ctx* ctx_ = nullptr;
    ParticleData particles_;
    Grid grid_;
    SimulationParams params_;
    std::vector<T> x_, v_, m_;
    bool host_fresh_ = true;
    T current_time_ = T{0};
    size_t step_count_ = 0;
};

}  // namespace mpm
}  // namespace physgrad

#endif
