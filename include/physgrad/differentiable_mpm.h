// Drop-in for the reference's differentiable MPM header (src/differentiable_mpm.h).
//
// Keeps the names a caller touches: namespace physgrad::mpm::differentiable,
// GradientMode, CheckpointStrategy, DifferentiableMPMConfig<T> (:54-94),
// ParticleGradients<T> (chunk 64, component-major within a chunk, :103-259),
// GridGradients<T> (:264-340), TapeNode<T> / MPMGradientTape<T> (:349-467),
// DifferentiableMPMSystem<T> with initializeSystem, enableGradientComputation,
// simulationStep, computeGradients, the accessors and resetSimulation
// (:675-815), and verifyGradients (:873-940).
//
// The reference's backward is empty (computeGradients reverses a tape whose
// nodes do nothing).  Here simulationStep runs MLS-MPM steps on the GPU
// (APIC, F update, elastic stress from config.material, sticky walls) through
// one C-ABI context, and computeGradients runs the exact reverse of the steps
// recorded since enableGradientComputation(true) (or resetSimulation()):
// the rollout is replayed from its initial state with a checkpoint every
// config.checkpoint_every steps and reversed segment by segment on the GPU.
//   * The seed: before computeGradients, put dL/d(final x, v, F) into
//     getParticleGradients() (position, velocity, deformation-gradient
//     entries; e.g. from the objective).  computeGradients replaces them with
//     dL/d(initial x, v, F) and fills getYoungModulusGradient(p) and
//     getAffineGradient(p) (dL/d initial C) as well.  Mass, volume and stress
//     gradients stay zero (mass and volume are constants of the model; stress
//     is not a state).  User nodes recorded on the tape run afterwards, in
//     reverse order, as in the reference.
//   * dt and gravity must stay the same for every step of a recorded rollout.
//   * The particles' state lives on the GPU between steps.  getParticles()
//     (non-const) copies it to the host container and the next step uploads
//     the container again, so edits through the reference are seen.
//   * verifyGradients restores the initial particle state before every
//     objective evaluation (the reference re-runs the objective from the
//     state left by the previous run).
//
// Link with libphysgrad_mpm.so.  T must be float.
#ifndef PHYSGRAD_DROPIN_DIFFERENTIABLE_MPM_H
#define PHYSGRAD_DROPIN_DIFFERENTIABLE_MPM_H

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../physgrad_mpm.h"
#include "material_point_method.h"
#include "mpm_g2p2g_kernels.h"

namespace physgrad {
namespace mpm {
namespace differentiable {

enum class GradientMode { FORWARD, REVERSE, HYBRID };
enum class CheckpointStrategy { NONE, UNIFORM, ADAPTIVE, LOGARITHMIC };

template <typename T>
struct DifferentiableMPMConfig {
    GradientMode gradient_mode = GradientMode::REVERSE;
    CheckpointStrategy checkpoint_strategy = CheckpointStrategy::ADAPTIVE;
    T finite_difference_eps = T{1e-6};
    bool enable_second_order = false;
    bool use_checkpointing = true;
    size_t max_checkpoints = 100;
    T memory_budget_gb = T{4.0};
    T gradient_clip_threshold = T{1e3};
    bool enable_gradient_accumulation = true;
    bool parallelize_gradient_computation = true;
    size_t gradient_block_size = 256;

    // not in the reference: the model the steps integrate
    int material = PG_MAT_FIXED_COROTATED;   // pg_material
    T youngs_modulus = T{1e4};               // for particles whose E entry is 0
    T poisson_ratio = T{0.3};
    T particle_volume = T{0};                // 0: the particles' volume entries (must agree)
    int boundary_nodes = 3;                  // sticky wall thickness in nodes
    int checkpoint_every = 32;               // steps between checkpoints (UNIFORM)
};

// Steps between checkpoints of a recorded rollout of `steps` steps, as the
// config's strategy asks (the reverse pass recomputes every segment but the
// last from its checkpoint, and holds one segment's tape):
//   NONE, or use_checkpointing off: one segment (no recompute, the whole tape);
//   UNIFORM: checkpoint_every;
//   ADAPTIVE: the longest segment whose device memory (checkpoints plus one
//     segment's tape, estimated from ckpt_bytes and step_bytes) fits
//     memory_budget_gb, with at most max_checkpoints checkpoints when that also
//     fits; the least memory when nothing fits;
//   LOGARITHMIC: ceil(sqrt(steps)), the single-level spacing with the least
//     memory (the reference declares the strategy without implementing it; a
//     multi-level schedule is not implemented here either).
// ckpt_bytes: device bytes of one checkpoint; step_bytes: of one tape step.
template <typename T>
int checkpointInterval(const DifferentiableMPMConfig<T>& cfg, int steps, double ckpt_bytes, double step_bytes) {
    if (steps <= 1) return 1;
    if (!cfg.use_checkpointing || cfg.checkpoint_strategy == CheckpointStrategy::NONE) return steps;
    switch (cfg.checkpoint_strategy) {
        case CheckpointStrategy::UNIFORM:
            return std::min(steps, std::max(1, cfg.checkpoint_every));
        case CheckpointStrategy::LOGARITHMIC:
            return std::max(1, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(steps)))));
        default: {
            auto bytes = [&](int K) { return ((steps + K - 1) / K) * ckpt_bytes + (K + 1.0) * step_bytes; };
            const double budget = static_cast<double>(cfg.memory_budget_gb) * 1e9;
            const size_t maxc = std::max<size_t>(1, cfg.max_checkpoints);
            int best = -1, least = 1;
            for (int K = steps; K >= 1; --K) {
                if (bytes(K) < bytes(least)) least = K;
                if (best < 0 && bytes(K) <= budget && static_cast<size_t>((steps + K - 1) / K) <= maxc) best = K;
            }
            if (best < 0)   // the checkpoint cap cannot be met within the budget: the budget rules
                for (int K = steps; K >= 1 && best < 0; --K)
                    if (bytes(K) <= budget) best = K;
            return best > 0 ? best : least;
        }
    }
}

template <typename T>
class ParticleGradients {
public:
    using vector_type = ConceptVector3D<T>;
    static constexpr size_t chunk_size = 64;

    ParticleGradients() = default;

    void resize(size_t n) {
        num_particles_ = n;
        const size_t c = (n + chunk_size - 1) / chunk_size;
        position_gradients_.resize(c);
        velocity_gradients_.resize(c);
        mass_gradients_.resize(c);
        deformation_grad_gradients_.resize(c);
        volume_gradients_.resize(c);
        stress_gradients_.resize(c);
        young_modulus_gradients_.resize(c);
        affine_gradients_.resize(c);
        zero();
    }

    vector_type getPositionGradient(size_t p) const { return get3(position_gradients_, p); }
    void setPositionGradient(size_t p, const vector_type& g) { set3(position_gradients_, p, g); }
    vector_type getVelocityGradient(size_t p) const { return get3(velocity_gradients_, p); }
    void setVelocityGradient(size_t p, const vector_type& g) { set3(velocity_gradients_, p, g); }
    T getMassGradient(size_t p) const { return mass_gradients_[p / chunk_size][p % chunk_size]; }
    void setMassGradient(size_t p, T g) { mass_gradients_[p / chunk_size][p % chunk_size] = g; }
    void accumulatePositionGradient(size_t p, const vector_type& g) {
        setPositionGradient(p, getPositionGradient(p) + g);
    }
    void accumulateVelocityGradient(size_t p, const vector_type& g) {
        setVelocityGradient(p, getVelocityGradient(p) + g);
    }
    // dL/dF, row-major 3x3
    void getDeformationGradientGradient(size_t p, T* g) const { get9(deformation_grad_gradients_, p, g); }
    void setDeformationGradientGradient(size_t p, const T* g) { set9(deformation_grad_gradients_, p, g); }
    // not in the reference: dL/dE and dL/dC (row-major 3x3)
    T getYoungModulusGradient(size_t p) const { return young_modulus_gradients_[p / chunk_size][p % chunk_size]; }
    void setYoungModulusGradient(size_t p, T g) { young_modulus_gradients_[p / chunk_size][p % chunk_size] = g; }
    void getAffineGradient(size_t p, T* g) const { get9(affine_gradients_, p, g); }
    void setAffineGradient(size_t p, const T* g) { set9(affine_gradients_, p, g); }

    void zero() {
        for (auto& c : position_gradients_) c.fill(T{0});
        for (auto& c : velocity_gradients_) c.fill(T{0});
        for (auto& c : mass_gradients_) c.fill(T{0});
        for (auto& c : deformation_grad_gradients_) c.fill(T{0});
        for (auto& c : volume_gradients_) c.fill(T{0});
        for (auto& c : stress_gradients_) c.fill(T{0});
        for (auto& c : young_modulus_gradients_) c.fill(T{0});
        for (auto& c : affine_gradients_) c.fill(T{0});
    }
    size_t size() const { return num_particles_; }
    size_t getMemoryUsageBytes() const {
        return position_gradients_.size() * sizeof(position_gradients_[0]) +
               velocity_gradients_.size() * sizeof(velocity_gradients_[0]) +
               mass_gradients_.size() * sizeof(mass_gradients_[0]) +
               deformation_grad_gradients_.size() * sizeof(deformation_grad_gradients_[0]) +
               volume_gradients_.size() * sizeof(volume_gradients_[0]) +
               stress_gradients_.size() * sizeof(stress_gradients_[0]) +
               young_modulus_gradients_.size() * sizeof(young_modulus_gradients_[0]) +
               affine_gradients_.size() * sizeof(affine_gradients_[0]);
    }

private:
    using C3 = std::vector<std::array<T, chunk_size * 3>>;
    using C9 = std::vector<std::array<T, chunk_size * 9>>;
    static vector_type get3(const C3& v, size_t p) {
        const auto& c = v[p / chunk_size];
        const size_t l = p % chunk_size;
        return {c[l], c[l + chunk_size], c[l + 2 * chunk_size]};
    }
    static void set3(C3& v, size_t p, const vector_type& g) {
        auto& c = v[p / chunk_size];
        const size_t l = p % chunk_size;
        c[l] = g[0]; c[l + chunk_size] = g[1]; c[l + 2 * chunk_size] = g[2];
    }
    static void get9(const C9& v, size_t p, T* g) {
        const auto& c = v[p / chunk_size];
        for (int i = 0; i < 9; ++i) g[i] = c[p % chunk_size + i * chunk_size];
    }
    static void set9(C9& v, size_t p, const T* g) {
        auto& c = v[p / chunk_size];
        for (int i = 0; i < 9; ++i) c[p % chunk_size + i * chunk_size] = g[i];
    }

    size_t num_particles_ = 0;
    C3 position_gradients_, velocity_gradients_;
    std::vector<std::array<T, chunk_size>> mass_gradients_;
    C9 deformation_grad_gradients_;
    std::vector<std::array<T, chunk_size>> volume_gradients_;
    std::vector<std::array<T, chunk_size * 6>> stress_gradients_;
    std::vector<std::array<T, chunk_size>> young_modulus_gradients_;
    C9 affine_gradients_;
};

template <typename T>
class GridGradients {
public:
    using vector_type = ConceptVector3D<T>;

    GridGradients() = default;
    explicit GridGradients(const int3& dims) : grid_dims_(dims) {
        resize(static_cast<size_t>(dims.x) * dims.y * dims.z);
    }
    void resize(size_t n) {
        total_nodes_ = n;
        velocity_gradients_.assign(n, vector_type{});
        momentum_gradients_.assign(n, vector_type{});
        force_gradients_.assign(n, vector_type{});
        mass_gradients_.assign(n, T{0});
    }
    vector_type getVelocityGradient(size_t i) const { return velocity_gradients_[i]; }
    void setVelocityGradient(size_t i, const vector_type& g) { velocity_gradients_[i] = g; }
    vector_type getMomentumGradient(size_t i) const { return momentum_gradients_[i]; }
    void setMomentumGradient(size_t i, const vector_type& g) { momentum_gradients_[i] = g; }
    void accumulateVelocityGradient(size_t i, const vector_type& g) { velocity_gradients_[i] = velocity_gradients_[i] + g; }
    void accumulateMomentumGradient(size_t i, const vector_type& g) { momentum_gradients_[i] = momentum_gradients_[i] + g; }
    void zero() {
        std::fill(velocity_gradients_.begin(), velocity_gradients_.end(), vector_type{});
        std::fill(momentum_gradients_.begin(), momentum_gradients_.end(), vector_type{});
        std::fill(force_gradients_.begin(), force_gradients_.end(), vector_type{});
        std::fill(mass_gradients_.begin(), mass_gradients_.end(), T{0});
    }
    size_t size() const { return total_nodes_; }
    size_t getMemoryUsageBytes() const {
        return velocity_gradients_.size() * sizeof(vector_type) * 3 + mass_gradients_.size() * sizeof(T);
    }

private:
    int3 grid_dims_;
    size_t total_nodes_ = 0;
    std::vector<vector_type> velocity_gradients_, momentum_gradients_, force_gradients_;
    std::vector<T> mass_gradients_;
};

template <typename T>
struct TapeNode {
    enum class OperationType {
        P2G_TRANSFER,
        G2P_TRANSFER,
        GRID_UPDATE,
        CONSTITUTIVE_UPDATE,
        FORCE_COMPUTATION,
        BOUNDARY_CONDITIONS
    };
    OperationType operation;
    size_t timestep;
    std::vector<size_t> input_indices, output_indices;
    std::vector<T> intermediate_data;
    std::function<void(const TapeNode<T>&, ParticleGradients<T>&, GridGradients<T>&)> backward_fn;
    TapeNode(OperationType op, size_t step) : operation(op), timestep(step) {}
};

// A record of operations whose backward functions run in reverse order.
template <typename T>
class MPMGradientTape {
public:
    explicit MPMGradientTape(size_t max_memory_mb = 1000) : max_memory_bytes_(max_memory_mb * 1024 * 1024) {}

    void startRecording() {
        recording_ = true;
        clear();
    }
    void stopRecording() { recording_ = false; }
    bool isRecording() const { return recording_; }
    void nextTimestep() { current_timestep_++; }

    void recordOperation(typename TapeNode<T>::OperationType op, const std::vector<size_t>& inputs,
                         const std::vector<size_t>& outputs, const std::vector<T>& intermediate,
                         std::function<void(const TapeNode<T>&, ParticleGradients<T>&, GridGradients<T>&)> fn) {
        if (!recording_) return;
        TapeNode<T> node(op, current_timestep_);
        node.input_indices = inputs;
        node.output_indices = outputs;
        node.intermediate_data = intermediate;
        node.backward_fn = std::move(fn);
        memory_usage_bytes_ += sizeof(TapeNode<T>) + (inputs.size() + outputs.size()) * sizeof(size_t) +
                               intermediate.size() * sizeof(T);
        if (memory_usage_bytes_ > max_memory_bytes_ / 4) createCheckpoint();
        tape_.push_back(std::move(node));
    }
    void backward(ParticleGradients<T>& pg, GridGradients<T>& gg) {
        for (auto it = tape_.rbegin(); it != tape_.rend(); ++it)
            if (it->backward_fn) it->backward_fn(*it, pg, gg);
    }
    void createCheckpoint() { checkpoint_indices_.push_back(tape_.size()); }
    void clear() {
        tape_.clear();
        checkpoint_indices_.clear();
        memory_usage_bytes_ = 0;
        current_timestep_ = 0;
    }
    size_t size() const { return tape_.size(); }
    size_t getMemoryUsage() const { return memory_usage_bytes_; }
    const TapeNode<T>& getNode(size_t i) const { return tape_[i]; }
    // bytes of device memory held for the last reverse pass (checkpoints + segment tape)
    void addDeviceBytes(size_t b) { memory_usage_bytes_ += b; }

private:
    std::vector<TapeNode<T>> tape_;
    std::vector<size_t> checkpoint_indices_;
    bool recording_ = false;
    size_t current_timestep_ = 0;
    size_t memory_usage_bytes_ = 0;
    size_t max_memory_bytes_;
};

template <typename T>
class DifferentiableMPMSystem {
    static_assert(std::is_same<T, float>::value, "the GPU MPM path computes in float");

public:
    // The full particle state (host copy), for save / restore around evaluations.
    struct State {
        std::vector<T> x, v, C, F;
    };

    explicit DifferentiableMPMSystem(const DifferentiableMPMConfig<T>& config = DifferentiableMPMConfig<T>{})
        : config_(config), gradient_tape_(static_cast<size_t>(config.memory_budget_gb * 1000)) {
        typename kernels::G2P2GKernelLauncher<T>::PerformanceConfig kc;
        kernel_launcher_ = std::make_unique<kernels::G2P2GKernelLauncher<T>>(kc);
    }
    ~DifferentiableMPMSystem() { destroyContext(); }
    DifferentiableMPMSystem(const DifferentiableMPMSystem&) = delete;
    DifferentiableMPMSystem& operator=(const DifferentiableMPMSystem&) = delete;

    void initializeSystem(const int3& grid_dims, const ConceptVector3D<T>& cell_size,
                          const ConceptVector3D<T>& origin, size_t num_particles) {
        destroyContext();
        grid_ = std::make_unique<MPMGrid<T>>(grid_dims, cell_size, origin);
        particles_ = std::make_unique<ParticleAoSoA<T>>();
        particles_->resize(num_particles);
        const T I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (size_t p = 0; p < num_particles; ++p) particles_->setDeformationGradient(p, I);
        particle_gradients_ = std::make_unique<ParticleGradients<T>>();
        particle_gradients_->resize(num_particles);
        grid_gradients_ = std::make_unique<GridGradients<T>>(grid_dims);
        C_.assign(9 * num_particles, T{0});
        host_dirty_ = true;
    }

    void enableGradientComputation(bool enable = true) {
        gradient_computation_enabled_ = enable;
        if (enable) gradient_tape_.startRecording();
        else gradient_tape_.stopRecording();
        recorded_steps_ = 0;
    }

    void simulationStep(T dt, const ConceptVector3D<T>& gravity = {T{0}, T{-9.81}, T{0}}) {
        if (!particles_ || !grid_) throw std::runtime_error("System not initialized");
        ensureContext(dt, gravity);
        if (gradient_computation_enabled_) {
            if (recorded_steps_ == 0) {
                pull();
                initial_ = captureState();
            }
            gradient_tape_.nextTimestep();
            recorded_steps_++;
        }
        if (host_dirty_) push();
        check(pg_mpm_step(ctx_, 1, nullptr));
        device_ahead_ = true;
        current_time_ += dt;
        current_step_++;
    }

    // Reverse of the recorded rollout; see the header comment for the seed.
    void computeGradients() {
        if (!gradient_computation_enabled_) throw std::runtime_error("Gradient computation not enabled");
        const size_t n = particles_->size();
        std::vector<T> ax(3 * n), av(3 * n), aF(9 * n), gx(3 * n), gv(3 * n), gC(9 * n), gF(9 * n), gE(n);
        for (size_t p = 0; p < n; ++p) {
            const auto a = particle_gradients_->getPositionGradient(p), b = particle_gradients_->getVelocityGradient(p);
            for (int c = 0; c < 3; ++c) {
                ax[3 * p + c] = a[c];
                av[3 * p + c] = b[c];
            }
            particle_gradients_->getDeformationGradientGradient(p, &aF[9 * p]);
        }
        particle_gradients_->zero();
        grid_gradients_->zero();
        if (recorded_steps_ == 0) {   // nothing recorded: the gradient is the seed itself
            gx = ax; gv = av; gF = aF;
        } else {
            pull();
            const State final_state = captureState();
            restoreState(initial_);
            push();
            // checkpoint: state, caller index, E and mass per particle; tape step:
            // the same plus the binning's permutation and source maps, and the
            // step's grid (m, p) and (v, m), on each of the two tapes (the
            // backward recomputes one segment while it reverses the other)
            const double ckpt_bytes = 108.0 * n;
            const double nodes = static_cast<double>(grid_->dims.x) * grid_->dims.y * grid_->dims.z;
            const double step_bytes = 2.0 * (116.0 * n + 32.0 * nodes);
            const int K = checkpointInterval(config_, static_cast<int>(recorded_steps_), ckpt_bytes, step_bytes);
            check(pg_mpm_forward_checkpointed(ctx_, static_cast<int>(recorded_steps_), K, nullptr));
            check(pg_mpm_backward(ctx_, ax.data(), av.data(), nullptr, aF.data(), gx.data(), gv.data(), gC.data(),
                                  gF.data(), gE.data(), nullptr));
            // keep the recorded end state (the replay may differ from it in the
            // last bits: P2G sums are accumulated in a different order)
            restoreState(final_state);
            gradient_tape_.addDeviceBytes(static_cast<size_t>(((recorded_steps_ + K - 1) / K) * ckpt_bytes +
                                                              (K + 1.0) * step_bytes));
        }
        for (size_t p = 0; p < n; ++p) {
            particle_gradients_->setPositionGradient(p, {gx[3 * p], gx[3 * p + 1], gx[3 * p + 2]});
            particle_gradients_->setVelocityGradient(p, {gv[3 * p], gv[3 * p + 1], gv[3 * p + 2]});
            particle_gradients_->setDeformationGradientGradient(p, &gF[9 * p]);
            particle_gradients_->setAffineGradient(p, &gC[9 * p]);
            particle_gradients_->setYoungModulusGradient(p, gE[p]);
        }
        gradient_tape_.backward(*particle_gradients_, *grid_gradients_);
    }

    ParticleAoSoA<T>& getParticles() {
        pull();
        host_dirty_ = true;   // the caller may edit the container
        return *particles_;
    }
    const ParticleAoSoA<T>& getParticles() const {
        const_cast<DifferentiableMPMSystem*>(this)->pull();
        return *particles_;
    }
    MPMGrid<T>& getGrid() {
        syncGrid();
        return *grid_;
    }
    const MPMGrid<T>& getGrid() const {
        const_cast<DifferentiableMPMSystem*>(this)->syncGrid();
        return *grid_;
    }
    ParticleGradients<T>& getParticleGradients() { return *particle_gradients_; }
    const ParticleGradients<T>& getParticleGradients() const { return *particle_gradients_; }
    GridGradients<T>& getGridGradients() { return *grid_gradients_; }
    const GridGradients<T>& getGridGradients() const { return *grid_gradients_; }

    T getCurrentTime() const { return current_time_; }
    size_t getCurrentStep() const { return current_step_; }
    size_t getGradientTapeSize() const { return gradient_tape_.size(); }
    size_t getGradientMemoryUsage() const {
        return gradient_tape_.getMemoryUsage() + particle_gradients_->getMemoryUsageBytes() +
               grid_gradients_->getMemoryUsageBytes();
    }

    void resetSimulation() {
        current_time_ = T{0};
        current_step_ = 0;
        gradient_tape_.clear();
        recorded_steps_ = 0;
        if (particle_gradients_) particle_gradients_->zero();
        if (grid_gradients_) grid_gradients_->zero();
    }

    // not in the reference
    State captureState() const {
        const_cast<DifferentiableMPMSystem*>(this)->pull();
        const size_t n = particles_->size();
        State s;
        s.x.resize(3 * n);
        s.v.resize(3 * n);
        s.F.resize(9 * n);
        s.C = C_;
        for (size_t p = 0; p < n; ++p) {
            const auto x = particles_->getPosition(p), v = particles_->getVelocity(p);
            for (int c = 0; c < 3; ++c) {
                s.x[3 * p + c] = x[c];
                s.v[3 * p + c] = v[c];
            }
            particles_->getDeformationGradient(p, &s.F[9 * p]);
        }
        return s;
    }
    void restoreState(const State& s) {
        const size_t n = particles_->size();
        if (s.x.size() != 3 * n) throw std::runtime_error("restoreState: particle count differs");
        for (size_t p = 0; p < n; ++p) {
            particles_->setPosition(p, {s.x[3 * p], s.x[3 * p + 1], s.x[3 * p + 2]});
            particles_->setVelocity(p, {s.v[3 * p], s.v[3 * p + 1], s.v[3 * p + 2]});
            particles_->setDeformationGradient(p, &s.F[9 * p]);
        }
        C_ = s.C;
        device_ahead_ = false;
        host_dirty_ = true;
    }
    size_t recordedSteps() const { return recorded_steps_; }

private:
    void check(int rc) const {
        if (rc != PG_OK) throw std::runtime_error(std::string("physgrad_mpm: ") + pg_mpm_last_error(ctx_));
    }
    void destroyContext() {
        if (ctx_) pg_mpm_destroy(ctx_);
        ctx_ = nullptr;
    }

    void ensureContext(T dt, const ConceptVector3D<T>& g) {
        if (ctx_ && dt == dt_ && g[0] == g_[0] && g[1] == g_[1] && g[2] == g_[2]) return;
        if (ctx_ && gradient_computation_enabled_ && recorded_steps_ > 0)
            throw std::runtime_error("simulationStep: dt and gravity must stay fixed within a recorded rollout");
        pull();
        destroyContext();
        const MPMGrid<T>& G = *grid_;
        if (!(G.cell_size[0] == G.cell_size[1] && G.cell_size[1] == G.cell_size[2]))
            throw std::runtime_error("DifferentiableMPMSystem: the MLS-MPM path needs cubic cells");
        const size_t n = particles_->size();
        T vol = config_.particle_volume;
        if (!(vol > T{0})) {
            vol = n ? particles_->volume_chunks[0][0] : T{1e-6};
            for (size_t p = 0; p < n; ++p) {
                const T v = particles_->volume_chunks[p / ParticleAoSoA<T>::chunk_size][p % ParticleAoSoA<T>::chunk_size];
                if (std::fabs(v - vol) > T{1e-5} * vol)
                    throw std::runtime_error("DifferentiableMPMSystem: particle volumes differ; set config.particle_volume");
            }
            if (!(vol > T{0})) throw std::runtime_error("DifferentiableMPMSystem: particle volume must be > 0");
        }
        pg_mpm_params p;
        pg_mpm_default_params(&p);
        p.mode = PG_MODE_MLS_MPM;
        p.impl = PG_IMPL_TILED;
        p.nx = G.dimensions.x; p.ny = G.dimensions.y; p.nz = G.dimensions.z;
        p.dx = p.dy = p.dz = G.cell_size[0];
        p.ox = G.origin[0]; p.oy = G.origin[1]; p.oz = G.origin[2];
        p.dt = dt;
        p.gx = g[0]; p.gy = g[1]; p.gz = g[2];
        p.material = config_.material;
        p.youngs_modulus = config_.youngs_modulus;
        p.poisson_ratio = config_.poisson_ratio;
        p.particle_volume = vol;
        p.density = n && particles_->getMass(0) > T{0} ? particles_->getMass(0) / vol : T{1000};
        p.boundary_thickness = config_.boundary_nodes;
        p.max_particles = static_cast<int>(std::max<size_t>(n, 1));
        check(pg_mpm_create(&p, &ctx_));
        dt_ = dt;
        g_ = g;
        host_dirty_ = true;
    }

    void push() {
        const size_t n = particles_->size();
        std::vector<T> x(3 * n), v(3 * n), F(9 * n), m(n), E(n);
        const size_t cs = ParticleAoSoA<T>::chunk_size;
        for (size_t p = 0; p < n; ++p) {
            const auto xp = particles_->getPosition(p), vp = particles_->getVelocity(p);
            for (int c = 0; c < 3; ++c) {
                x[3 * p + c] = xp[c];
                v[3 * p + c] = vp[c];
            }
            particles_->getDeformationGradient(p, &F[9 * p]);
            m[p] = particles_->getMass(p);
            const T e = particles_->young_modulus_chunks[p / cs][p % cs];
            E[p] = e > T{0} ? e : config_.youngs_modulus;
        }
        C_.resize(9 * n, T{0});
        check(pg_mpm_set_particles(ctx_, static_cast<int>(n), x.data(), v.data(), C_.data(), F.data(), m.data(),
                                   E.data(), PG_HOST, nullptr));
        host_dirty_ = false;
        device_ahead_ = false;
    }

    void pull() {
        if (!ctx_ || !device_ahead_) return;
        const size_t n = particles_->size();
        std::vector<T> x(3 * n), v(3 * n), F(9 * n);
        C_.resize(9 * n);
        check(pg_mpm_get_particles(ctx_, x.data(), v.data(), C_.data(), F.data(), PG_HOST, nullptr));
        for (size_t p = 0; p < n; ++p) {
            particles_->setPosition(p, {x[3 * p], x[3 * p + 1], x[3 * p + 2]});
            particles_->setVelocity(p, {v[3 * p], v[3 * p + 1], v[3 * p + 2]});
            particles_->setDeformationGradient(p, &F[9 * p]);
        }
        device_ahead_ = false;
    }

    void syncGrid() {
        if (!ctx_) return;
        const size_t G = grid_->total_nodes;
        std::vector<T> m(G), v(3 * G);
        check(pg_mpm_get_grid(ctx_, m.data(), nullptr, v.data(), PG_HOST, nullptr));
        for (size_t i = 0; i < G; ++i) {
            grid_->mass[i] = m[i];
            grid_->velocity[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
            grid_->momentum[i] = {m[i] * v[3 * i], m[i] * v[3 * i + 1], m[i] * v[3 * i + 2]};
        }
    }

    DifferentiableMPMConfig<T> config_;
    MPMGradientTape<T> gradient_tape_;
    std::unique_ptr<ParticleAoSoA<T>> particles_;
    std::unique_ptr<MPMGrid<T>> grid_;
    std::unique_ptr<kernels::G2P2GKernelLauncher<T>> kernel_launcher_;
    std::unique_ptr<ParticleGradients<T>> particle_gradients_;
    std::unique_ptr<GridGradients<T>> grid_gradients_;
    pg_mpm_ctx* ctx_ = nullptr;
    T dt_ = T{0};
    ConceptVector3D<T> g_;
    std::vector<T> C_;
    State initial_;
    size_t recorded_steps_ = 0;
    bool host_dirty_ = true, device_ahead_ = false;
    T current_time_ = T{0};
    size_t current_step_ = 0;
    bool gradient_computation_enabled_ = false;
};

// Central finite differences of the objective against the adjoint for the
// first min(n, 10) particles' positions (reference :873-940).  The objective
// runs a rollout with simulationStep and leaves the seed dL/d(final state) in
// getParticleGradients(); every evaluation starts from the same initial state.
template <typename T>
bool verifyGradients(DifferentiableMPMSystem<T>& system, const std::function<T()>& objective_function,
                     T epsilon = T{1e-6}, T tolerance = T{1e-3}) {
    const auto s0 = system.captureState();
    const size_t n = system.getParticles().size();
    system.resetSimulation();
    system.enableGradientComputation(true);
    objective_function();
    system.computeGradients();
    const size_t m = std::min(n, size_t{10});
    std::vector<ConceptVector3D<T>> analytic(m);
    for (size_t p = 0; p < m; ++p) analytic[p] = system.getParticleGradients().getPositionGradient(p);
    auto eval = [&](size_t p, int d, T delta) {
        system.restoreState(s0);
        auto& P = system.getParticles();
        auto x = P.getPosition(p);
        x[d] += delta;
        P.setPosition(p, x);
        system.resetSimulation();
        return objective_function();
    };
    bool ok = true;
    for (size_t p = 0; p < m; ++p)
        for (int d = 0; d < 3; ++d) {
            const T fp = eval(p, d, epsilon), fm = eval(p, d, -epsilon);
            const T num = (fp - fm) / (2 * epsilon), ana = analytic[p][d];
            const T rel = std::fabs(num - ana) / (std::fabs(num) + T{1e-8});
            if (rel > tolerance) {
                ok = false;
                std::cout << "Gradient verification failed for particle " << p << " dimension " << d
                          << ": numerical=" << num << " analytical=" << ana << " rel_error=" << rel << "\n";
            }
        }
    system.restoreState(s0);
    system.resetSimulation();
    return ok;
}

}  // namespace differentiable
}  // namespace mpm
}  // namespace physgrad

#endif
