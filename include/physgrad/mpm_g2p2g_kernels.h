// Drop-in for the reference's transfer-launcher header (src/mpm_g2p2g_kernels.h).
//
// Keeps the names a caller touches: namespace physgrad::mpm::kernels,
// KernelConfig (:68-75), FastShapeFunctions<T> (:167-205) and
// G2P2GKernelLauncher<T> with PerformanceConfig and launchG2P2G
// (:462-515).  launchG2P2G takes the host containers of the CPU solver header
// (ParticleAoSoA, MPMGrid from physgrad/material_point_method.h), as the
// reference does, and runs the launcher's sequence on the GPU through
// pg_launch_g2p2g: clear, P2G with gravity added to the particle velocity,
// v = p/m, un-normalised G2P gather, FLIP blend when use_affine_pic is set
// (pure PIC otherwise), x += v dt.  The particle masses are read from the
// container; positions, velocities and the grid's mass, momentum and velocity
// are written back.  One launcher works on one stream (the null stream); the
// reference's three unordered streams (:476-480) are not reproduced.
//
// Link with libphysgrad_mpm.so.  T must be float.
#ifndef PHYSGRAD_DROPIN_MPM_G2P2G_KERNELS_H
#define PHYSGRAD_DROPIN_MPM_G2P2G_KERNELS_H

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../physgrad_mpm.h"
#include "material_point_method.h"

namespace physgrad {
namespace mpm {
namespace kernels {

constexpr int WARP_SIZE = 64;   // a CDNA4 wavefront
constexpr int MAX_BLOCK_SIZE = 1024;
constexpr int SHARED_MEMORY_SIZE = 160 * 1024;   // LDS per CU on MI355X

struct KernelConfig {
    static constexpr int P2G_BLOCK_SIZE = 256;
    static constexpr int G2P_BLOCK_SIZE = 256;
    static constexpr int G2P2G_BLOCK_SIZE = 256;
    static constexpr int GRID_BLOCK_SIZE_X = 16;
    static constexpr int GRID_BLOCK_SIZE_Y = 16;
    static constexpr int GRID_BLOCK_SIZE_Z = 4;
};

// B-spline weights and their derivatives (quadratic support 1.5, cubic 2).
template <typename T>
struct FastShapeFunctions {
    static void evalQuadratic(T x, T& w, T& dw) {
        const T a = x < T{0} ? -x : x;
        if (a >= T{1.5}) {
            w = T{0};
            dw = T{0};
        } else if (a < T{0.5}) {
            w = T{0.75} - x * x;
            dw = T{-2} * x;
        } else {
            const T t = T{1.5} - a;
            w = T{0.5} * t * t;
            dw = x < T{0} ? t : -t;
        }
    }
    static void evalCubic(T x, T& w, T& dw) {
        const T a = x < T{0} ? -x : x;
        const T sgn = x < T{0} ? T{-1} : T{1};
        if (a >= T{2}) {
            w = T{0};
            dw = T{0};
        } else if (a < T{1}) {
            w = static_cast<T>(2.0 / 3.0) - a * a + T{0.5} * a * a * a;
            dw = sgn * (T{-2} * a + T{1.5} * a * a);
        } else {
            const T t = T{2} - a;
            w = static_cast<T>(1.0 / 6.0) * t * t * t;
            dw = -sgn * T{0.5} * t * t;
        }
    }
};

template <typename T>
class G2P2GKernelLauncher {
    static_assert(std::is_same<T, float>::value, "the GPU transfer path computes in float");

public:
    struct PerformanceConfig {
        bool use_kernel_fusion = true;
        bool use_shared_memory = true;
        bool use_coalesced_access = true;
        int max_particles_per_block = KernelConfig::G2P2G_BLOCK_SIZE;
        bool enable_async_execution = true;
    };

    G2P2GKernelLauncher(const PerformanceConfig& config = {}) : config_(config) {}

    // One transfer step (reference :494-515).  interpolation_order 2 or 3.
    template <size_t ChunkSize>
    void launchG2P2G(ParticleAoSoA<T, ChunkSize>& particles, MPMGrid<T>& grid, T dt,
                     const ConceptVector3D<T>& gravity, bool use_affine_pic, T flip_ratio,
                     int interpolation_order) {
        const size_t n = particles.num_particles;
        pos_.resize(4 * n);
        vel_.resize(4 * n);
        for (size_t p = 0; p < n; ++p) {
            const ConceptVector3D<T> x = particles.getPosition(p), v = particles.getVelocity(p);
            for (int a = 0; a < 3; ++a) {
                pos_[4 * p + a] = x[a];
                vel_[4 * p + a] = v[a];
            }
            pos_[4 * p + 3] = particles.getMass(p);
            vel_[4 * p + 3] = T{0};
        }
        const size_t G = grid.total_nodes;
        gm_.resize(G);
        gp_.resize(3 * G);
        gv_.resize(3 * G);
        const float o[3] = {grid.origin[0], grid.origin[1], grid.origin[2]};
        const float h[3] = {grid.cell_size[0], grid.cell_size[1], grid.cell_size[2]};
        const int d[3] = {grid.dimensions.x, grid.dimensions.y, grid.dimensions.z};
        const float g[3] = {gravity[0], gravity[1], gravity[2]};
        const int rc = pg_launch_g2p2g(pos_.data(), vel_.data(), gm_.data(), gp_.data(), gv_.data(), o, h, d,
                                       static_cast<int>(n), dt, g, use_affine_pic ? 1 : 0, flip_ratio,
                                       interpolation_order, PG_HOST, nullptr);
        if (rc != PG_OK)
            throw std::runtime_error(std::string("launchG2P2G: ") + pg_mpm_last_error(nullptr) +
                                     " (status " + std::to_string(rc) + ")");
        for (size_t p = 0; p < n; ++p) {
            particles.setPosition(p, {pos_[4 * p], pos_[4 * p + 1], pos_[4 * p + 2]});
            particles.setVelocity(p, {vel_[4 * p], vel_[4 * p + 1], vel_[4 * p + 2]});
        }
        for (size_t i = 0; i < G; ++i) {
            grid.mass[i] = gm_[i];
            grid.momentum[i] = {gp_[3 * i], gp_[3 * i + 1], gp_[3 * i + 2]};
            grid.velocity[i] = {gv_[3 * i], gv_[3 * i + 1], gv_[3 * i + 2]};
        }
    }

    const PerformanceConfig& config() const { return config_; }

private:
    PerformanceConfig config_;
    std::vector<float> pos_, vel_, gm_, gp_, gv_;   // host staging in the launcher's float4 layout
};

}  // namespace kernels
}  // namespace mpm
}  // namespace physgrad

#endif
