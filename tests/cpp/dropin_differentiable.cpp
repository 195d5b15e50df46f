// Drives the launcher and differentiable-system drop-ins
// (include/physgrad/mpm_g2p2g_kernels.h, include/physgrad/differentiable_mpm.h).
//
//   dropin_differentiable launcher in.bin out.bin
//     in: int32 nx ny nz, n, steps, order, use_affine_pic; float32 o[3], h[3],
//         dt, g[3], flip, pos4[4n] (x y z mass), vel3[3n]
//     out: float32 x[3n], v[3n], grid mass[G], grid velocity[3G]
//   dropin_differentiable diff in.bin out.bin
//     in: int32 nx, n, steps, K, material; float32 dt, E, nu, vol, g[3],
//         x[3n], v[3n], mass[n], Ep[n], seed_x[3n], seed_v[3n]
//     out: float32 gx[3n], gv[3n], gF[9n], gC[9n], gE[n]; prints verifyGradients' result
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "physgrad/differentiable_mpm.h"

using namespace physgrad;
using namespace physgrad::mpm;

template <typename T>
static bool rd(FILE* f, T* p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }

static int launcher(const char* in, const char* out) {
    FILE* f = std::fopen(in, "rb");
    if (!f) return 2;
    int h[7];
    float fl[11];
    if (!rd(f, h, 7) || !rd(f, fl, 11)) return 2;
    const int n = h[3], steps = h[4], order = h[5], affine = h[6];
    std::vector<float> pos(4 * n), vel(3 * n);
    if (!rd(f, pos.data(), pos.size()) || !rd(f, vel.data(), vel.size())) return 2;
    std::fclose(f);
    ParticleAoSoA<float> parts;
    parts.resize(n);
    for (int i = 0; i < n; ++i) {
        parts.setPosition(i, {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2]});
        parts.setMass(i, pos[4 * i + 3]);
        parts.setVelocity(i, {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]});
    }
    MPMGrid<float> grid(int3{h[0], h[1], h[2]}, {fl[3], fl[4], fl[5]}, {fl[0], fl[1], fl[2]});
    kernels::G2P2GKernelLauncher<float> L;
    for (int s = 0; s < steps; ++s)
        L.launchG2P2G(parts, grid, fl[6], {fl[7], fl[8], fl[9]}, affine != 0, fl[10], order);
    FILE* o = std::fopen(out, "wb");
    if (!o) return 2;
    for (int i = 0; i < n; ++i) {
        const auto x = parts.getPosition(i);
        std::fwrite(&x[0], 4, 3, o);
    }
    for (int i = 0; i < n; ++i) {
        const auto v = parts.getVelocity(i);
        std::fwrite(&v[0], 4, 3, o);
    }
    std::fwrite(grid.mass.data(), 4, grid.total_nodes, o);
    for (size_t i = 0; i < grid.total_nodes; ++i) std::fwrite(&grid.velocity[i][0], 4, 3, o);
    std::fclose(o);
    std::printf("launcher steps %d\n", steps);
    return 0;
}

static int diff(const char* in, const char* out) {
    using namespace physgrad::mpm::differentiable;
    FILE* f = std::fopen(in, "rb");
    if (!f) return 2;
    int h[5];
    float fl[7];
    if (!rd(f, h, 5) || !rd(f, fl, 7)) return 2;
    const int nx = h[0], n = h[1], steps = h[2], K = h[3], material = h[4];
    std::vector<float> x(3 * n), v(3 * n), m(n), Ep(n), sx(3 * n), sv(3 * n);
    if (!rd(f, x.data(), x.size()) || !rd(f, v.data(), v.size()) || !rd(f, m.data(), m.size()) ||
        !rd(f, Ep.data(), Ep.size()) || !rd(f, sx.data(), sx.size()) || !rd(f, sv.data(), sv.size()))
        return 2;
    std::fclose(f);
    const float dt = fl[0];
    const ConceptVector3D<float> g{fl[4], fl[5], fl[6]};

    DifferentiableMPMConfig<float> cfg;
    cfg.material = material;
    cfg.youngs_modulus = fl[1];
    cfg.poisson_ratio = fl[2];
    cfg.particle_volume = fl[3];
    cfg.checkpoint_strategy = CheckpointStrategy::UNIFORM;   // recompute every K steps
    cfg.checkpoint_every = K;
    DifferentiableMPMSystem<float> sys(cfg);
    sys.initializeSystem(int3{nx, nx, nx}, {1.0f / nx, 1.0f / nx, 1.0f / nx}, {0.f, 0.f, 0.f}, n);
    {
        auto& P = sys.getParticles();
        for (int i = 0; i < n; ++i) {
            P.setPosition(i, {x[3 * i], x[3 * i + 1], x[3 * i + 2]});
            P.setVelocity(i, {v[3 * i], v[3 * i + 1], v[3 * i + 2]});
            P.setMass(i, m[i]);
            P.young_modulus_chunks[i / 64][i % 64] = Ep[i];
        }
    }
    // the rollout and its seed
    auto objective = [&]() -> float {
        for (int s = 0; s < steps; ++s) sys.simulationStep(dt, g);
        const auto& P = static_cast<const DifferentiableMPMSystem<float>&>(sys).getParticles();
        double L = 0.0;
        for (int i = 0; i < n; ++i) {
            const auto xp = P.getPosition(i), vp = P.getVelocity(i);
            // offsets by the (constant) initial values keep L small, so that its
            // float rounding stays far below the central differences
            for (int c = 0; c < 3; ++c)
                L += (double)sx[3 * i + c] * (xp[c] - x[3 * i + c]) + (double)sv[3 * i + c] * (vp[c] - v[3 * i + c]);
            sys.getParticleGradients().setPositionGradient(i, {sx[3 * i], sx[3 * i + 1], sx[3 * i + 2]});
            sys.getParticleGradients().setVelocityGradient(i, {sv[3 * i], sv[3 * i + 1], sv[3 * i + 2]});
        }
        return static_cast<float>(L);
    };
    sys.enableGradientComputation(true);
    objective();
    sys.computeGradients();
    const auto& G = sys.getParticleGradients();
    std::vector<float> gx(3 * n), gv(3 * n), gF(9 * n), gC(9 * n), gE(n);
    for (int i = 0; i < n; ++i) {
        const auto a = G.getPositionGradient(i), b = G.getVelocityGradient(i);
        for (int c = 0; c < 3; ++c) {
            gx[3 * i + c] = a[c];
            gv[3 * i + c] = b[c];
        }
        G.getDeformationGradientGradient(i, &gF[9 * i]);
        G.getAffineGradient(i, &gC[9 * i]);
        gE[i] = G.getYoungModulusGradient(i);
    }
    std::printf("steps %zu recorded %zu\n", sys.getCurrentStep(), sys.recordedSteps());
    FILE* o = std::fopen(out, "wb");
    if (!o) return 2;
    std::fwrite(gx.data(), 4, gx.size(), o);
    std::fwrite(gv.data(), 4, gv.size(), o);
    std::fwrite(gF.data(), 4, gF.size(), o);
    std::fwrite(gC.data(), 4, gC.size(), o);
    std::fwrite(gE.data(), 4, gE.size(), o);
    std::fclose(o);
    // restart from the initial state: verifyGradients drives the same rollout
    const bool ok = verifyGradients<float>(sys, objective, 1e-3f, 2e-2f);
    std::printf("verifyGradients %s\n", ok ? "passed" : "failed");
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && !std::strcmp(argv[1], "launcher")) return launcher(argv[2], argv[3]);
    if (argc == 4 && !std::strcmp(argv[1], "diff")) return diff(argv[2], argv[3]);
    return 2;
}
