// ref-compat mode: the reference CPU solver's step on the GPU.
//
// Semantics of MPMSolver<T,64>::step (src/material_point_method.h:479-666):
// PIC transfer with the reference's stencil (base = floor((x-o)/h), offsets
// {-k/2 .. k-k/2-1} for kernel size k = 2,3,4 by interpolation order), the
// weight threshold 1e-10, gravity on grid momentum, v = p/m, unnormalised G2P,
// then x += v dt.  Floating-point contraction is disabled in this file so
// every per-contribution value is the same float the reference computes; only
// the order of the P2G sums differs (hardware float atomics), which keeps the
// step within 1e-5 of the reference.
//
// Layout: x, v [3N] xyz-interleaved, m [N]; grid mass [G], momentum [3G],
// velocity [3G] (AoS), node (k*ny + j)*nx + i: the reference's own layouts.
#include "pg_device.h"
#include "pg_kernels.h"

#pragma clang fp contract(off)

namespace pg {
namespace {

// src/material_point_method.h:277-311
__device__ __forceinline__ float rc_shape(int order, float x) {
    float a = x < 0.0f ? -x : x;
    if (order == 1) {
        if (a >= 1.0f) return 0.0f;
        return 1.0f - a;
    }
    if (order == 3) {
        if (a >= 2.0f) return 0.0f;
        if (a < 1.0f) return 0.666666686534881591796875f - a * a + 0.5f * a * a * a;
        float t = 2.0f - a;
        return 0.16666667163372039794921875f * t * t * t;
    }
    if (a >= 1.5f) return 0.0f;
    if (a < 0.5f) return 0.75f - x * x;
    return 0.5f * (1.5f - a) * (1.5f - a);
}

__device__ __forceinline__ void rc_base(const float* pos, const RefGeom& g, int* b, float* f) {
#pragma unroll
    for (int a = 0; a < 3; ++a) b[a] = (int)floorf((pos[a] - g.o[a]) / g.h[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = (pos[a] - g.o[a]) / g.h[a] - (float)b[a];
}

__global__ void rc_clear(float* __restrict__ gm, float* __restrict__ gp, float* __restrict__ gv, size_t G) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x) {
        gm[i] = 0.f;
        gp[3 * i] = 0.f; gp[3 * i + 1] = 0.f; gp[3 * i + 2] = 0.f;
        gv[3 * i] = 0.f; gv[3 * i + 1] = 0.f; gv[3 * i + 2] = 0.f;
    }
}

// particleToGrid (:515-570)
__global__ void rc_p2g(const float* __restrict__ x, const float* __restrict__ v, const float* __restrict__ m, int n,
                       RefGeom g, float* __restrict__ gm, float* __restrict__ gp) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float pos[3] = {x[3 * q], x[3 * q + 1], x[3 * q + 2]};
    const float vel[3] = {v[3 * q], v[3 * q + 1], v[3 * q + 2]};
    const float mass = m[q];
    int b[3];
    float f[3];
    rc_base(pos, g, b, f);
    const int ks = (g.order == 1) ? 2 : (g.order == 2) ? 3 : 4;
    const int off = ks / 2;
    for (int di = -off; di < ks - off; ++di)
        for (int dj = -off; dj < ks - off; ++dj)
            for (int dk = -off; dk < ks - off; ++dk) {
                int gi = b[0] + di, gj = b[1] + dj, gk = b[2] + dk;
                if (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || gk < 0 || gk >= g.nz) continue;
                float w = rc_shape(g.order, f[0] - (float)di) * rc_shape(g.order, f[1] - (float)dj);
                w = w * rc_shape(g.order, f[2] - (float)dk);
                if (w > 1.00000001335143196e-10f) {
                    size_t id = ((size_t)gk * g.ny + gj) * g.nx + gi;
                    float wm = w * mass;
                    atomic_add_f32(gm + id, wm);
                    atomic_add_f32(gp + 3 * id + 0, wm * vel[0]);
                    atomic_add_f32(gp + 3 * id + 1, wm * vel[1]);
                    atomic_add_f32(gp + 3 * id + 2, wm * vel[2]);
                }
            }
}

// updateGridMomentum (:572-585) + calculateGridVelocity (:587-597)
__global__ void rc_grid(const float* __restrict__ gm, float* __restrict__ gp, float* __restrict__ gv, size_t G,
                        float dt, float gx, float gy, float gz) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x) {
        float m = gm[i];
        if (m > 1.00000001335143196e-10f) {
            float p0 = gp[3 * i] + m * gx * dt;
            float p1 = gp[3 * i + 1] + m * gy * dt;
            float p2 = gp[3 * i + 2] + m * gz * dt;
            gp[3 * i] = p0; gp[3 * i + 1] = p1; gp[3 * i + 2] = p2;
            gv[3 * i] = p0 / m; gv[3 * i + 1] = p1 / m; gv[3 * i + 2] = p2 / m;
        }
    }
}

// gridToParticle (:599-649) + updateParticlePositions (:651-666)
__global__ void rc_g2p(float* __restrict__ x, float* __restrict__ v, int n, RefGeom g, const float* __restrict__ gv,
                       float dt) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float pos[3] = {x[3 * q], x[3 * q + 1], x[3 * q + 2]};
    int b[3];
    float f[3];
    rc_base(pos, g, b, f);
    const int ks = (g.order == 1) ? 2 : (g.order == 2) ? 3 : 4;
    const int off = ks / 2;
    float nv0 = 0.f, nv1 = 0.f, nv2 = 0.f;
    for (int di = -off; di < ks - off; ++di)
        for (int dj = -off; dj < ks - off; ++dj)
            for (int dk = -off; dk < ks - off; ++dk) {
                int gi = b[0] + di, gj = b[1] + dj, gk = b[2] + dk;
                if (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || gk < 0 || gk >= g.nz) continue;
                float w = rc_shape(g.order, f[0] - (float)di) * rc_shape(g.order, f[1] - (float)dj);
                w = w * rc_shape(g.order, f[2] - (float)dk);
                if (w > 1.00000001335143196e-10f) {
                    size_t id = ((size_t)gk * g.ny + gj) * g.nx + gi;
                    nv0 = nv0 + w * gv[3 * id];
                    nv1 = nv1 + w * gv[3 * id + 1];
                    nv2 = nv2 + w * gv[3 * id + 2];
                }
            }
    v[3 * q] = nv0; v[3 * q + 1] = nv1; v[3 * q + 2] = nv2;
    x[3 * q] = pos[0] + nv0 * dt;
    x[3 * q + 1] = pos[1] + nv1 * dt;
    x[3 * q + 2] = pos[2] + nv2 * dt;
}

}  // namespace

hipError_t refcompat_step(const RefGeom& g, int n, float* x, float* v, const float* m, float* gm, float* gp, float* gv,
                          float dt, const float* grav, hipStream_t st, KernelProf* prof) {
    ProfScope ps(prof, PK_REF_STEP, st);
    const size_t G = (size_t)g.nx * g.ny * g.nz;
    int gb = (int)((G + 255) / 256);
    if (gb > 8192) gb = 8192;
    rc_clear<<<gb, 256, 0, st>>>(gm, gp, gv, G);
    if (n > 0) rc_p2g<<<(n + 255) / 256, 256, 0, st>>>(x, v, m, n, g, gm, gp);
    rc_grid<<<gb, 256, 0, st>>>(gm, gp, gv, G, dt, grav[0], grav[1], grav[2]);
    if (n > 0) rc_g2p<<<(n + 255) / 256, 256, 0, st>>>(x, v, n, g, gv, dt);
    return hipGetLastError();
}

}  // namespace pg
