// MLS-MPM "plain GPU port": the straightforward structure of the reference's
// transfer kernels (src/mpm_g2p2g_kernels.cu:72-271, src/mpm_kernels.cu:247-545):
// one thread per particle, one global float atomic per node and channel in
// P2G, a dense grid clear and a dense grid update, separate kernels.  It is the
// baseline the tiled path is measured against, and a second independent GPU
// implementation for parity tests.  Also holds the layout helpers.
#include "pg_tile.h"

namespace pg {
namespace {

__global__ void k_fill_f4(float4* __restrict__ p, size_t n, float4 v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_pack(int n, size_t cap, float4* __restrict__ S, const float* __restrict__ x,
                       const float* __restrict__ v, const float* __restrict__ C, const float* __restrict__ F) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float xx[3], vv[3], CC[9], FF[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) { xx[a] = x[3 * i + a]; vv[a] = v ? v[3 * i + a] : 0.f; }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        CC[a] = C ? C[9 * i + a] : 0.f;
        FF[a] = F ? F[9 * i + a] : ((a % 4) == 0 ? 1.f : 0.f);
    }
    store_state(S, cap, i, xx, FF, vv, CC);
}

__global__ void k_unpack(int n, size_t cap, const float4* __restrict__ S, const uint32_t* __restrict__ pid,
                         float* __restrict__ x, float* __restrict__ v, float* __restrict__ C, float* __restrict__ F) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xx[3], vv[3], CC[9], FF[9];
    load_state(S, cap, s, xx, FF, vv, CC);
    size_t i = pid ? pid[s] : (uint32_t)s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (x) x[3 * i + a] = xx[a];
        if (v) v[3 * i + a] = vv[a];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        if (C) C[9 * i + a] = CC[a];
        if (F) F[9 * i + a] = FF[a];
    }
}

__global__ void k_unpermute(int n, const float* __restrict__ src, const uint32_t* __restrict__ pid,
                            float* __restrict__ dst) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[pid ? pid[s] : (uint32_t)s] = src[s];
}

__global__ void k_grid_split(size_t G, const float4* __restrict__ mp, const float4* __restrict__ gv,
                             float* __restrict__ mass, float* __restrict__ mom, float* __restrict__ vel) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x) {
        float4 a = mp[i];
        float4 b = gv[i];
        if (mass) mass[i] = b.w;
        if (mom) { mom[3 * i] = a.y; mom[3 * i + 1] = a.z; mom[3 * i + 2] = a.w; }
        if (vel) { vel[3 * i] = b.x; vel[3 * i + 1] = b.y; vel[3 * i + 2] = b.z; }
    }
}

// P2G: mp[node] = (m, px, py, pz) += w (m, m v + A (x_i - x_p))
__global__ void __launch_bounds__(256) k_p2g_plain(int n, size_t cap, const float4* __restrict__ S,
                                                   const float* __restrict__ mass, const float* __restrict__ E,
                                                   MlsGeom g, MatTable mt, float4* __restrict__ mp) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float x[3], F[9], v[3], C[9];
    load_state(S, cap, q, x, F, v, C);
    const float mq = mass ? mass[q] : g.p_mass;
    const PMat pm = particle_mat(g, mt, (uint32_t)q, E, q, mq);
    float tau[9], A[9];
    stress_of(pm, F, C, tau);
    const float kscale = -g.dt * pm.vol * 4.f * g.inv_dx * g.inv_dx;
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = kscale * tau[i] + mq * C[i];
    Stencil s;
    make_stencil(x, g.o, g.inv_dx, s);
    float* base = reinterpret_cast<float*>(mp);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                int gi = s.b[0] + i, gj = s.b[1] + j, gk = s.b[2] + k;
                if (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || gk < 0 || gk >= g.nz) continue;
                float w = s.w[0][i] * s.w[1][j] * s.w[2][k];
                float d0 = ((float)i - s.f[0]) * g.dx, d1 = ((float)j - s.f[1]) * g.dx, d2 = ((float)k - s.f[2]) * g.dx;
                size_t id = ((size_t)gk * g.ny + gj) * g.nx + gi;
                float* c = base + 4 * id;
                atomic_add_f32(c + 0, w * mq);
                atomic_add_f32(c + 1, w * (mq * v[0] + A[0] * d0 + A[1] * d1 + A[2] * d2));
                atomic_add_f32(c + 2, w * (mq * v[1] + A[3] * d0 + A[4] * d1 + A[5] * d2));
                atomic_add_f32(c + 3, w * (mq * v[2] + A[6] * d0 + A[7] * d1 + A[8] * d2));
            }
}

// grid update: v = p/m + dt g (m > 0), sticky walls; gv = (vx, vy, vz, m)
__global__ void k_grid_plain(MlsGeom g, const float4* __restrict__ mp, float4* __restrict__ gv,
                             const uint32_t* __restrict__ node_bc) {
    const size_t G = (size_t)g.nx * g.ny * g.nz;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < G; id += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(id % g.nx);
        int j = (int)((id / g.nx) % g.ny);
        int k = (int)(id / ((size_t)g.nx * g.ny));
        gv[id] = node_velocity(g, i, j, k, node_bc ? node_bc[id] : 0u, mp[id]);
    }
}

// G2P: v = sum w v_i, C = 4/dx^2 sum w v_i (x_i - x_p)^T, F <- (I + dt C) F, x += dt v
__global__ void __launch_bounds__(256) k_g2p_plain(int n, size_t cap, float4* __restrict__ S, MlsGeom g,
                                                   const float4* __restrict__ gv) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float x[3], F[9];
    load_xF(S, cap, q, x, F);
    Stencil s;
    make_stencil(x, g.o, g.inv_dx, s);
    float nv[3] = {0.f, 0.f, 0.f}, nc[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                int gi = s.b[0] + i, gj = s.b[1] + j, gk = s.b[2] + k;
                if (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || gk < 0 || gk >= g.nz) continue;
                float w = s.w[0][i] * s.w[1][j] * s.w[2][k];
                float d[3] = {((float)i - s.f[0]) * g.dx, ((float)j - s.f[1]) * g.dx, ((float)k - s.f[2]) * g.dx};
                size_t id = ((size_t)gk * g.ny + gj) * g.nx + gi;
                float4 u = gv[id];
                float uv[3] = {u.x, u.y, u.z};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float wu = w * uv[c];
                    nv[c] += wu;
#pragma unroll
                    for (int e = 0; e < 3; ++e) nc[3 * c + e] += wu * d[e];
                }
            }
    const float s4 = 4.f * g.inv_dx * g.inv_dx;
#pragma unroll
    for (int i = 0; i < 9; ++i) nc[i] *= s4;
    float Fn[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Fn[3 * i + j] = F[3 * i + j] + g.dt * (nc[3 * i] * F[j] + nc[3 * i + 1] * F[3 + j] + nc[3 * i + 2] * F[6 + j]);
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] += g.dt * nv[c];
    store_state(S, cap, q, x, Fn, nv, nc);
}

inline int blocks_for(size_t n, int bs, int cap_blocks = 1 << 20) {
    size_t b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    return (int)(b > (size_t)cap_blocks ? cap_blocks : b);
}

}  // namespace

hipError_t fill_f4(float4* p, size_t n, float4 v, hipStream_t st) {
    if (n) k_fill_f4<<<blocks_for(n, 256, 16384), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}
hipError_t fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
    if (n) k_fill_u32<<<blocks_for(n, 256, 16384), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}
hipError_t iota_u32(uint32_t* p, size_t n, hipStream_t st) {
    if (n) k_iota_u32<<<blocks_for(n, 256, 16384), 256, 0, st>>>(p, n);
    return hipGetLastError();
}
hipError_t pack_state(int n, size_t cap, float4* S, const float* x, const float* v, const float* C, const float* F,
                      hipStream_t st) {
    if (n) k_pack<<<blocks_for(n, 256), 256, 0, st>>>(n, cap, S, x, v, C, F);
    return hipGetLastError();
}
hipError_t unpack_state(int n, size_t cap, const float4* S, const uint32_t* pid, float* x, float* v, float* C, float* F,
                        hipStream_t st) {
    if (n) k_unpack<<<blocks_for(n, 256), 256, 0, st>>>(n, cap, S, pid, x, v, C, F);
    return hipGetLastError();
}
hipError_t unpermute_f32(int n, const float* src, const uint32_t* pid, float* dst, hipStream_t st) {
    if (n) k_unpermute<<<blocks_for(n, 256), 256, 0, st>>>(n, src, pid, dst);
    return hipGetLastError();
}
hipError_t grid_split(size_t G, const float4* mp, const float4* gv, float* mass, float* mom, float* vel,
                      hipStream_t st) {
    k_grid_split<<<blocks_for(G, 256, 16384), 256, 0, st>>>(G, mp, gv, mass, mom, vel);
    return hipGetLastError();
}

hipError_t mls_plain_step(const MlsGeom& g, int n, size_t cap, float4* S, const float* mass, const float* E,
                          const MatTable& mt, const uint32_t* node_bc, float4* mp, float4* gv, hipStream_t st,
                          KernelProf* prof) {
    const size_t G = (size_t)g.nx * g.ny * g.nz;
    {
        ProfScope ps(prof, PK_CLEAR, st);
        k_fill_f4<<<blocks_for(G, 256, 16384), 256, 0, st>>>(mp, G, make_float4(0.f, 0.f, 0.f, 0.f));
    }
    if (n > 0) {
        ProfScope ps(prof, PK_PLAIN_P2G, st);
        k_p2g_plain<<<blocks_for(n, 256), 256, 0, st>>>(n, cap, S, mass, E, g, mt, mp);
    }
    {
        ProfScope ps(prof, PK_GRID, st);
        k_grid_plain<<<blocks_for(G, 256, 16384), 256, 0, st>>>(g, mp, gv, node_bc);
    }
    if (n > 0) {
        ProfScope ps(prof, PK_PLAIN_G2P, st);
        k_g2p_plain<<<blocks_for(n, 256), 256, 0, st>>>(n, cap, S, g, gv);
    }
    return hipGetLastError();
}

}  // namespace pg
