// Tile geometry shared by the tiled forward and adjoint kernels.
//
// A tile is a TT^3 block of stencil-base cells; its node window is the
// (TT+2)^3 nodes [t0, t0+TT+2) per axis that the stencils of its particles
// touch.  Tile ids are Morton codes of the tile coordinates, so the tile list
// produced by the counting sort is spatially ordered.
#pragma once
#include "pg_device.h"
#include "pg_kernels.h"

namespace pg {

constexpr int TT = 4;            // tile edge (cells)
constexpr int WN = TT + 2;       // window edge (nodes)
constexpr int WV = WN * WN * WN; // window volume (216)
constexpr int PN = TT + 4;       // P2G window edge (nodes [t0 - 1, t0 + TT + 3)): also takes
                                 // particles that moved one cell out of the tile
constexpr int PV = PN * PN * PN; // P2G window volume (512)
constexpr int TB = 256;          // threads per tile workgroup
constexpr int NW = TB / 64;      // waves per tile workgroup
constexpr int GB = 8;            // grid-update block edge (nodes)

__device__ __forceinline__ uint32_t expand3(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t compact3(uint32_t v) {
    v &= 0x49249249u;
    v = (v ^ (v >> 2)) & 0xC30C30C3u;
    v = (v ^ (v >> 4)) & 0x0F00F00Fu;
    v = (v ^ (v >> 8)) & 0xFF0000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

__device__ __forceinline__ uint32_t tile_key(const float* x, const MlsGeom& g, int ntx, int nty, int ntz) {
    int t[3];
    const int nt[3] = {ntx, nty, ntz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int b = stencil_base(grid_coord(x[a], g.o[a], g.inv_dx));
        int tc = b >= 0 ? b / TT : 0;
        t[a] = tc < nt[a] ? tc : nt[a] - 1;
    }
    return expand3((uint32_t)t[0]) | (expand3((uint32_t)t[1]) << 1) | (expand3((uint32_t)t[2]) << 2);
}

__device__ __forceinline__ uint32_t morton_tile(int tx, int ty, int tz) {
    return expand3((uint32_t)tx) | (expand3((uint32_t)ty) << 1) | (expand3((uint32_t)tz) << 2);
}

__device__ __forceinline__ void tile_origin(uint32_t tile, int* t0) {
    t0[0] = (int)compact3(tile) * TT;
    t0[1] = (int)compact3(tile >> 1) * TT;
    t0[2] = (int)compact3(tile >> 2) * TT;
}

// Work items of the active list, dealt round-robin: workgroup b takes items
// b, b + gridDim, ...  The list holds the longest items first (the binning
// orders it by size class), and the hardware starts workgroups in index order
// as slots free up and spreads consecutive ones over the 8 XCDs, so long
// items start first on every XCD and short ones fill the end of the launch.
struct ItemStride {
    uint32_t first, end, stride;
    __device__ __forceinline__ explicit ItemStride(uint32_t nact) : first(blockIdx.x), end(nact), stride(gridDim.x) {}
};

// number of set bits of a wave mask below this lane
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Keeps the compiler from moving LDS accesses across this point.  The window
// updates of one wave are read-modify-writes whose lanes touch overlapping
// nodes in successive instructions; the LDS executes a wave's instructions in
// issue order, so program order is all they need.
__device__ __forceinline__ void lds_order() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// G2P sums v = sum_n w u_n and C = 4/dx sum_n w u_n (n - f)^T over the 27
// stencil nodes, factored along the axes (x innermost).  load(x, y, z) gives
// the (v, m) of stencil node (x, y, z).
template <class Load>
__device__ __forceinline__ void g2p_sum(Load load, const Stencil& st, float inv_dx, float* v, float* C) {
    float wd[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int n = 0; n < 3; ++n) wd[a][n] = st.w[a][n] * ((float)n - st.f[a]);
    float vv[3] = {0.f, 0.f, 0.f}, cx[3] = {0.f, 0.f, 0.f}, cy[3] = {0.f, 0.f, 0.f}, cz[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int z = 0; z < 3; ++z) {
        float R[3] = {0.f, 0.f, 0.f}, Ry[3] = {0.f, 0.f, 0.f}, Rx[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int y = 0; y < 3; ++y) {
            float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const float4 u = load(x, y, z);
                const float uu[3] = {u.x, u.y, u.z};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    P[c] += st.w[0][x] * uu[c];
                    Q[c] += wd[0][x] * uu[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                R[c] += st.w[1][y] * P[c];
                Ry[c] += wd[1][y] * P[c];
                Rx[c] += st.w[1][y] * Q[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vv[c] += st.w[2][z] * R[c];
            cz[c] += wd[2][z] * R[c];
            cy[c] += st.w[2][z] * Ry[c];
            cx[c] += st.w[2][z] * Rx[c];
        }
        // at most one 9-node slice of loads in flight: bounds the registers
        __builtin_amdgcn_sched_barrier(0);
    }
    const float s = 4.f * inv_dx;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v[c] = vv[c];
        C[3 * c + 0] = s * cx[c];
        C[3 * c + 1] = s * cy[c];
        C[3 * c + 2] = s * cz[c];
    }
}

// scene owning an index, by binary search of a monotone offset field
template <typename F>
__device__ __forceinline__ int scene_of(const SceneInfo* sc, int ns, uint32_t idx, F off) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off(sc[mid]) <= idx) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool in_grid(const MlsGeom& g, int i, int j, int k) {
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}

__device__ __forceinline__ size_t node_id(const MlsGeom& g, int i, int j, int k) {
    return ((size_t)k * g.ny + j) * g.nx + i;
}

__device__ __forceinline__ bool is_sticky(const MlsGeom& g, int i, int j, int k) {
    return i < g.bc || i >= g.nx - g.bc || j < g.bc || j >= g.ny - g.bc || k < g.bc || k >= g.nz - g.bc;
}

// Material of the particle with caller index uid: its table entry when the
// context has a table, else the scene's.  Young's modulus: Ep[ei] when the
// particle array is given, else the material's (or scene's); initial volume:
// mass / density for a table entry with a density, else the scene's V0.
__device__ __forceinline__ PMat particle_mat(const MlsGeom& g, const MatTable& mt, uint32_t uid, const float* Ep,
                                             size_t ei, float mq) {
    PMat m;
    if (mt.recs) {
        const MatRec& r = mt.recs[mt.ids[uid]];
        const float E = Ep ? Ep[ei] : r.E;
        m.model = r.model;
        m.emu = r.emu;
        m.elam = r.elam;
        m.mu = E * r.emu;
        m.lam = E * r.elam;
        m.visc = r.visc;
        m.vol = r.inv_rho > 0.f ? mq * r.inv_rho : g.p_vol;
    } else {
        const float E = Ep ? Ep[ei] : g.E;
        m.model = g.material;
        m.emu = 1.f / (2.f * (1.f + g.nu));
        m.elam = g.nu / ((1.f + g.nu) * (1.f - 2.f * g.nu));
        m.mu = E * m.emu;
        m.lam = E * m.elam;
        m.visc = 0.f;
        m.vol = g.p_vol;
    }
    return m;
}

// Grid velocity of a node from its summed (m, p): v = p/m + dt g masked by
// the walls and the node's boundary flags (zero on an empty node).
__device__ __forceinline__ float4 node_velocity(const MlsGeom& g, int i, int j, int k, uint32_t flags, float4 q) {
    float4 r = make_float4(0.f, 0.f, 0.f, q.x);
    if (q.x > 0.f) {
        const float im = 1.f / q.x;
        const float v[3] = {q.y * im + g.dt * g.g[0], q.z * im + g.dt * g.g[1], q.w * im + g.dt * g.g[2]};
        const int idx[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
        const uint32_t keep = bc_keep(idx, n, g.bc, flags, v);
        r.x = (keep & 1u) ? v[0] : 0.f;
        r.y = (keep & 2u) ? v[1] : 0.f;
        r.z = (keep & 4u) ? v[2] : 0.f;
    }
    return r;
}

// Reverse of node_velocity: the adjoints (am, ap) of the node's (m, p) from
// the adjoint av of its velocity (components the mask dropped pass nothing).
__device__ __forceinline__ float4 node_velocity_vjp(const MlsGeom& g, int i, int j, int k, uint32_t flags, float4 q,
                                                    float4 av) {
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q.x > 0.f) {
        const float im = 1.f / q.x;
        const float v[3] = {q.y * im + g.dt * g.g[0], q.z * im + g.dt * g.g[1], q.w * im + g.dt * g.g[2]};
        const int idx[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
        const uint32_t keep = bc_keep(idx, n, g.bc, flags, v);
        const float a0 = (keep & 1u) ? av.x : 0.f, a1 = (keep & 2u) ? av.y : 0.f, a2 = (keep & 4u) ? av.z : 0.f;
        r.y = a0 * im;
        r.z = a1 * im;
        r.w = a2 * im;
        r.x = -(a0 * q.y + a1 * q.z + a2 * q.w) * im * im;
    }
    return r;
}

}  // namespace pg
