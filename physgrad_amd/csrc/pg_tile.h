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

__device__ __forceinline__ void tile_origin(uint32_t tile, int* t0) {
    t0[0] = (int)compact3(tile) * TT;
    t0[1] = (int)compact3(tile >> 1) * TT;
    t0[2] = (int)compact3(tile >> 2) * TT;
}

// Workgroups are dispatched round-robin over the 8 XCDs; each XCD takes a
// contiguous chunk of the (Morton-ordered) tile list so that tiles sharing
// window nodes share an L2.
struct XcdChunk {
    uint32_t first, end, stride;
    __device__ __forceinline__ XcdChunk(uint32_t nact) {
        const uint32_t xcd = blockIdx.x % 8u;
        stride = (gridDim.x + 7u) / 8u;
        const uint32_t lane = blockIdx.x / 8u;
        const uint32_t c0 = (uint32_t)(((uint64_t)nact * xcd) / 8u);
        end = (uint32_t)(((uint64_t)nact * (xcd + 1)) / 8u);
        first = c0 + lane;
    }
};

// Wave-level transpose-reduce scatter into an LDS node window.
//
// LDS float atomics cost ~3 clocks per lane on gfx950, so a per-contribution
// ds_add (108 per particle in P2G) is the bottleneck of a naive LDS-staged
// scatter.  Instead each lane writes the NCH-channel contributions of its
// particle to the 9 nodes of one stencil slice (fixed first offset ii) into a
// per-wave scratch row; then lane (q, c) = (node-in-slice, channel) walks the
// 64 rows in order, summing runs of consecutive particles that share a
// stencil base (particles are sorted by cell inside a tile, so runs are long)
// and issuing one LDS atomic per run.  Rows of particles that do not scatter
// into this window carry base = -1 and break runs without contributing.
//   scr    this wave's scratch, 64 rows of 9*NCH floats, row p = lane p
//   base   the lane's window index of its stencil base node, or -1
//   win    channel-major window, win[c*WV + node]
template <int NCH>
__device__ __forceinline__ void wave_slice_reduce(const float* scr, int base, int ii, float* win) {
    constexpr int NP = 9 * NCH;
    const int lane = threadIdx.x & 63;
    const bool pl = lane < NP;
    const int q = lane / NCH, c = lane - (lane / NCH) * NCH;
    const int off = ii + WN * (q / 3) + WN * WN * (q % 3);  // (ii, jj, kk) -> window offset
    float acc = 0.f;
    int cur = -1;
    for (int p0 = 0; p0 < 64; p0 += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pl ? scr[(p0 + u) * NP + lane] : 0.f;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int cp = __builtin_amdgcn_readlane(base, p0 + u);
            if (cp != cur) {
                if (cur >= 0 && pl) atomicAdd(&win[c * WV + cur + off], acc);
                acc = 0.f;
                cur = cp;
            }
            acc += v[u];
        }
    }
    if (cur >= 0 && pl) atomicAdd(&win[c * WV + cur + off], acc);
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

}  // namespace pg
