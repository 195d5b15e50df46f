// MLS-MPM tiled path: cell-binned particles with LDS-staged transfers.
//
// Particles are binned every step by the tile (T^3 block of base cells) of
// their stencil base; a counting sort (per-tile atomic rank, one-workgroup
// scan, scatter) gives perm = sorted position -> storage slot.  One workgroup
// processes one tile at a time:
//   G2P  reads the tile's (T+2)^3 node window of grid velocity into LDS and
//        gathers from it;
//   P2G  accumulates the updated particles into an LDS window with LDS float
//        atomics, then flushes the window with one global atomic per node and
//        channel (instead of one per particle, node and channel).
// G2P of step n and P2G of step n+1 are fused (the particle's new state never
// leaves registers).  The updated state is written at the particle's sorted
// position, so storage stays sorted by tile and the next step's gathers are
// nearly coalesced; pid carries the user's particle index along.  A particle
// whose stencil leaves the tile window (it crossed a tile face this step)
// falls back to global atomics, so correctness never depends on the binning.
//
// The grid update runs over 8^3 node blocks flagged by the P2G flush; it
// writes (v, m) and clears the accumulator, so there is no dense grid clear.
#include "pg_device.h"
#include "pg_kernels.h"

namespace pg {
namespace {

constexpr int TT = 4;            // tile edge (cells)
constexpr int WN = TT + 2;       // window edge (nodes)
constexpr int WV = WN * WN * WN; // window volume (216)
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
        int b = (int)floorf((x[a] - g.o[a]) * g.inv_dx - 0.5f);
        int tc = b >= 0 ? b / TT : 0;
        t[a] = tc < nt[a] ? tc : nt[a] - 1;
    }
    return expand3((uint32_t)t[0]) | (expand3((uint32_t)t[1]) << 1) | (expand3((uint32_t)t[2]) << 2);
}

struct TileArgs {
    const float4* Sin;
    float4* Sout;
    const uint32_t* pid_in;
    uint32_t* pid_out;
    const float* E_in;
    float* E_out;
    const float* m_in;
    float* m_out;
    const uint32_t* perm;
    const uint4* alist;      // (tile, start, count, 0)
    const uint32_t* nactive;
    uint32_t* key_out;
    uint32_t* rank_out;
    uint32_t* count;
    const float4* gv;
    float4* mp;
    uint32_t* flags;
    size_t cap;
    int ntx, nty, ntz;
    int gbx, gby, gbz;
};

__device__ __forceinline__ void flag_node(uint32_t* flags, const TileArgs& a, int gi, int gj, int gk) {
    flags[((gk / GB) * a.gby + gj / GB) * a.gbx + gi / GB] = 1u;
}

// MODE bit 1: G2P (reads x, F; writes the new state at the sorted position)
// MODE bit 2: P2G (scatters the current/updated state into the grid accumulator)
template <int MODE>
__global__ void __launch_bounds__(TB) k_tile(MlsGeom g, TileArgs a) {
    constexpr bool G2P = (MODE & 1) != 0;
    constexpr bool P2G = (MODE & 2) != 0;
    __shared__ float sv[3][WV];
    __shared__ float sm[4][WV];
    const uint32_t nact = *a.nactive;
    // XCD-aware assignment: workgroups are dispatched round-robin over 8 XCDs;
    // give each XCD a contiguous (Morton-ordered, hence compact) chunk of tiles
    // so tiles sharing window nodes share an L2.
    const uint32_t nx8 = 8;
    const uint32_t xcd = blockIdx.x % nx8;
    const uint32_t per_xcd = (gridDim.x + nx8 - 1) / nx8;
    const uint32_t lane = blockIdx.x / nx8;
    const uint32_t c0 = (uint32_t)(((uint64_t)nact * xcd) / nx8);
    const uint32_t c1 = (uint32_t)(((uint64_t)nact * (xcd + 1)) / nx8);
    const float s4 = 4.f * g.inv_dx * g.inv_dx;
    const float kscale = -g.dt * g.p_vol * s4;
    const int tid = threadIdx.x;

    for (uint32_t ti = c0 + lane; ti < c1; ti += per_xcd) {
        const uint4 te = a.alist[ti];
        const uint32_t tile = te.x, beg = te.y, cnt = te.z;
        const int t0[3] = {(int)compact3(tile) * TT, (int)compact3(tile >> 1) * TT, (int)compact3(tile >> 2) * TT};
        for (int l = tid; l < WV; l += TB) {
            if (G2P) {
                int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
                float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
                if (gi < g.nx && gj < g.ny && gk < g.nz) u = a.gv[((size_t)gk * g.ny + gj) * g.nx + gi];
                sv[0][l] = u.x; sv[1][l] = u.y; sv[2][l] = u.z;
            }
            if (P2G) { sm[0][l] = 0.f; sm[1][l] = 0.f; sm[2][l] = 0.f; sm[3][l] = 0.f; }
        }
        __syncthreads();

        for (uint32_t r = tid; r < cnt; r += TB) {
            const uint32_t i = beg + r;
            const uint32_t s = a.perm[i];
            float x[3], F[9], v[3], C[9];
            float Eq = g.E, mq = g.p_mass;
            if (G2P) {
                load_xF(a.Sin, a.cap, s, x, F);
                Stencil st;
                make_stencil(x, g.o, g.inv_dx, st);
                const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
                const bool inwin = l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c] = 0.f;
#pragma unroll
                for (int c = 0; c < 9; ++c) C[c] = 0.f;
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                        for (int kk = 0; kk < 3; ++kk) {
                            float w = st.w[0][ii] * st.w[1][jj] * st.w[2][kk];
                            float d[3] = {((float)ii - st.f[0]) * g.dx, ((float)jj - st.f[1]) * g.dx,
                                          ((float)kk - st.f[2]) * g.dx};
                            float u[3];
                            if (inwin) {
                                int l = ((l2 + kk) * WN + (l1 + jj)) * WN + (l0 + ii);
                                u[0] = sv[0][l]; u[1] = sv[1][l]; u[2] = sv[2][l];
                            } else {
                                int gi = st.b[0] + ii, gj = st.b[1] + jj, gk = st.b[2] + kk;
                                u[0] = u[1] = u[2] = 0.f;
                                if (gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz) {
                                    float4 q = a.gv[((size_t)gk * g.ny + gj) * g.nx + gi];
                                    u[0] = q.x; u[1] = q.y; u[2] = q.z;
                                }
                            }
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                float wu = w * u[c];
                                v[c] += wu;
#pragma unroll
                                for (int e = 0; e < 3; ++e) C[3 * c + e] += wu * d[e];
                            }
                        }
#pragma unroll
                for (int c = 0; c < 9; ++c) C[c] *= s4;
                float Fn[9];
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj)
                        Fn[3 * ii + jj] = F[3 * ii + jj] + g.dt * (C[3 * ii] * F[jj] + C[3 * ii + 1] * F[3 + jj] +
                                                                   C[3 * ii + 2] * F[6 + jj]);
#pragma unroll
                for (int c = 0; c < 9; ++c) F[c] = Fn[c];
#pragma unroll
                for (int c = 0; c < 3; ++c) x[c] += g.dt * v[c];
                store_state(a.Sout, a.cap, i, x, F, v, C);
                a.pid_out[i] = a.pid_in[s];
                if (a.E_in) { Eq = a.E_in[s]; a.E_out[i] = Eq; }
                if (a.m_in) { mq = a.m_in[s]; a.m_out[i] = mq; }
                if (P2G) {
                    uint32_t k = tile_key(x, g, a.ntx, a.nty, a.ntz);
                    a.key_out[i] = k;
                    a.rank_out[i] = atomicAdd(a.count + k, 1u);
                }
            } else {
                load_state(a.Sin, a.cap, s, x, F, v, C);
                if (a.E_in) Eq = a.E_in[s];
                if (a.m_in) mq = a.m_in[s];
            }
            if (P2G) {
                const float mu = Eq / (2.f * (1.f + g.nu));
                const float lam = Eq * g.nu / ((1.f + g.nu) * (1.f - 2.f * g.nu));
                float tau[9], A[9];
                kirchhoff(g.material, mu, lam, F, tau);
#pragma unroll
                for (int c = 0; c < 9; ++c) A[c] = kscale * tau[c] + mq * C[c];
                Stencil st;
                make_stencil(x, g.o, g.inv_dx, st);
                const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
                const bool inwin = l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
                const float mv[3] = {mq * v[0], mq * v[1], mq * v[2]};
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                        for (int kk = 0; kk < 3; ++kk) {
                            float w = st.w[0][ii] * st.w[1][jj] * st.w[2][kk];
                            float d0 = ((float)ii - st.f[0]) * g.dx, d1 = ((float)jj - st.f[1]) * g.dx,
                                  d2 = ((float)kk - st.f[2]) * g.dx;
                            float c0v = w * mq;
                            float c1v = w * (mv[0] + A[0] * d0 + A[1] * d1 + A[2] * d2);
                            float c2v = w * (mv[1] + A[3] * d0 + A[4] * d1 + A[5] * d2);
                            float c3v = w * (mv[2] + A[6] * d0 + A[7] * d1 + A[8] * d2);
                            if (inwin) {
                                int l = ((l2 + kk) * WN + (l1 + jj)) * WN + (l0 + ii);
                                atomicAdd(&sm[0][l], c0v);
                                atomicAdd(&sm[1][l], c1v);
                                atomicAdd(&sm[2][l], c2v);
                                atomicAdd(&sm[3][l], c3v);
                            } else {
                                int gi = st.b[0] + ii, gj = st.b[1] + jj, gk = st.b[2] + kk;
                                if (gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz) {
                                    float* cp = reinterpret_cast<float*>(a.mp + ((size_t)gk * g.ny + gj) * g.nx + gi);
                                    atomic_add_f32(cp + 0, c0v);
                                    atomic_add_f32(cp + 1, c1v);
                                    atomic_add_f32(cp + 2, c2v);
                                    atomic_add_f32(cp + 3, c3v);
                                    flag_node(a.flags, a, gi, gj, gk);
                                }
                            }
                        }
            }
        }
        __syncthreads();
        if (P2G) {
            for (int l = tid; l < WV; l += TB) {
                float m = sm[0][l];
                if (m != 0.f) {
                    int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
                    if (gi < g.nx && gj < g.ny && gk < g.nz) {
                        float* cp = reinterpret_cast<float*>(a.mp + ((size_t)gk * g.ny + gj) * g.nx + gi);
                        atomic_add_f32(cp + 0, m);
                        atomic_add_f32(cp + 1, sm[1][l]);
                        atomic_add_f32(cp + 2, sm[2][l]);
                        atomic_add_f32(cp + 3, sm[3][l]);
                    }
                }
            }
            if (tid < 8) {
                int bx = t0[0] / GB + (tid & 1), by = t0[1] / GB + ((tid >> 1) & 1), bz = t0[2] / GB + (tid >> 2);
                // the window spans nodes [t0, t0 + WN); flag each 8^3 block it touches
                bool touch = bx * GB < t0[0] + WN && by * GB < t0[1] + WN && bz * GB < t0[2] + WN;
                if (touch && bx < a.gbx && by < a.gby && bz < a.gbz) a.flags[(bz * a.gby + by) * a.gbx + bx] = 1u;
            }
        }
        __syncthreads();
    }
}

// Grid update over flagged 8^3 blocks: gv = (p/m + dt g, m) with sticky walls,
// accumulator cleared for the next P2G.
__global__ void __launch_bounds__(512) k_grid_tiled(MlsGeom g, float4* __restrict__ mp, float4* __restrict__ gv,
                                                    uint32_t* __restrict__ flags, int gbx, int gby) {
    const int b = blockIdx.x;
    if (flags[b] == 0u) return;
    const int bx = b % gbx, by = (b / gbx) % gby, bz = b / (gbx * gby);
    const int i = bx * GB + (threadIdx.x & 7), j = by * GB + ((threadIdx.x >> 3) & 7), k = bz * GB + (threadIdx.x >> 6);
    if (i < g.nx && j < g.ny && k < g.nz) {
        size_t id = ((size_t)k * g.ny + j) * g.nx + i;
        float4 a = mp[id];
        bool sticky = i < g.bc || i >= g.nx - g.bc || j < g.bc || j >= g.ny - g.bc || k < g.bc || k >= g.nz - g.bc;
        float4 r = make_float4(0.f, 0.f, 0.f, a.x);
        if (a.x > 0.f && !sticky) {
            float im = 1.f / a.x;
            r.x = a.y * im + g.dt * g.g[0];
            r.y = a.z * im + g.dt * g.g[1];
            r.z = a.w * im + g.dt * g.g[2];
        }
        gv[id] = r;
        mp[id] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    if (threadIdx.x == 0) flags[b] = 0u;
}

__global__ void k_init_keys(int n, size_t cap, const float4* __restrict__ S, MlsGeom g, int ntx, int nty, int ntz,
                            uint32_t* __restrict__ key, uint32_t* __restrict__ rank, uint32_t* __restrict__ count) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float4 q = S[s];
    float x[3] = {q.x, q.y, q.z};
    uint32_t k = tile_key(x, g, ntx, nty, ntz);
    key[s] = k;
    rank[s] = atomicAdd(count + k, 1u);
}

// One workgroup: exclusive scan of per-tile counts, compaction of non-empty
// tiles into the active list, counts cleared for the next binning.
constexpr int SCAN_T = 1024;
__global__ void __launch_bounds__(SCAN_T) k_scan(uint32_t* __restrict__ count, uint32_t* __restrict__ start,
                                                 uint4* __restrict__ alist, uint32_t* __restrict__ nactive,
                                                 int slots) {
    __shared__ uint32_t s_sum[SCAN_T];
    __shared__ uint32_t s_act[SCAN_T];
    const int tid = threadIdx.x;
    const int per = (slots + SCAN_T - 1) / SCAN_T;
    const int lo = tid * per;
    const int hi = min(lo + per, slots);
    uint32_t sum = 0, act = 0;
    for (int c = lo; c < hi; ++c) {
        uint32_t v = count[c];
        sum += v;
        act += v ? 1u : 0u;
    }
    s_sum[tid] = sum;
    s_act[tid] = act;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        uint32_t a = tid >= off ? s_sum[tid - off] : 0u;
        uint32_t b = tid >= off ? s_act[tid - off] : 0u;
        __syncthreads();
        s_sum[tid] += a;
        s_act[tid] += b;
        __syncthreads();
    }
    uint32_t run = s_sum[tid] - sum;
    uint32_t pos = s_act[tid] - act;
    for (int c = lo; c < hi; ++c) {
        uint32_t v = count[c];
        start[c] = run;
        if (v) alist[pos++] = make_uint4((uint32_t)c, run, v, 0u);
        run += v;
        count[c] = 0u;
    }
    if (tid == SCAN_T - 1) *nactive = s_act[tid];
}

__global__ void k_scatter(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ start, uint32_t* __restrict__ perm) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    perm[start[key[s]] + rank[s]] = (uint32_t)s;
}

hipError_t bin(TiledWork& w, int n, hipStream_t st) {
    ProfScope ps(w.prof, PK_BIN, st);
    k_scan<<<1, SCAN_T, 0, st>>>(w.count, w.start, reinterpret_cast<uint4*>(w.active), w.nactive, w.tile_slots);
    if (n > 0) k_scatter<<<(n + 255) / 256, 256, 0, st>>>(n, w.key, w.rank, w.start, w.perm);
    return hipGetLastError();
}

}  // namespace

hipError_t tiled_init_keys(const MlsGeom& g, TiledWork& w, int n, size_t cap, const float4* S, hipStream_t st) {
    if (n > 0)
        k_init_keys<<<(n + 255) / 256, 256, 0, st>>>(n, cap, S, g, w.ntx, w.nty, w.ntz, w.key, w.rank, w.count);
    return bin(w, n, st);
}

hipError_t tiled_steps(const MlsGeom& g, TiledWork& w, int n, size_t cap, float4*& S, float4*& S2, uint32_t*& pid,
                       uint32_t*& pid2, float*& E, float*& E2, float*& mass, float*& mass2, float4* mp, float4* gv,
                       int nsteps, hipStream_t st) {
    if (nsteps <= 0) return hipSuccess;
    const int nblk = w.gbx * w.gby * w.gbz;
    TileArgs a;
    a.perm = w.perm;
    a.alist = reinterpret_cast<const uint4*>(w.active);
    a.nactive = w.nactive;
    a.key_out = w.key;
    a.rank_out = w.rank;
    a.count = w.count;
    a.gv = gv;
    a.mp = mp;
    a.flags = w.flags;
    a.cap = cap;
    a.ntx = w.ntx; a.nty = w.nty; a.ntz = w.ntz;
    a.gbx = w.gbx; a.gby = w.gby; a.gbz = w.gbz;
    auto set_io = [&]() {
        a.Sin = S; a.Sout = S2;
        a.pid_in = pid; a.pid_out = pid2;
        a.E_in = E; a.E_out = E2;
        a.m_in = mass; a.m_out = mass2;
    };
    hipError_t e = tiled_init_keys(g, w, n, cap, S, st);
    if (e != hipSuccess) return e;
    set_io();
    {
        ProfScope ps(w.prof, PK_TILE_P2G, st);
        k_tile<2><<<w.nblocks_launch, TB, 0, st>>>(g, a);
    }
    {
        ProfScope ps(w.prof, PK_GRID, st);
        k_grid_tiled<<<nblk, 512, 0, st>>>(g, mp, gv, w.flags, w.gbx, w.gby);
    }
    for (int s = 0; s + 1 < nsteps; ++s) {
        set_io();
        {
            ProfScope ps(w.prof, PK_TILE_G2P2G, st);
            k_tile<3><<<w.nblocks_launch, TB, 0, st>>>(g, a);
        }
        {
            ProfScope ps(w.prof, PK_GRID, st);
            k_grid_tiled<<<nblk, 512, 0, st>>>(g, mp, gv, w.flags, w.gbx, w.gby);
        }
        e = bin(w, n, st);
        if (e != hipSuccess) return e;
        std::swap(S, S2); std::swap(pid, pid2); std::swap(E, E2); std::swap(mass, mass2);
    }
    set_io();
    {
        ProfScope ps(w.prof, PK_TILE_G2P, st);
        k_tile<1><<<w.nblocks_launch, TB, 0, st>>>(g, a);
    }
    std::swap(S, S2); std::swap(pid, pid2); std::swap(E, E2); std::swap(mass, mass2);
    return hipGetLastError();
}

}  // namespace pg
