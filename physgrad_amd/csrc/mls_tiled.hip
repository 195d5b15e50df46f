// MLS-MPM tiled path: cell-binned particles with LDS-staged transfers.
//
// Binning.  Particles are binned every step by the tile (TT^3 block of
// stencil-base cells, Morton-numbered) of their stencil base.  The fused
// kernel that moves a particle also counts it into its new tile: per
// workgroup, counts go to 27 LDS counters (one per neighbour direction) and
// one global atomic per (tile, direction) reserves the ranks, so there is no
// per-particle global atomic.  A two-kernel scan turns the tile counts into
// starts and the list of non-empty tiles, and a scatter builds
// perm = sorted position -> storage slot.
//
// Transfers.  One workgroup processes one work item at a time: a chunk of at
// most `chunk` particles of one tile (dense tiles have several).  Per pass of up
// to PASS particles, x and the storage slots are loaded in bulk and ordered by
// stencil-base cell with an LDS counting sort (64 cells) and dealt over
// groups of up to 64 (one per lane) that hold at most two particles of a cell.
//   G2P  gathers the 27 node velocities from the tile's (TT+2)^3 window of
//        grid velocity held in LDS;
//   P2G  each wave owns two LDS windows, one for each of the two particles a
//        cell may have in a group; the lanes of one window touch distinct
//        nodes in each instruction, so they add with plain read-modify-writes
//        (a wave's LDS instructions execute in order).  A particle that moved
//        to another cell adds into a shared, one node wider window with LDS
//        atomics, after the rounds.
// G2P of step n and P2G of step n+1 are fused (the new state never leaves
// registers); the new state is written round by round at consecutive
// positions of the tile's output range, so stores and the next step's loads
// coalesce; pid carries the user's particle index.  A particle whose stencil
// leaves the P2G window (more than one cell per step) scatters with global
// atomics, so correctness never depends on the binning.
//
// Grid.  Each work item accumulates P2G in an (TT+4)^3-node window of its
// tile (one node wider than the stencils of its particles on every side, so a
// particle that moved one cell out of the tile still lands in it) and stores
// the window in a buffer indexed by its place in the active list.  The grid
// update runs over 4^3 node blocks: each node gathers the windows of every
// work item of the 8 tiles that cover it (plus the global accumulator where
// the rare out-of-window P2G landed), computes v = p/m + dt g with sticky walls, and clears what it
// read.  There are no atomics on the regular path and no dense grid clear.
#include <algorithm>

#include "pg_tile.h"

namespace pg {
namespace {

constexpr int PASS = 1280;       // particles staged per LDS pass of a work item (a larger work item takes
                                 // several passes)
constexpr uint16_t FAR = 0xFFFF; // direction code of a particle that jumped more than one tile
constexpr int QCAP = 48;         // P2G values of cell changers kept in LDS per pass (more are reloaded)

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
    const uint4* alist;      // (tile, start, count, scene)
    const uint32_t* nactive;
    uint32_t* key_out;
    uint32_t* rank_out;
    uint32_t* count;
    uint32_t* src_out;       // output position -> input slot (segment tapes only)
    const float4* gv;        // concatenated grids of all scenes
    float4* mp;
    uint32_t* bflag;         // 4^3 node blocks to update (by tile slot): bit 0 window, bit 1 global atomics
    uint32_t* blist;         // flagged blocks, first-flagged order
    uint32_t* nblist;
    float4* wbuf;            // window sums by active-list index
    const SceneInfo* scenes;
    size_t cap;
    uint32_t nslots;         // tile slots (capacity of bflag and blist)
    uint32_t* err;           // binning error word (ERR_* bits)
    MatTable mt;             // per-particle materials (recs NULL: the scene's)
};


// marks 4^3 node block B (tile slot) for the grid update; the first mark of a
// step appends it to the block list
__device__ __forceinline__ void flag_block(uint32_t* bflag, uint32_t* blist, uint32_t* nblist, uint32_t B,
                                           uint32_t bit, uint32_t nslots, uint32_t* err) {
    if (B >= nslots) {   // a block id past the slot space: report, never write
        atomicOr(err, ERR_BLOCK);
        return;
    }
    if (atomicOr(bflag + B, bit) == 0u) {
        const uint32_t i = atomicAdd(nblist, 1u);
        if (i < nslots) blist[i] = B;
        else atomicOr(err, ERR_BLOCK);
    }
}

// tile coordinates of a stencil base, clamped like tile_key
__device__ __forceinline__ void tile_of_base(const int* b, int ntx, int nty, int ntz, int* t) {
    const int nt[3] = {ntx, nty, ntz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int tc = b[a] >= 0 ? b[a] / TT : 0;
        t[a] = tc < nt[a] ? tc : nt[a] - 1;
    }
}

// P2G of one particle: node n = base + (x, y, z) receives w_n (m, q + a0 x + a1 y + a2 z)
// with A = -dt V0 4/dx^2 tau + m C, q = m v - dx A f, a_j = dx A[:, j].
struct P2GVal {
    float m;
    float q[3], a0[3], a1[3], a2[3];
};

__device__ __forceinline__ void p2g_prep(const MlsGeom& g, const PMat& pm, float mq, const float* F, const float* C,
                                         const float* v, const Stencil& st, P2GVal& p) {
    const float kscale = -g.dt * pm.vol * 4.f * g.inv_dx * g.inv_dx;
    float tau[9];
    stress_of(pm, F, C, tau);
    float A[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) A[c] = kscale * tau[c] + mq * C[c];
    p.m = mq;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        p.q[c] = mq * v[c] - g.dx * (A[3 * c] * st.f[0] + A[3 * c + 1] * st.f[1] + A[3 * c + 2] * st.f[2]);
        p.a0[c] = g.dx * A[3 * c];
        p.a1[c] = g.dx * A[3 * c + 1];
        p.a2[c] = g.dx * A[3 * c + 2];
    }
}

// visits the 27 contributions: emit(x, y, z, float4(w m, w p)).  The node
// offsets a0 x + a1 y + a2 z are built by additions along each axis: products
// by the constants 0 and 1 are not folded away in IEEE arithmetic.
template <class Emit>
__device__ __forceinline__ void p2g_visit(const Stencil& st, const P2GVal& p, Emit emit) {
    float bz[3] = {p.q[0], p.q[1], p.q[2]};
#pragma unroll
    for (int z = 0; z < 3; ++z) {
        float by[3] = {bz[0], bz[1], bz[2]};
#pragma unroll
        for (int y = 0; y < 3; ++y) {
            const float wyz = st.w[1][y] * st.w[2][z];
            float bx[3] = {by[0], by[1], by[2]};
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const float w = wyz * st.w[0][x];
                emit(x, y, z, make_float4(w * p.m, w * bx[0], w * bx[1], w * bx[2]));
                if (x < 2)
#pragma unroll
                    for (int c = 0; c < 3; ++c) bx[c] += p.a0[c];
            }
            if (y < 2)
#pragma unroll
                for (int c = 0; c < 3; ++c) by[c] += p.a1[c];
        }
        if (z < 2)
#pragma unroll
            for (int c = 0; c < 3; ++c) bz[c] += p.a2[c];
    }
}

// Rare paths, written as rolled loops so that they stay out of the register
// budget of the rounds.  Stencil values are picked with selects rather than by
// indexing, which would put the stencil in scratch memory.
__device__ __forceinline__ float pick3(const float* w, int i) { return i == 0 ? w[0] : (i == 1 ? w[1] : w[2]); }

// G2P from the global grid (particle whose stencil is not its lane's cell).
__device__ __forceinline__ void g2p_global(const float4* __restrict__ gvs, const MlsGeom& g, const Stencil& st,
                                           float* v, float* C) {
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = 0.f;
#pragma unroll
    for (int c = 0; c < 9; ++c) C[c] = 0.f;
#pragma unroll 1
    for (int n = 0; n < 27; ++n) {
        const int ix = n % 3, iy = (n / 3) % 3, iz = n / 9;
        const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
        if (!in_grid(g, gi, gj, gk)) continue;
        const float4 u = gvs[node_id(g, gi, gj, gk)];
        const float w = pick3(st.w[0], ix) * pick3(st.w[1], iy) * pick3(st.w[2], iz);
        const float d[3] = {(float)ix - st.f[0], (float)iy - st.f[1], (float)iz - st.f[2]};
        const float uu[3] = {u.x, u.y, u.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] += w * uu[c];
#pragma unroll
            for (int e = 0; e < 3; ++e) C[3 * c + e] += w * uu[c] * d[e];
        }
    }
    const float s = 4.f * g.inv_dx;
#pragma unroll
    for (int c = 0; c < 9; ++c) C[c] *= s;
}

__device__ __forceinline__ float4 p2g_contrib(const Stencil& st, const P2GVal& p, int ix, int iy, int iz) {
    const float w = pick3(st.w[0], ix) * pick3(st.w[1], iy) * pick3(st.w[2], iz);
    return make_float4(w * p.m, w * (p.q[0] + p.a0[0] * (float)ix + p.a1[0] * (float)iy + p.a2[0] * (float)iz),
                       w * (p.q[1] + p.a0[1] * (float)ix + p.a1[1] * (float)iy + p.a2[1] * (float)iz),
                       w * (p.q[2] + p.a0[2] * (float)ix + p.a1[2] * (float)iy + p.a2[2] * (float)iz));
}

// P2G into the wide LDS window with atomics (nb = window index of the stencil base).
__device__ __forceinline__ void p2g_window_atomic(float* wf, int nb, const Stencil& st, const P2GVal& p) {
    p2g_visit(st, p, [&](int ix, int iy, int iz, float4 c) {
        float* t = wf + 4 * (nb + (iz * PN + iy) * PN + ix);
        atomicAdd(t + 0, c.x);
        atomicAdd(t + 1, c.y);
        atomicAdd(t + 2, c.z);
        atomicAdd(t + 3, c.w);
    });
}

// P2G into the global grid with atomics; the node blocks hit are flagged for
// the grid update.
__device__ __forceinline__ void p2g_global(float4* __restrict__ mps, uint32_t* bflag, uint32_t* blist,
                                           uint32_t* nblist, uint32_t nslots, uint32_t* err, const SceneInfo& sc,
                                           const MlsGeom& g, const Stencil& st, const P2GVal& p) {
#pragma unroll 1
    for (int n = 0; n < 27; ++n) {
        const int ix = n % 3, iy = (n / 3) % 3, iz = n / 9;
        const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
        if (!in_grid(g, gi, gj, gk)) continue;
        const float4 c = p2g_contrib(st, p, ix, iy, iz);
        float* cp = reinterpret_cast<float*>(mps + node_id(g, gi, gj, gk));
        atomic_add_f32(cp + 0, c.x);
        atomic_add_f32(cp + 1, c.y);
        atomic_add_f32(cp + 2, c.z);
        atomic_add_f32(cp + 3, c.w);
        flag_block(bflag, blist, nblist, sc.toff + morton_tile(gi / TT, gj / TT, gk / TT), 2u, nslots, err);
    }
}

// Tile kernel.  MODE bit 1: G2P (reads x, F; writes the new state);
// MODE bit 2: P2G (scatters the current (MODE 2) or updated (MODE 3) state).
//
// Per pass of up to PASS particles of the work item: the particles' (x, F00)
// and storage slots are loaded in bulk, counted by stencil-base cell (64
// cells, plus "outside" for clamped keys) and staged in LDS in cell order.
// Rounds: the N in-tile particles are dealt, in cell order, over
// G = max(ceil(N / 64), ceil(maxcount / 2)) groups (cell-order index j goes to
// group j mod G, lane j div G), one particle per lane, and waves take groups
// in turn.  Consecutive particles of a cell go to consecutive groups, so a
// group holds at most two particles of a cell: the one of rank < G in the
// cell adds into the wave's first LDS window, the one of rank >= G into the
// second.  Within a window the lanes of a group sit in distinct cells, so
// their 27 node updates never meet in one instruction and add with plain
// read-modify-writes in program order; groups are nearly full whatever the
// spread of the cell counts.  A particle that left its cell during G2P adds
// into a shared, one node wider window after the rounds: the first QCAP of a
// pass leave their P2G values in LDS entries, which one wave adds with
// read-modify-writes (one entry per stencil base at a time, the rare others
// with LDS atomics); the rest are queued, reloaded and added with LDS atomics,
// or, having left even the wide window (more than a cell per step), with
// global atomics.  Outputs are written
// group by group, so stores coalesce.  At the end of the tile the windows are
// summed into the tile's buffer.
// BT: threads per workgroup (256: 4 waves, 3 waves per SIMD; 512: 8 waves,
// 2 waves per SIMD), a launch-shape knob of the autotuner.
template <int MODE, int BT>
__global__ void __launch_bounds__(BT, BT == 256 ? 3 : 2) k_tile(TileArgs a) {
    constexpr int TB = BT;        // threads of this workgroup shape
    constexpr int NW = BT / 64;   // its waves
    constexpr bool G2P = (MODE & 1) != 0;
    constexpr bool P2G = (MODE & 2) != 0;
    constexpr int PT = (PASS + TB - 1) / TB;      // particles per thread in the bulk phases
    __shared__ float4 sv[G2P ? WV : 1];          // G2P window: grid (v, m) on nodes [t0, t0 + WN)
    __shared__ float4 pw[P2G ? NW : 1][P2G ? 2 : 1][WV];  // per-wave P2G windows (m, p) on nodes [t0, t0 + WN):
                                                        // rank in cell < G, >= G
    __shared__ float4 mv[P2G ? PV : 1];          // P2G of moved particles on nodes [t0 - 1, t0 + PN - 1)
    __shared__ uint32_t lslot[PASS];             // storage slot, cell order
    __shared__ uint16_t lrank[PASS];             // rank in its cell, cell order (then the P2G queue)
    __shared__ uint32_t lqueue[(G2P && P2G) ? PASS / 32 : 1];  // bit oo: P2G after the rounds (changed cell)
    __shared__ uint16_t ldir[(G2P && P2G) ? PASS : 1];  // (direction << 11 | rank) of the new tile, or FAR
    __shared__ uint32_t lcnt[65], lstart[65];
    __shared__ uint32_t dcnt[27], dbase[27];
    __shared__ float4 qval[4][(G2P && P2G) ? QCAP : 1];  // changers' (m, q, a0, a1, a2, f)
    __shared__ int qnb[(G2P && P2G) ? QCAP : 1];         // and their wide-window stencil base
    __shared__ uint32_t s_max, s_tot, s_nq, s_qe;
    static_assert(PV <= 4 * PASS, "the changers' cell tags fit in lslot");
    static_assert(PASS / 32 <= 64, "one wave compacts the P2G queue");

    const ItemStride ch(*a.nactive);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (uint32_t ti = ch.first; ti < ch.end; ti += ch.stride) {
        const uint4 te = a.alist[ti];
        const SceneInfo sc = a.scenes[te.w];   // by value: kept in registers across the stores
        const MlsGeom g = sc.g;
        const float4* __restrict__ gvs = a.gv + sc.goff;
        float4* __restrict__ mps = a.mp + sc.goff;
        const uint32_t tile = te.x - sc.toff, beg = te.y, cnt = te.z;
        int t0[3];
        tile_origin(tile, t0);
        const int tc[3] = {t0[0] / TT, t0[1] / TT, t0[2] / TT};
        // the G2P window is loaded here and written to LDS after the first pass's
        // bulk loads have been issued, so that the two latencies overlap
        static_assert(WV <= TB, "one window node per thread");
        float4 wvel = make_float4(0.f, 0.f, 0.f, 0.f);
        if (G2P && tid < WV) {
            const int gi = t0[0] + tid % WN, gj = t0[1] + (tid / WN) % WN, gk = t0[2] + tid / (WN * WN);
            if (gi < g.nx && gj < g.ny && gk < g.nz) wvel = gvs[node_id(g, gi, gj, gk)];
        }
        if (P2G) {
            for (int l = tid; l < WV; l += TB) {
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    pw[w][0][l] = make_float4(0.f, 0.f, 0.f, 0.f);
                    pw[w][1][l] = make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
            for (int l = tid; l < PV; l += TB) mv[l] = make_float4(0.f, 0.f, 0.f, 0.f);
        }

        for (uint32_t pb = 0; pb < cnt; pb += PASS) {
            const uint32_t pc = min((uint32_t)PASS, cnt - pb);
            // ---- bulk load and count by stencil-base cell
            if (tid < 65) lcnt[tid] = 0;
            if (tid < 27) dcnt[tid] = 0;
            if (G2P && P2G) {
                for (int q = tid; q < PASS / 32; q += TB) lqueue[q] = 0u;
                if (tid == 0) s_qe = 0u;
            }
            __syncthreads();
            uint32_t key[PT], ks[PT];
            float4 k0[PT];
#pragma unroll
            for (int k = 0; k < PT; ++k) {
                const uint32_t r = tid + k * TB;
                key[k] = 0u;
                if (r < pc) {
                    const uint32_t s = a.perm[beg + pb + r];
                    ks[k] = s;
                    k0[k] = a.Sin[s];
                }
            }
            if (G2P && pb == 0 && tid < WV) sv[tid] = wvel;
#pragma unroll
            for (int k = 0; k < PT; ++k) {
                const uint32_t r = tid + k * TB;
                if (r < pc) {
                    const float xx[3] = {k0[k].x, k0[k].y, k0[k].z};
                    int l3[3];
                    bool in = true;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        l3[d] = stencil_base(grid_coord(xx[d], g.o[d], g.inv_dx)) - t0[d];
                        in = in && l3[d] >= 0 && l3[d] < TT;
                    }
                    const int cell = in ? l3[0] + TT * l3[2] + TT * TT * l3[1] : 64;
                    const uint32_t rk = atomicAdd(&lcnt[cell], 1u);
                    key[k] = ((uint32_t)cell << 16) | rk;
                }
            }
            __syncthreads();
            if (wave == 0) {
                // cell starts (wave scan) and the largest cell count
                const uint32_t cl = lcnt[lane];
                uint32_t inc = cl;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t t = __shfl_up(inc, off);
                    if (lane >= off) inc += t;
                }
                lstart[lane] = inc - cl;
                uint32_t mx = cl;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
                if (lane == 63) {
                    lstart[64] = inc;
                    s_tot = inc;
                    s_max = mx;
                }
            }
            __syncthreads();
            // ---- stage in cell order
#pragma unroll
            for (int k = 0; k < PT; ++k) {
                const uint32_t r = tid + k * TB;
                if (r < pc) {
                    const uint32_t j = lstart[key[k] >> 16] + (key[k] & 0xFFFFu);
                    lslot[j] = ks[k];
                    lrank[j] = (uint16_t)(key[k] & 0xFFFFu);
                }
            }
            __syncthreads();

            // binning of a moved particle: key, and its rank in the new tile through the
            // per-direction LDS counters (FAR: a global atomic on the tile count)
            auto bin_out = [&](const Stencil& stn, uint32_t o, uint32_t oo, bool aggregate) {
                int nt[3];
                tile_of_base(stn.b, sc.ntx, sc.nty, sc.ntz, nt);
                const int dx_ = nt[0] - tc[0], dy_ = nt[1] - tc[1], dz_ = nt[2] - tc[2];
                const uint32_t k = sc.toff + morton_tile(nt[0], nt[1], nt[2]);
                a.key_out[o] = k;
                const bool near = dx_ >= -1 && dx_ <= 1 && dy_ >= -1 && dy_ <= 1 && dz_ >= -1 && dz_ <= 1;
                const int dir = near ? (dx_ + 1) + 3 * (dy_ + 1) + 9 * (dz_ + 1) : -1;
                uint32_t rk = 0;
                bool done = false;
                if (aggregate) {
                    // most particles stay in the tile: one LDS atomic per wave for those
                    const uint64_t stay = __ballot(dir == 13);
                    if (stay) {
                        const int leader = __ffsll((unsigned long long)stay) - 1;
                        uint32_t b0 = 0;
                        if (lane == leader) b0 = atomicAdd(&dcnt[13], (uint32_t)__popcll(stay));
                        b0 = __builtin_amdgcn_readlane(b0, leader);
                        rk = b0 + lanes_below(stay);
                        done = dir == 13;
                    }
                }
                if (dir >= 0 && !done) rk = atomicAdd(&dcnt[dir], 1u);
                if (dir >= 0) {
                    ldir[oo] = (uint16_t)((dir << 11) | rk);
                } else {
                    ldir[oo] = FAR;
                    a.rank_out[o] = atomicAdd(a.count + k, 1u);
                }
            };

            // P2G of a particle with its (new) state: into the wave's window by
            // read-modify-write when its base is this lane's cell (own).  Otherwise
            // the fused kernel queues it for a pass after the rounds (so that the
            // rounds never branch off into its 27 node updates): its P2G values,
            // computed alongside the own lanes', go to an LDS entry while there is
            // one, else the particle is reloaded after the rounds.  The P2G-only
            // kernel adds it into the shared wide window with LDS atomics when it
            // is nearby, else with global atomics.
            // ob = window index of the particle's base before its G2P (own_ok only);
            // win = which of the wave's two windows
            auto scatter = [&](const Stencil& st, const float* F, const float* C, const float* v, const PMat& pm, float mq,
                               bool own_ok, uint32_t oo, int ob, int win) {
                const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
                const int nb = (l2 * WN + l1) * WN + l0;
                const bool inside = l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
                const bool own = own_ok && inside && nb == ob;
                const bool near = l0 >= -1 && l0 <= TT && l1 >= -1 && l1 <= TT && l2 >= -1 && l2 <= TT;
                P2GVal p;
                p2g_prep(g, pm, mq, F, C, v, st, p);
                if (G2P && !own) {
                    const uint32_t e = near ? atomicAdd(&s_qe, 1u) : (uint32_t)QCAP;
                    if (e < QCAP) {
                        qval[0][e] = make_float4(p.m, p.q[0], p.q[1], p.q[2]);
                        qval[1][e] = make_float4(p.a0[0], p.a0[1], p.a0[2], p.a1[0]);
                        qval[2][e] = make_float4(p.a1[1], p.a1[2], p.a2[0], p.a2[1]);
                        qval[3][e] = make_float4(p.a2[2], st.f[0], st.f[1], st.f[2]);
                        qnb[e] = ((l2 + 1) * PN + l1 + 1) * PN + l0 + 1;
                    } else {
                        atomicOr(&lqueue[oo >> 5], 1u << (oo & 31));
                    }
                    return;
                }
                if (G2P || own) {
                    float4* __restrict__ pwin = pw[wave][win];
                    p2g_visit(st, p, [&](int ix, int iy, int iz, float4 c) {
                        float4* t = pwin + nb + (iz * WN + iy) * WN + ix;
                        float4 u = *t;
                        u.x += c.x;
                        u.y += c.y;
                        u.z += c.z;
                        u.w += c.w;
                        *t = u;
                        lds_order();
                    });
                } else if (near) {
                    p2g_window_atomic(reinterpret_cast<float*>(mv), ((l2 + 1) * PN + l1 + 1) * PN + l0 + 1, st, p);
                } else {
                    p2g_global(mps, a.bflag, a.blist, a.nblist, a.nslots, a.err, sc, g, st, p);
                }
                lds_order();
            };

            // one particle: j = its cell-order index, oo = its output offset, own_ok = it
            // is in a round (its base is a cell of the tile: G2P from the window),
            // win = which of the wave's two P2G windows it adds into; s = its storage
            // slot and q0 = its (x, F00), loaded by the caller
            auto particle_q = [&](uint32_t j, uint32_t oo, bool own_ok, int win, uint32_t s, float4 q0) {
                const float4 q1 = a.Sin[a.cap + s], q2 = a.Sin[2 * a.cap + s];
                float x[3], F[9], v[3], C[9];
                const uint32_t o = beg + pb + oo;
                x[0] = q0.x; x[1] = q0.y; x[2] = q0.z;
                F[0] = q0.w;
                int ob = -1;
                float mq = g.p_mass;
                if (a.m_in) mq = a.m_in[s];
                const uint32_t uid = a.pid_in[s];
                Stencil stn;   // stencil of the position P2G scatters from
                if (G2P) {
                    Stencil st;
                    make_stencil(x, g.o, g.inv_dx, st);
                    ob = ((st.b[2] - t0[2]) * WN + (st.b[1] - t0[1])) * WN + (st.b[0] - t0[0]);
                    if (own_ok)
                        g2p_sum([&](int ix, int iy, int iz) { return sv[ob + (iz * WN + iy) * WN + ix]; }, st,
                                g.inv_dx, v, C);
                    else
                        g2p_global(gvs, g, st, v, C);
                    F[1] = q1.x; F[2] = q1.y; F[3] = q1.z; F[4] = q1.w;
                    F[5] = q2.x; F[6] = q2.y; F[7] = q2.z; F[8] = q2.w;
                    float Fn[9];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int jj = 0; jj < 3; ++jj)
                            Fn[3 * i + jj] = F[3 * i + jj] + g.dt * (C[3 * i] * F[jj] + C[3 * i + 1] * F[3 + jj] +
                                                                     C[3 * i + 2] * F[6 + jj]);
#pragma unroll
                    for (int c = 0; c < 9; ++c) F[c] = Fn[c];
#pragma unroll
                    for (int c = 0; c < 3; ++c) x[c] += g.dt * v[c];
                    store_state(a.Sout, a.cap, o, x, F, v, C);
                    a.pid_out[o] = uid;
                    if (a.E_out && a.E_in) a.E_out[o] = a.E_in[s];
                    if (a.m_out && a.m_in) a.m_out[o] = mq;
                    if (a.src_out) a.src_out[o] = s;
                    if (P2G) {
                        make_stencil(x, g.o, g.inv_dx, stn);
                        bin_out(stn, o, oo, own_ok);
                    }
                } else {
                    F[1] = q1.x; F[2] = q1.y; F[3] = q1.z; F[4] = q1.w;
                    F[5] = q2.x; F[6] = q2.y; F[7] = q2.z; F[8] = q2.w;
                    const float4 q3 = a.Sin[3 * a.cap + s], q4 = a.Sin[4 * a.cap + s], q5 = a.Sin[5 * a.cap + s];
                    v[0] = q3.x; v[1] = q3.y; v[2] = q3.z; C[0] = q3.w;
                    C[1] = q4.x; C[2] = q4.y; C[3] = q4.z; C[4] = q4.w;
                    C[5] = q5.x; C[6] = q5.y; C[7] = q5.z; C[8] = q5.w;
                    // no G2P: the base is still the staged cell
                    make_stencil(x, g.o, g.inv_dx, stn);
                    ob = ((stn.b[2] - t0[2]) * WN + (stn.b[1] - t0[1])) * WN + (stn.b[0] - t0[0]);
                }
                if (P2G) scatter(stn, F, C, v, particle_mat(g, a.mt, uid, a.E_in, s, mq), mq, own_ok, oo, ob, win);
            };
            // the same with its slot and (x, F00) loaded here
            auto particle = [&](uint32_t j, uint32_t oo, bool own_ok, int win) {
                const uint32_t s = lslot[j];
                particle_q(j, oo, own_ok, win, s, a.Sin[s]);
            };

            // ---- rounds: cell-order index j = grp + lane G of group grp (grp < G); the
            // groups' output ranges follow each other (group grp holds ceil((tot - grp) / G)
            // particles and starts at grp (tot div G) + min(grp, tot mod G))
            const uint32_t maxcnt = s_max, tot = s_tot;
            const uint32_t G = max((tot + 63u) / 64u, (maxcnt + 1u) / 2u);
            if (G > 0) {
                const uint32_t gq = tot / G, gr = tot % G;
                // a lane loads the slot and (x, F00) of its particle in the wave's next
                // group while it works on the current one (the G2P addresses need x)
                uint32_t jn = wave + lane * G;
                bool vn = wave < G && jn < tot;
                uint32_t sn = vn ? lslot[jn] : 0u;
                float4 q0n = make_float4(0.f, 0.f, 0.f, 0.f);
                if (vn) q0n = a.Sin[sn];
                for (uint32_t grp = wave; grp < G; grp += NW) {
                    const uint32_t j = jn;
                    const bool val = vn;
                    const uint32_t s0 = sn;
                    const float4 q0 = q0n;
                    jn = grp + NW + lane * G;
                    vn = grp + NW < G && jn < tot;
                    if (vn) { sn = lslot[jn]; q0n = a.Sin[sn]; }
                    if (val) particle_q(j, grp * gq + min(grp, gr) + lane, true, lrank[j] >= G ? 1 : 0, s0, q0);
                }
            }
            // ---- particles whose stencil base is outside the tile (clamped keys)
            const uint32_t nout = lcnt[64], sout = lstart[64];
            for (uint32_t t = tid; t < nout; t += TB) particle(sout + t, tot + t, false, 0);
            __syncthreads();
            // ---- after the rounds: the last wave adds the changers' LDS entries into
            // the shared wide window while wave 0 compacts the queue of the rest (into
            // lrank, free after the rounds) and takes the ranks in the new tiles (one
            // global atomic per (tile, direction)).  Of 64 entries with the same
            // stencil base one, picked by a tag per base (in lslot, free after the
            // rounds), goes in by plain read-modify-writes in program order as in the
            // rounds (the picked lanes' bases differ, so one instruction never touches
            // a node twice); the others, rare, with LDS atomics.  Then the queued
            // particles are reloaded, up to 64 a wave, and added with LDS atomics
            // (global atomics for one that left the wide window), and the new-tile
            // ranks are written.
            if (P2G && G2P) {
                if (wave == 0) {
                    const uint32_t bits = lane < (pc + 31) / 32 ? lqueue[lane] : 0u;
                    const uint32_t nb = (uint32_t)__popc(bits);
                    uint32_t inc = nb;
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const uint32_t t = __shfl_up(inc, off);
                        if (lane >= off) inc += t;
                    }
                    uint32_t q = inc - nb;
                    for (uint32_t r = bits; r; r &= r - 1u) lrank[q++] = (uint16_t)(lane * 32 + __ffs(r) - 1);
                    if (lane == 63) s_nq = inc;
                    if (lane < 27) {
                        const uint32_t nd = dcnt[lane];
                        if (nd) {
                            const int dx_ = lane % 3 - 1, dy_ = (lane / 3) % 3 - 1, dz_ = lane / 9 - 1;
                            const uint32_t k = sc.toff + morton_tile(tc[0] + dx_, tc[1] + dy_, tc[2] + dz_);
                            dbase[lane] = atomicAdd(a.count + k, nd);
                        }
                    }
                } else if (wave == NW - 1) {
                    uint8_t* __restrict__ qtag = reinterpret_cast<uint8_t*>(lslot);
                    const uint32_t ne = min(s_qe, (uint32_t)QCAP);
                    for (uint32_t e0 = 0; e0 < ne; e0 += 64) {
                        const uint32_t e = min(e0 + lane, ne - 1u);
                        bool pend = e0 + lane < ne;
                        const float4 u0 = qval[0][e], u1 = qval[1][e], u2 = qval[2][e], u3 = qval[3][e];
                        const P2GVal p = {u0.x, {u0.y, u0.z, u0.w}, {u1.x, u1.y, u1.z}, {u1.w, u2.x, u2.y},
                                          {u2.z, u2.w, u3.x}};
                        Stencil st;
                        stencil_weights(u3.y, st.w[0]);
                        stencil_weights(u3.z, st.w[1]);
                        stencil_weights(u3.w, st.w[2]);
                        const int nb = qnb[e];
                        if (pend) qtag[nb] = (uint8_t)lane;
                        lds_order();
                        const bool turn = pend && qtag[nb] == (uint8_t)lane;
                        if (turn) {
                            p2g_visit(st, p, [&](int ix, int iy, int iz, float4 c) {
                                float4* t = mv + nb + (iz * PN + iy) * PN + ix;
                                float4 u = *t;
                                u.x += c.x;
                                u.y += c.y;
                                u.z += c.z;
                                u.w += c.w;
                                *t = u;
                                lds_order();
                            });
                        }
                        lds_order();
                        if (pend && !turn) p2g_window_atomic(reinterpret_cast<float*>(mv), nb, st, p);
                    }
                }
                __syncthreads();
                const uint32_t nq = s_nq;
                for (uint32_t q = tid; q < nq; q += TB) {
                    const uint32_t o = beg + pb + lrank[q];
                    float x[3], F[9], v[3], C[9];
                    load_state(a.Sout, a.cap, o, x, F, v, C);
                    float mq = g.p_mass;
                    if (a.m_out && a.m_in) mq = a.m_out[o];
                    const PMat pm = particle_mat(g, a.mt, a.pid_out[o], (a.E_out && a.E_in) ? a.E_out : nullptr, o,
                                                 mq);
                    Stencil st;
                    make_stencil(x, g.o, g.inv_dx, st);
                    const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
                    P2GVal p;
                    p2g_prep(g, pm, mq, F, C, v, st, p);
                    if (l0 >= -1 && l0 <= TT && l1 >= -1 && l1 <= TT && l2 >= -1 && l2 <= TT)
                        p2g_window_atomic(reinterpret_cast<float*>(mv), ((l2 + 1) * PN + l1 + 1) * PN + l0 + 1, st,
                                          p);
                    else
                        p2g_global(mps, a.bflag, a.blist, a.nblist, a.nslots, a.err, sc, g, st, p);
                }
                for (uint32_t jj = tid; jj < pc; jj += TB) {
                    const uint16_t dr = ldir[jj];
                    if (dr != FAR) a.rank_out[beg + pb + jj] = dbase[dr >> 11] + (dr & 0x7FFu);
                }
            }
            __syncthreads();
        }

        if (P2G) {
            // the tile's window sum goes to its buffer (one per active-list entry);
            // node l of the wide window is node (l - (1,1,1)) of the per-wave windows
            float4* __restrict__ wb = a.wbuf + (size_t)ti * PV;
            for (int l = tid; l < PV; l += TB) {
                const int px = l % PN, py = (l / PN) % PN, pz = l / (PN * PN);
                float4 m = mv[l];
                if (px >= 1 && px <= WN && py >= 1 && py <= WN && pz >= 1 && pz <= WN) {
                    const int l6 = ((pz - 1) * WN + py - 1) * WN + px - 1;
#pragma unroll
                    for (int w = 0; w < NW; ++w)
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const float4 u = pw[w][h][l6];
                            m.x += u.x;
                            m.y += u.y;
                            m.z += u.z;
                            m.w += u.w;
                        }
                }
                wb[l] = m;
            }
            if (tid < 27) {
                // node blocks this window covers: the tile's and its 26 neighbours
                const int bx = tc[0] + tid % 3 - 1, by = tc[1] + (tid / 3) % 3 - 1, bz = tc[2] + tid / 9 - 1;
                if (bx >= 0 && by >= 0 && bz >= 0 && bx < sc.ntx && by < sc.nty && bz < sc.ntz)
                    flag_block(a.bflag, a.blist, a.nblist, sc.toff + morton_tile(bx, by, bz), 1u, a.nslots, a.err);
            }
        }
        __syncthreads();
    }
}

// the tile kernel in the workgroup shape the context chose
template <int MODE>
void launch_tile(const TiledWork& w, const TileArgs& a, hipStream_t st) {
    if (w.block == 512) k_tile<MODE, 512><<<w.nblocks_launch, 512, 0, st>>>(a);
    else k_tile<MODE, 256><<<w.nblocks_launch, 256, 0, st>>>(a);
}

// Grid update of one flagged 4^3 node block (one wave, lane = node).  The
// P2G window of tile T covers nodes [4T - 1, 4T + 7) per axis, so node block
// B is covered by tiles B + {-1, 0, 1}^3, each node by 8 of them; the node's
// (m, p) is the sum of the buffered windows of every work item of those tiles
// plus the global accumulator where the rare global-atomic P2G landed.
// Writes gv = (p/m + dt g, m) with sticky walls and clears the accumulator
// and the block's flags.  TAPE (segment recompute for the adjoint): also
// stores the summed (m, p) in the step's grid and flags its 8^3 block, both
// of which the reverse pass uses and clears.
template <bool TAPE>
__device__ __forceinline__ void grid_block(uint32_t B, const SceneInfo* __restrict__ scenes,
                                           const uint32_t* __restrict__ tscene,
                                           const uint32_t* __restrict__ tslot, uint32_t* __restrict__ bflag,
                                           const float4* __restrict__ wbuf, float4* __restrict__ mp_all,
                                           float4* __restrict__ gv_all, uint32_t* __restrict__ flags8,
                                           const uint32_t* __restrict__ node_bc) {
    // covered by a buffered window (bit 0), global-atomic P2G here (bit 1)
    const uint32_t fl = bflag[B];
    const bool fb = (fl & 2u) != 0u, wb = (fl & 1u) != 0u;
    const int lane = threadIdx.x & 63;
    const SceneInfo sc = scenes[tscene[B]];
    const uint32_t lb = B - sc.toff;
    const int bc[3] = {(int)compact3(lb), (int)compact3(lb >> 1), (int)compact3(lb >> 2)};
    const int nt[3] = {sc.ntx, sc.nty, sc.ntz};
    // first active-list entry of a tile and its number of entries (chunks): (first, 0) outside the grid
    auto slot_of = [&](int tx, int ty, int tz) -> uint2 {
        if (tx < 0 || ty < 0 || tz < 0 || tx >= nt[0] || ty >= nt[1] || tz >= nt[2]) return make_uint2(0u, 0u);
        return reinterpret_cast<const uint2*>(tslot)[sc.toff + morton_tile(tx, ty, tz)];
    };
    const MlsGeom g = sc.g;
    const int lc[3] = {lane & 3, (lane >> 2) & 3, lane >> 4};
    const int i = bc[0] * TT + lc[0], j = bc[1] * TT + lc[1], k = bc[2] * TT + lc[2];
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (wb) {
        // the entry ranges of the 8 covering tiles, then their first windows, as
        // independent loads (two round trips); the further windows of a tile split
        // into several work items after that
        uint2 f[8];
        int wo[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int dT[3], w[3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const int bit = (e >> ax) & 1;
                dT[ax] = lc[ax] < 3 ? bit - 1 : bit;
                w[ax] = lc[ax] + 1 - TT * dT[ax];
            }
            f[e] = slot_of(bc[0] + dT[0], bc[1] + dT[1], bc[2] + dT[2]);
            wo[e] = (w[2] * PN + w[1]) * PN + w[0];
        }
        float4 u[8];
#pragma unroll
        for (int e = 0; e < 8; ++e)
            u[e] = f[e].y ? wbuf[(size_t)f[e].x * PV + wo[e]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            q.x += u[e].x;
            q.y += u[e].y;
            q.z += u[e].z;
            q.w += u[e].w;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e)
            for (uint32_t c = 1; c < f[e].y; ++c) {
                const float4 t = wbuf[(size_t)(f[e].x + c) * PV + wo[e]];
                q.x += t.x;
                q.y += t.y;
                q.z += t.z;
                q.w += t.w;
            }
    }
    float4* __restrict__ mp = mp_all + sc.goff;
    float4* __restrict__ gv = gv_all + sc.goff;
    if (i < g.nx && j < g.ny && k < g.nz && bc[0] < nt[0] && bc[1] < nt[1] && bc[2] < nt[2]) {
        const size_t id = node_id(g, i, j, k);
        if (fb) {
            const float4 u = mp[id];
            q.x += u.x;
            q.y += u.y;
            q.z += u.z;
            q.w += u.w;
            if (!TAPE) mp[id] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        gv[id] = node_velocity(g, i, j, k, node_bc ? node_bc[sc.goff + id] : 0u, q);
        if (TAPE) mp[id] = q;
    }
    if (lane == 0) {
        bflag[B] = 0u;
        if (TAPE && bc[0] < nt[0] && bc[1] < nt[1] && bc[2] < nt[2])
            flags8[sc.boff + ((bc[2] * TT / GB) * sc.gby + bc[1] * TT / GB) * sc.gbx + bc[0] * TT / GB] = 1u;
    }
}

// The flagged blocks of the list, one per wave, waves striding over the list
// (nwg workgroups of 4 waves).  The list is emptied by the next binning
// (k_scan_b) or, after a call's first grid update, by k_list_reset.
template <bool TAPE>
__device__ __forceinline__ void grid_list(uint32_t nwg, const SceneInfo* __restrict__ scenes,
                                          const uint32_t* __restrict__ tscene, const uint32_t* __restrict__ tslot,
                                          uint32_t* __restrict__ bflag, const uint32_t* __restrict__ blist,
                                          const uint32_t* __restrict__ nblist, const float4* __restrict__ wbuf,
                                          float4* __restrict__ mp_all, float4* __restrict__ gv_all,
                                          uint32_t* __restrict__ flags8, const uint32_t* __restrict__ node_bc) {
    const uint32_t nl = *nblist;
    for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < nl; i += nwg * 4)
        grid_block<TAPE>(blist[i], scenes, tscene, tslot, bflag, wbuf, mp_all, gv_all, flags8, node_bc);
}

__global__ void k_list_reset(uint32_t* __restrict__ nblist) { *nblist = 0u; }

template <bool TAPE>
__global__ void __launch_bounds__(256) k_grid_nodes(const SceneInfo* __restrict__ scenes,
                                                    const uint32_t* __restrict__ tscene,
                                                    const uint32_t* __restrict__ tslot, uint32_t* __restrict__ bflag,
                                                    const uint32_t* __restrict__ blist,
                                                    const uint32_t* __restrict__ nblist,
                                                    const float4* __restrict__ wbuf,
                                                    float4* __restrict__ mp_all, float4* __restrict__ gv_all,
                                                    uint32_t* __restrict__ flags8,
                                                    const uint32_t* __restrict__ node_bc) {
    grid_list<TAPE>(gridDim.x, scenes, tscene, tslot, bflag, blist, nblist, wbuf, mp_all, gv_all, flags8, node_bc);
}

// First binning of a call: tile key per storage slot, ranks from one global
// atomic per distinct key in each wave (storage is normally tile-sorted
// already, so a wave holds one or two keys).
__global__ void k_init_keys(int n, const float4* __restrict__ S, const uint32_t* __restrict__ pid,
                            const SceneInfo* __restrict__ scenes, int nscenes, uint32_t* __restrict__ key,
                            uint32_t* __restrict__ rank, uint32_t* __restrict__ count) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < n;
    uint32_t k = 0xFFFFFFFFu;
    if (valid) {
        const float4 q = S[s];
        const float x[3] = {q.x, q.y, q.z};
        const SceneInfo& sc = scenes[scene_of(scenes, nscenes, pid[s], [](const SceneInfo& t) { return t.poff; })];
        k = sc.toff + tile_key(x, sc.g, sc.ntx, sc.nty, sc.ntz);
    }
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint64_t todo = __ballot(valid);
    uint32_t rk = 0;
    // up to four distinct keys per wave with one atomic each (tile-sorted
    // storage); the lanes left after that (unsorted storage, e.g. the caller's
    // order) take their ranks with one atomic each, all in flight at once
    for (int it = 0; it < 4 && todo; ++it) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t kl = __builtin_amdgcn_readlane(k, leader);
        const uint64_t peers = __ballot(k == kl) & todo;
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(count + kl, (uint32_t)__popcll(peers));
        b = __builtin_amdgcn_readlane(b, leader);
        if (k == kl && valid) rk = b + (uint32_t)__popcll(peers & lt);
        todo &= ~peers;
    }
    if ((todo >> lane) & 1ull) rk = atomicAdd(count + k, 1u);
    if (valid) {
        key[s] = k;
        rank[s] = rk;
    }
}

// Scan of the tile counts in two kernels over blocks of SCAN_B slots:
// block sums, then per-block exclusive scans with the preceding blocks'
// totals; writes starts, the active list (one entry per chunk of at most
// `chunk` particles of a tile), the tile -> entries map, and clears the counts.
//
// Order of the active list: longest work first.  A tile's entries stay
// together; tiles are grouped by the size class of their first entry
// (quarters of `chunk`), the largest class first, tile order within a class.
// Dealt round-robin to workgroups that start in index order (ItemStride), the
// long items start first and the short ones fill the end of the launch.
constexpr int SCAN_T = 256;
constexpr int SCAN_B = 1024;
constexpr int NCLS = 4;            // size classes of work items
constexpr int BSUM_W = 1 + NCLS;   // words per scan block in bsum: particles, entries of each class

// size class of a tile with v > 0 particles: its first entry holds min(v, chunk)
__device__ __forceinline__ uint32_t item_class(uint32_t v, uint32_t chunk) {
    return min((4u * min(v, chunk) - 1u) / chunk, (uint32_t)NCLS - 1u);
}

// The entries per class of a thread's tiles, packed two classes to a 64-bit
// word (class 3 | class 2, class 1 | class 0) so that one 64-bit scan sums both
// halves; no half overflows (fewer than 2^32 entries).
__device__ __forceinline__ void add_class(uint64_t* e, uint32_t cls, uint32_t ne) {
    const uint64_t d = (uint64_t)ne << (32 * (cls & 1u));
    if (cls >= 2u) e[0] += d;
    else e[1] += d;
}
__device__ __forceinline__ uint32_t class_part(const uint64_t* e, uint32_t cls) {
    return (uint32_t)(e[cls >= 2u ? 0 : 1] >> (32 * (cls & 1u)));
}

// Exclusive scans over the workgroup (SCAN_T threads) of K sums at once: wave
// scans by shuffles, then the wave totals through LDS (sh: K * SCAN_T/64).
// v[k] becomes the exclusive prefix, tot[k] the workgroup total.
template <int K>
__device__ __forceinline__ void block_scan(uint64_t (&v)[K], uint64_t (&tot)[K], uint64_t* sh) {
    constexpr int NWV = SCAN_T / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        inc[k] = v[k];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t t = __shfl_up(inc[k], off);
            if (lane >= off) inc[k] += t;
        }
        if (lane == 63) sh[k * NWV + wave] = inc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint64_t pre = 0, all = 0;
#pragma unroll
        for (int w = 0; w < NWV; ++w) {
            const uint64_t t = sh[k * NWV + w];
            pre += w < wave ? t : 0;
            all += t;
        }
        v[k] = inc[k] - v[k] + pre;
        tot[k] = all;
    }
    __syncthreads();
}

// particles and per-class entries of this thread's 4 tile slots of block blk
__device__ __forceinline__ void thread_tiles(int blk, const uint32_t* __restrict__ count, int slots, uint32_t chunk,
                                             uint32_t* v, uint64_t* e, uint64_t& s) {
    const int base = blk * SCAN_B + threadIdx.x * 4;
    s = 0;
    e[0] = e[1] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < slots ? count[base + k] : 0u;
        s += v[k];
        if (v[k]) add_class(e, item_class(v[k], chunk), (v[k] + chunk - 1) / chunk);
    }
}

__device__ __forceinline__ void scan_a_block(int blk, const uint32_t* __restrict__ count, int slots, uint32_t chunk,
                                             uint32_t* __restrict__ bsum, uint64_t* sh) {
    uint32_t v[4];
    uint64_t x[3], tot[3];
    thread_tiles(blk, count, slots, chunk, v, x + 1, x[0]);
    block_scan<3>(x, tot, sh);
    if (threadIdx.x == 0) {
        uint32_t* b = bsum + BSUM_W * blk;
        b[0] = (uint32_t)tot[0];
#pragma unroll
        for (uint32_t c = 0; c < NCLS; ++c) b[1 + c] = class_part(tot + 1, c);
    }
}

__global__ void __launch_bounds__(SCAN_T) k_scan_a(const uint32_t* __restrict__ count, int slots, uint32_t chunk,
                                                   uint32_t* __restrict__ bsum) {
    __shared__ uint64_t sh[3 * SCAN_T / 64];
    scan_a_block(blockIdx.x, count, slots, chunk, bsum, sh);
}

// The grid update of step n and the first scan pass over the tile counts of
// binning n + 1 are independent: one launch runs both (workgroups [0, ngrid)
// update the grid, the rest sum blocks of tile counts).
template <bool TAPE>
__global__ void __launch_bounds__(256) k_grid_and_count(const SceneInfo* __restrict__ scenes,
                                                        const uint32_t* __restrict__ tscene, int slots,
                                                        const uint32_t* __restrict__ tslot,
                                                        uint32_t* __restrict__ bflag,
                                                        const uint32_t* __restrict__ blist,
                                                        const uint32_t* __restrict__ nblist,
                                                        const float4* __restrict__ wbuf,
                                                        float4* __restrict__ mp_all, float4* __restrict__ gv_all,
                                                        uint32_t* __restrict__ flags8, int ngrid,
                                                        const uint32_t* __restrict__ count, uint32_t chunk,
                                                        uint32_t* __restrict__ bsum,
                                                        const uint32_t* __restrict__ node_bc) {
    static_assert(SCAN_T == 256, "both parts run 256 threads");
    __shared__ uint64_t sh[3 * SCAN_T / 64];
    if ((int)blockIdx.x < ngrid)
        grid_list<TAPE>((uint32_t)ngrid, scenes, tscene, tslot, bflag, blist, nblist, wbuf, mp_all, gv_all, flags8,
                        node_bc);
    else
        scan_a_block(blockIdx.x - ngrid, count, slots, chunk, bsum, sh);
}

__global__ void __launch_bounds__(SCAN_T) k_scan_b(uint32_t* __restrict__ count, int slots,
                                                   const uint32_t* __restrict__ bsum, uint32_t* __restrict__ start,
                                                   uint4* __restrict__ alist, uint32_t* __restrict__ nactive,
                                                   uint2* __restrict__ tslot, const uint32_t* __restrict__ tscene,
                                                   uint32_t* __restrict__ nblist, uint32_t max_entries,
                                                   uint32_t chunk, uint32_t* __restrict__ err) {
    __shared__ uint64_t sh[5 * SCAN_T / 64];
    // the grid update of the last binning has read the block list: empty it
    if (blockIdx.x == 0 && threadIdx.x == 0) *nblist = 0u;
    // particles and class entries of the preceding blocks; class entries of all blocks
    uint64_t g[5] = {0, 0, 0, 0, 0}, gt[5];
    for (int b = threadIdx.x; b < (int)gridDim.x; b += SCAN_T) {
        const uint32_t* s = bsum + BSUM_W * b;
        uint64_t e[2] = {0, 0};
#pragma unroll
        for (uint32_t c = 0; c < NCLS; ++c) add_class(e, c, s[1 + c]);
        if (b < (int)blockIdx.x) {
            g[0] += s[0];
            g[1] += e[0];
            g[2] += e[1];
        }
        g[3] += e[0];
        g[4] += e[1];
    }
    block_scan<5>(g, gt, sh);
    // first list position of each class: the larger classes come first
    uint32_t cbase[NCLS];
    uint32_t acc = 0;
#pragma unroll
    for (int c = NCLS - 1; c >= 0; --c) {
        cbase[c] = acc;
        acc += class_part(gt + 3, (uint32_t)c);
    }
    uint32_t v[4];
    uint64_t x[3], xt[3];
    thread_tiles(blockIdx.x, count, slots, chunk, v, x + 1, x[0]);
    block_scan<3>(x, xt, sh);
    uint32_t run = (uint32_t)(x[0] + gt[0]);
    uint32_t pos[NCLS];
#pragma unroll
    for (uint32_t c = 0; c < NCLS; ++c) pos[c] = cbase[c] + class_part(gt + 1, c) + class_part(x + 1, c);
    const int base = blockIdx.x * SCAN_B + threadIdx.x * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < slots) {
            start[base + k] = run;
            const uint32_t ne = (v[k] + chunk - 1) / chunk;
            uint32_t p = 0xFFFFFFFFu;
            if (v[k]) {
                const uint32_t cls = item_class(v[k], chunk);
                // pick pos[cls] without indexing the array by a runtime value
                p = cls == 0 ? pos[0] : cls == 1 ? pos[1] : cls == 2 ? pos[2] : pos[3];
                const uint32_t sc = tscene[base + k];
                for (uint32_t c = 0; c < ne; ++c) {
                    if (p + c < max_entries)
                        alist[p + c] = make_uint4((uint32_t)(base + k), run + c * chunk,
                                                  min(chunk, v[k] - c * chunk), (uint32_t)sc);
                    else
                        atomicOr(err, ERR_ALIST);
                }
#pragma unroll
                for (uint32_t c = 0; c < NCLS; ++c) pos[c] += c == cls ? ne : 0u;
            }
            tslot[base + k] = make_uint2(p, ne);
            run += v[k];
            count[base + k] = 0u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // an overflowing list is cut at its capacity (and reported), so that no
        // kernel reads past it
        *nactive = min(acc, max_entries);
    }
}

__global__ void k_scatter(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ start, uint32_t* __restrict__ perm, uint32_t nslots,
                          uint32_t* __restrict__ err) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t k = key[s];
    const uint32_t d = k < nslots ? start[k] + rank[s] : 0xFFFFFFFFu;
    if (d < (uint32_t)n) perm[d] = (uint32_t)s;
    else atomicOr(err, ERR_PERM);
}

// binning after the first scan pass (k_scan_a or k_grid_and_count)
hipError_t bin_rest(TiledWork& w, int n, uint32_t* perm, uint32_t* alist, uint32_t* nactive, uint32_t* tslot,
                    hipStream_t st) {
    ProfScope ps(w.prof, PK_BIN, st);
    const int nb = (w.tile_slots + SCAN_B - 1) / SCAN_B;
    k_scan_b<<<nb, SCAN_T, 0, st>>>(w.count, w.tile_slots, w.bsum, w.start, reinterpret_cast<uint4*>(alist),
                                    nactive, reinterpret_cast<uint2*>(tslot), w.tscene, w.nblist,
                                    (uint32_t)w.max_entries, (uint32_t)w.chunk, w.err);
    if (n > 0) k_scatter<<<(n + 255) / 256, 256, 0, st>>>(n, w.key, w.rank, w.start, perm, w.tile_slots, w.err);
    return hipGetLastError();
}

hipError_t bin_into(TiledWork& w, int n, uint32_t* perm, uint32_t* alist, uint32_t* nactive, uint32_t* tslot,
                    hipStream_t st) {
    ProfScope ps(w.prof, PK_BIN, st);
    const int nb = (w.tile_slots + SCAN_B - 1) / SCAN_B;
    k_scan_a<<<nb, SCAN_T, 0, st>>>(w.count, w.tile_slots, (uint32_t)w.chunk, w.bsum);
    k_scan_b<<<nb, SCAN_T, 0, st>>>(w.count, w.tile_slots, w.bsum, w.start, reinterpret_cast<uint4*>(alist),
                                    nactive, reinterpret_cast<uint2*>(tslot), w.tscene, w.nblist,
                                    (uint32_t)w.max_entries, (uint32_t)w.chunk, w.err);
    if (n > 0) k_scatter<<<(n + 255) / 256, 256, 0, st>>>(n, w.key, w.rank, w.start, perm, w.tile_slots, w.err);
    return hipGetLastError();
}

hipError_t bin(TiledWork& w, int n, hipStream_t st) {
    return bin_into(w, n, w.perm, w.active, w.nactive, w.tslot, st);
}

void fill_args(TileArgs& a, const TiledWork& w, size_t cap) {
    a.key_out = w.key;
    a.rank_out = w.rank;
    a.count = w.count;
    a.src_out = nullptr;
    a.bflag = w.bflag;
    a.blist = w.blist;
    a.nblist = w.nblist;
    a.wbuf = w.wbuf;
    a.scenes = w.scenes;
    a.cap = cap;
    a.nslots = (uint32_t)w.tile_slots;
    a.err = w.err;
    a.mt = w.mt;
}

// workgroups of the grid update (4 waves each, striding over the block list):
// one wave per possible block, up to 4,096 workgroups
int grid_wgs(const TiledWork& w) { return std::max(1, std::min((w.tile_slots + 3) / 4, 4096)); }

// tslot: the binning whose tile kernel filled the window buffers
template <bool TAPE>
void grid_update(const TiledWork& w, const uint32_t* tslot, float4* mp, float4* gv, uint32_t* flags8,
                 hipStream_t st) {
    ProfScope ps(w.prof, PK_GRID, st);
    k_grid_nodes<TAPE><<<grid_wgs(w), 256, 0, st>>>(w.scenes, w.tscene, tslot, w.bflag, w.blist, w.nblist, w.wbuf,
                                                     mp, gv, flags8, w.node_bc);
    // no binning follows a call's first grid update: empty the list here
    k_list_reset<<<1, 1, 0, st>>>(w.nblist);
}

// grid update of the binning whose tile kernel filled the window buffers
// (tslot), fused with the first scan pass of the next binning
template <bool TAPE>
void grid_update_and_count(const TiledWork& w, const uint32_t* tslot, float4* mp, float4* gv, uint32_t* flags8,
                           hipStream_t st) {
    ProfScope ps(w.prof, PK_GRID, st);
    const int ngrid = grid_wgs(w);
    const int nb = (w.tile_slots + SCAN_B - 1) / SCAN_B;
    k_grid_and_count<TAPE><<<ngrid + nb, 256, 0, st>>>(w.scenes, w.tscene, w.tile_slots, tslot, w.bflag, w.blist,
                                                        w.nblist, w.wbuf, mp, gv, flags8, ngrid, w.count,
                                                        (uint32_t)w.chunk, w.bsum, w.node_bc);
}

}  // namespace

hipError_t tiled_init_keys(TiledWork& w, int n, const float4* S, const uint32_t* pid, hipStream_t st) {
    if (n > 0)
        k_init_keys<<<(n + 255) / 256, 256, 0, st>>>(n, S, pid, w.scenes, w.nscenes, w.key, w.rank, w.count);
    return bin(w, n, st);
}

hipError_t tiled_steps(TiledWork& w, int n, size_t cap, float4*& S, float4*& S2, uint32_t*& pid,
                       uint32_t*& pid2, float*& E, float*& E2, float*& mass, float*& mass2, float4* mp, float4* gv,
                       int nsteps, hipStream_t st, bool primed, bool leave_primed) {
    if (nsteps <= 0) return hipSuccess;
    TileArgs a;
    fill_args(a, w, cap);
    a.perm = w.perm;
    a.alist = reinterpret_cast<const uint4*>(w.active);
    a.nactive = w.nactive;
    a.gv = gv;
    a.mp = mp;
    auto set_io = [&]() {
        a.Sin = S; a.Sout = S2;
        a.pid_in = pid; a.pid_out = pid2;
        a.E_in = E; a.E_out = E2;
        a.m_in = mass; a.m_out = mass2;
    };
    hipError_t e;
    if (primed) {
        // the previous call's fused kernel added this state's P2G into the windows
        // of its binning (w.tslot) and counted this state's tiles
        grid_update_and_count<false>(w, w.tslot, mp, gv, nullptr, st);
        e = bin_rest(w, n, w.perm, w.active, w.nactive, w.tslot, st);
        if (e != hipSuccess) return e;
    } else {
        e = tiled_init_keys(w, n, S, pid, st);
        if (e != hipSuccess) return e;
        set_io();
        {
            ProfScope ps(w.prof, PK_TILE_P2G, st);
            launch_tile<2>(w, a, st);
        }
        // the grid update reads the window buffers by this binning's active list,
        // so it runs before the next binning
        grid_update<false>(w, w.tslot, mp, gv, nullptr, st);
    }
    for (int s = 0; s + 1 < nsteps; ++s) {
        set_io();
        {
            ProfScope ps(w.prof, PK_TILE_G2P2G, st);
            launch_tile<3>(w, a, st);
        }
        grid_update_and_count<false>(w, w.tslot, mp, gv, nullptr, st);
        e = bin_rest(w, n, w.perm, w.active, w.nactive, w.tslot, st);
        if (e != hipSuccess) return e;
        std::swap(S, S2); std::swap(pid, pid2); std::swap(E, E2); std::swap(mass, mass2);
    }
    set_io();
    if (leave_primed) {
        ProfScope ps(w.prof, PK_TILE_G2P2G, st);
        launch_tile<3>(w, a, st);
    } else {
        ProfScope ps(w.prof, PK_TILE_G2P, st);
        launch_tile<1>(w, a, st);
    }
    std::swap(S, S2); std::swap(pid, pid2); std::swap(E, E2); std::swap(mass, mass2);
    return hipGetLastError();
}

hipError_t tiled_settle(TiledWork& w, int n, float4* mp, float4* gv, hipStream_t st) {
    grid_update_and_count<false>(w, w.tslot, mp, gv, nullptr, st);
    return bin_rest(w, n, w.perm, w.active, w.nactive, w.tslot, st);
}

namespace {

// clears the grid (m, p) of the flagged 8^3 blocks of a tape step and the flags
__global__ void __launch_bounds__(512) k_tape_clear(const SceneInfo* __restrict__ scenes,
                                                    const uint32_t* __restrict__ bscene, uint32_t* __restrict__ flags,
                                                    float4* __restrict__ mp_all) {
    const int b = blockIdx.x;
    if (flags[b] == 0u) return;
    const SceneInfo sc = scenes[bscene[b]];
    const int lb = b - (int)sc.boff;
    const int bx = lb % sc.gbx, by = (lb / sc.gbx) % sc.gby, bz = lb / (sc.gbx * sc.gby);
    const int i = bx * GB + (int)(threadIdx.x & 7), j = by * GB + (int)((threadIdx.x >> 3) & 7),
              k = bz * GB + (int)(threadIdx.x >> 6);
    if (i < sc.g.nx && j < sc.g.ny && k < sc.g.nz)
        mp_all[sc.goff + node_id(sc.g, i, j, k)] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    if (threadIdx.x == 0) flags[b] = 0u;
}

}  // namespace

hipError_t tape_clear(const TiledWork& w, const TapeStep& t, hipStream_t st) {
    if (w.total_blocks > 0) k_tape_clear<<<w.total_blocks, 512, 0, st>>>(w.scenes, w.bscene, t.flags, t.mp);
    return hipGetLastError();
}

hipError_t tiled_tape_forward(TiledWork& w, int n, size_t cap, TapeStep* tape, int L, hipStream_t st) {
    if (L <= 0) return hipSuccess;
    TileArgs a;
    fill_args(a, w, cap);
    if (n > 0)
        k_init_keys<<<(n + 255) / 256, 256, 0, st>>>(n, tape[0].S, tape[0].pid, w.scenes, w.nscenes, w.key, w.rank,
                                                     w.count);
    hipError_t e = bin_into(w, n, tape[0].perm, tape[0].alist, tape[0].nactive, tape[0].tslot, st);
    if (e != hipSuccess) return e;
    // P2G of step 0 and its grid
    a.Sin = tape[0].S;
    a.pid_in = tape[0].pid;
    a.E_in = tape[0].E;
    a.m_in = tape[0].mass;
    a.perm = tape[0].perm;
    a.alist = reinterpret_cast<const uint4*>(tape[0].alist);
    a.nactive = tape[0].nactive;
    a.mp = tape[0].mp;
    {
        ProfScope ps(w.prof, PK_TILE_P2G, st);
        launch_tile<2>(w, a, st);
    }
    grid_update<true>(w, tape[0].tslot, tape[0].mp, tape[0].gv, tape[0].flags, st);
    for (int s = 0; s < L; ++s) {
        TapeStep& cur = tape[s];
        TapeStep& nxt = tape[s + 1];
        a.Sin = cur.S;
        a.pid_in = cur.pid;
        a.E_in = cur.E;
        a.m_in = cur.mass;
        a.perm = cur.perm;
        a.alist = reinterpret_cast<const uint4*>(cur.alist);
        a.nactive = cur.nactive;
        a.gv = cur.gv;
        a.Sout = nxt.S;
        a.pid_out = nxt.pid;
        a.E_out = nxt.E;
        a.m_out = nxt.mass;
        a.src_out = cur.src;
        if (s + 1 < L) {
            // G2P of step s, P2G of step s+1 (window buffers by this binning), its grid,
            // then the binning of step s+1
            a.mp = nxt.mp;
            {
                ProfScope ps(w.prof, PK_TILE_G2P2G, st);
                launch_tile<3>(w, a, st);
            }
            grid_update_and_count<true>(w, cur.tslot, nxt.mp, nxt.gv, nxt.flags, st);
            e = bin_rest(w, n, nxt.perm, nxt.alist, nxt.nactive, nxt.tslot, st);
            if (e != hipSuccess) return e;
        } else {
            // last step of the segment: G2P only, to fix the end state's storage order
            ProfScope ps(w.prof, PK_TILE_G2P, st);
            launch_tile<1>(w, a, st);
        }
    }
    return hipGetLastError();
}

}  // namespace pg
