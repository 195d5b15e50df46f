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
// Transfers.  One workgroup processes one tile at a time.  Inside the tile
// the particles are ordered by stencil-base cell with an LDS counting sort
// (64 cells), then processed 256 at a time:
//   G2P  gathers the 27 node velocities from the tile's (TT+2)^3 window of
//        grid velocity held in LDS;
//   P2G  writes the particle's contributions to a per-wave scratch and
//        reduces them across the wave (wave_slice_reduce): runs of particles
//        sharing a cell become one LDS atomic per node and channel, and the
//        window is flushed with one global atomic per node and channel.
// G2P of step n and P2G of step n+1 are fused (the new state never leaves
// registers) and the new state is written at the particle's cell-sorted
// position, so storage stays sorted and the next step's gathers coalesce;
// pid carries the user's particle index.  A particle whose stencil leaves the
// window (it crossed a tile face this step) scatters with global atomics, so
// correctness never depends on the binning.
//
// Grid update.  Runs over 8^3 node blocks flagged by the P2G flush; writes
// (v, m) and clears the accumulator, so there is no dense grid clear.
#include "pg_tile.h"

namespace pg {
namespace {

constexpr int PASS = 2048;       // particles sorted per LDS pass of a tile
constexpr int NW = TB / 64;      // waves per workgroup
constexpr uint16_t FAR = 0xFFFF; // direction code of a particle that jumped more than one tile

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
    uint32_t* flags;         // concatenated 8^3 block flags of all scenes
    const SceneInfo* scenes;
    size_t cap;
};

__device__ __forceinline__ void flag_node(uint32_t* flags, const SceneInfo& sc, int gi, int gj, int gk) {
    flags[((gk / GB) * sc.gby + gj / GB) * sc.gbx + gi / GB] = 1u;
}

__device__ __forceinline__ uint32_t morton_tile(int tx, int ty, int tz) {
    return expand3((uint32_t)tx) | (expand3((uint32_t)ty) << 1) | (expand3((uint32_t)tz) << 2);
}

// tile coordinates of a position, clamped like tile_key
__device__ __forceinline__ void tile_coords(const float* x, const MlsGeom& g, int ntx, int nty, int ntz, int* t) {
    const int nt[3] = {ntx, nty, ntz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int b = (int)floorf((x[a] - g.o[a]) * g.inv_dx - 0.5f);
        int tc = b >= 0 ? b / TT : 0;
        t[a] = tc < nt[a] ? tc : nt[a] - 1;
    }
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

// P2G of one particle: node n = base + (x, y, z) receives w_n (m, q + a0 x + a1 y + a2 z)
// with A = -dt V0 4/dx^2 tau + m C, q = m v - dx A f, a_j = dx A[:, j].
struct P2GVal {
    float m;
    float q[3], a0[3], a1[3], a2[3];
};

__device__ __forceinline__ void p2g_prep(const MlsGeom& g, float Eq, float mq, const float* F, const float* C,
                                         const float* v, const Stencil& st, P2GVal& p) {
    const float kscale = -g.dt * g.p_vol * 4.f * g.inv_dx * g.inv_dx;
    const float mu = Eq / (2.f * (1.f + g.nu));
    const float lam = Eq * g.nu / ((1.f + g.nu) * (1.f - 2.f * g.nu));
    float tau[9];
    kirchhoff(g.material, mu, lam, F, tau);
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

// visits the 27 contributions: emit(x, y, z, float4(w m, w p))
template <class Emit>
__device__ __forceinline__ void p2g_visit(const Stencil& st, const P2GVal& p, Emit emit) {
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int y = 0; y < 3; ++y) {
            const float wyz = st.w[1][y] * st.w[2][z];
            float b[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) b[c] = p.q[c] + p.a1[c] * (float)y + p.a2[c] * (float)z;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const float w = wyz * st.w[0][x];
                emit(x, y, z, make_float4(w * p.m, w * (b[0] + p.a0[0] * (float)x), w * (b[1] + p.a0[1] * (float)x),
                                          w * (b[2] + p.a0[2] * (float)x)));
            }
        }
}

// Tile kernel.  MODE bit 1: G2P (reads x, F; writes the new state);
// MODE bit 2: P2G (scatters the current (MODE 2) or updated (MODE 3) state).
//
// Per pass of up to PASS particles of the tile: an LDS counting sort by
// stencil-base cell (64 cells, plus "outside" for clamped keys), then rounds:
// in round r lane c of a wave takes the r-th particle of cell c, waves take
// rounds in turn.  All lanes of a round sit in distinct cells, so the 27 node
// updates of a lane's P2G never meet another lane's in the same instruction,
// and each wave accumulates into its own LDS window with plain
// read-modify-writes in program order.  A particle that left its cell during
// G2P adds with LDS atomics (inside the window) or global atomics (outside).
// Outputs are written round by round, so stores and the next step's loads
// coalesce.  At the end of the tile the four windows are summed and added to
// the grid.
template <int MODE>
__global__ void __launch_bounds__(TB, 3) k_tile(TileArgs a) {
    constexpr bool G2P = (MODE & 1) != 0;
    constexpr bool P2G = (MODE & 2) != 0;
    __shared__ float4 sv[G2P ? WV : 1];          // grid (v, m) window
    __shared__ float4 pw[P2G ? NW : 1][WV];      // per-wave P2G windows (m, p)
    __shared__ uint16_t lperm[PASS];             // local cell-sorted order within the pass
    __shared__ uint16_t rs[PASS];                // first output offset of each round
    __shared__ uint16_t ldir[(G2P && P2G) ? PASS : 1];  // (direction << 11 | rank) of the new tile, or FAR
    __shared__ uint32_t lcnt[65], lstart[65];
    __shared__ uint32_t dcnt[27], dbase[27];
    __shared__ uint32_t s_max, s_tot;

    const XcdChunk ch(*a.nactive);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int ccx = lane & 3, ccy = (lane >> 2) & 3, ccz = lane >> 4;     // this lane's cell in a round
    const int cnb = (ccz * WN + ccy) * WN + ccx;                          // its base node in the window

    for (uint32_t ti = ch.first; ti < ch.end; ti += ch.stride) {
        const uint4 te = a.alist[ti];
        const SceneInfo& sc = a.scenes[te.w];
        const MlsGeom g = sc.g;
        const float4* __restrict__ gvs = a.gv + sc.goff;
        float4* __restrict__ mps = a.mp + sc.goff;
        uint32_t* __restrict__ flags = a.flags + sc.boff;
        const uint32_t tile = te.x - sc.toff, beg = te.y, cnt = te.z;
        int t0[3];
        tile_origin(tile, t0);
        const int tc[3] = {t0[0] / TT, t0[1] / TT, t0[2] / TT};
        for (int l = tid; l < WV; l += TB) {
            if (G2P) {
                const int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
                float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
                if (gi < g.nx && gj < g.ny && gk < g.nz) u = gvs[node_id(g, gi, gj, gk)];
                sv[l] = u;
            }
            if (P2G) {
#pragma unroll
                for (int w = 0; w < NW; ++w) pw[w][l] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }

        for (uint32_t pb = 0; pb < cnt; pb += PASS) {
            const uint32_t pc = min((uint32_t)PASS, cnt - pb);
            // ---- counting sort of the pass by stencil-base cell
            if (tid < 65) lcnt[tid] = 0;
            if (tid < 27) dcnt[tid] = 0;
            __syncthreads();
            uint32_t key[PASS / TB];
#pragma unroll
            for (int k = 0; k < PASS / TB; ++k) {
                const uint32_t r = tid + k * TB;
                key[k] = 0u;
                if (r < pc) {
                    const float4 q = a.Sin[a.perm[beg + pb + r]];
                    const float xx[3] = {q.x, q.y, q.z};
                    int l3[3];
                    bool in = true;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        l3[d] = (int)floorf((xx[d] - g.o[d]) * g.inv_dx - 0.5f) - t0[d];
                        in = in && l3[d] >= 0 && l3[d] < TT;
                    }
                    const int cell = in ? l3[0] + TT * l3[1] + TT * TT * l3[2] : 64;
                    const uint32_t rk = atomicAdd(&lcnt[cell], 1u);
                    key[k] = ((uint32_t)cell << 16) | rk;
                }
            }
            __syncthreads();
            if (wave == 0) {
                // cell starts (wave scan), the largest cell count, round offsets
                const uint32_t cl = lcnt[lane];
                uint32_t inc = cl;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t t = __shfl_up(inc, off);
                    if (lane >= off) inc += t;
                }
                lstart[lane] = inc - cl;
                if (lane == 63) lstart[64] = inc;
                uint32_t mx = cl;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
                uint32_t run = 0;
                for (uint32_t r = 0; r < mx; ++r) {
                    const uint64_t m = __ballot(cl > r);
                    if (lane == 0) rs[r] = (uint16_t)run;
                    run += (uint32_t)__popcll(m);
                }
                if (lane == 0) {
                    s_max = mx;
                    s_tot = run;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PASS / TB; ++k) {
                const uint32_t r = tid + k * TB;
                if (r < pc) lperm[lstart[key[k] >> 16] + (key[k] & 0xFFFFu)] = (uint16_t)r;
            }
            __syncthreads();

            // one particle: li = index within the pass (storage order), oo = output offset,
            // regular = it is this lane's cell in a round (G2P from the window, P2G by RMW)
            auto particle = [&](uint32_t li, uint32_t oo, bool regular) {
                const uint32_t s = a.perm[beg + pb + li];
                const uint32_t o = beg + pb + oo;
                float Eq = g.E, mq = g.p_mass;
                if (a.E_in) Eq = a.E_in[s];
                if (a.m_in) mq = a.m_in[s];
                float x[3], F[9], v[3], C[9];
                if (G2P) {
                    load_xF(a.Sin, a.cap, s, x, F);
                    Stencil st;
                    make_stencil(x, g.o, g.inv_dx, st);
                    // (the base equals the lane's cell by construction of the sort; checked so
                    // that a different rounding of the two computations cannot matter)
                    if (regular && st.b[0] - t0[0] == ccx && st.b[1] - t0[1] == ccy && st.b[2] - t0[2] == ccz) {
                        g2p_sum([&](int ix, int iy, int iz) { return sv[cnb + (iz * WN + iy) * WN + ix]; }, st,
                                g.inv_dx, v, C);
                    } else {
                        g2p_sum(
                            [&](int ix, int iy, int iz) {
                                const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
                                return in_grid(g, gi, gj, gk) ? gvs[node_id(g, gi, gj, gk)]
                                                              : make_float4(0.f, 0.f, 0.f, 0.f);
                            },
                            st, g.inv_dx, v, C);
                    }
                    float Fn[9];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            Fn[3 * i + j] = F[3 * i + j] + g.dt * (C[3 * i] * F[j] + C[3 * i + 1] * F[3 + j] +
                                                                   C[3 * i + 2] * F[6 + j]);
#pragma unroll
                    for (int c = 0; c < 9; ++c) F[c] = Fn[c];
#pragma unroll
                    for (int c = 0; c < 3; ++c) x[c] += g.dt * v[c];
                    store_state(a.Sout, a.cap, o, x, F, v, C);
                    a.pid_out[o] = a.pid_in[s];
                    if (a.E_out && a.E_in) a.E_out[o] = Eq;
                    if (a.m_out && a.m_in) a.m_out[o] = mq;
                    if (a.src_out) a.src_out[o] = s;
                    if (P2G) {
                        int nt[3];
                        tile_coords(x, g, sc.ntx, sc.nty, sc.ntz, nt);
                        const int dx_ = nt[0] - tc[0], dy_ = nt[1] - tc[1], dz_ = nt[2] - tc[2];
                        const uint32_t k = sc.toff + morton_tile(nt[0], nt[1], nt[2]);
                        a.key_out[o] = k;
                        const bool near = dx_ >= -1 && dx_ <= 1 && dy_ >= -1 && dy_ <= 1 && dz_ >= -1 && dz_ <= 1;
                        const int dir = near ? (dx_ + 1) + 3 * (dy_ + 1) + 9 * (dz_ + 1) : -1;
                        // most particles stay in the tile: one LDS atomic per wave for those
                        const uint64_t stay = __ballot(dir == 13);
                        uint32_t rk = 0;
                        if (stay) {
                            const int leader = __ffsll((unsigned long long)stay) - 1;
                            uint32_t b0 = 0;
                            if (lane == leader) b0 = atomicAdd(&dcnt[13], (uint32_t)__popcll(stay));
                            b0 = __builtin_amdgcn_readlane(b0, leader);
                            rk = b0 + (uint32_t)__popcll(stay & lt);
                        }
                        if (dir >= 0 && dir != 13) rk = atomicAdd(&dcnt[dir], 1u);
                        if (dir >= 0) {
                            ldir[oo] = (uint16_t)((dir << 11) | rk);
                        } else {
                            ldir[oo] = FAR;
                            a.rank_out[o] = atomicAdd(a.count + k, 1u);
                        }
                    }
                } else {
                    load_state(a.Sin, a.cap, s, x, F, v, C);
                }
                if (P2G) {
                    Stencil st;
                    make_stencil(x, g.o, g.inv_dx, st);
                    P2GVal p;
                    p2g_prep(g, Eq, mq, F, C, v, st, p);
                    const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
                    const bool inwin = l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
                    const int nb = (l2 * WN + l1) * WN + l0;
                    float4* __restrict__ win = pw[wave];
                    if (regular && inwin && nb == cnb) {
                        p2g_visit(st, p, [&](int ix, int iy, int iz, float4 c) {
                            float4* t = win + nb + (iz * WN + iy) * WN + ix;
                            float4 u = *t;
                            u.x += c.x;
                            u.y += c.y;
                            u.z += c.z;
                            u.w += c.w;
                            *t = u;
                            lds_order();
                        });
                    } else if (inwin) {
                        float* wf = reinterpret_cast<float*>(win);
                        p2g_visit(st, p, [&](int ix, int iy, int iz, float4 c) {
                            float* t = wf + 4 * (nb + (iz * WN + iy) * WN + ix);
                            atomicAdd(t + 0, c.x);
                            atomicAdd(t + 1, c.y);
                            atomicAdd(t + 2, c.z);
                            atomicAdd(t + 3, c.w);
                        });
                    } else {
                        p2g_visit(st, p, [&](int ix, int iy, int iz, float4 c) {
                            const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
                            if (!in_grid(g, gi, gj, gk)) return;
                            float* cp = reinterpret_cast<float*>(mps + node_id(g, gi, gj, gk));
                            atomic_add_f32(cp + 0, c.x);
                            atomic_add_f32(cp + 1, c.y);
                            atomic_add_f32(cp + 2, c.z);
                            atomic_add_f32(cp + 3, c.w);
                            flag_node(flags, sc, gi, gj, gk);
                        });
                    }
                    lds_order();
                }
            };

            // ---- rounds: lane = cell, round r = the r-th particle of every cell
            const uint32_t maxcnt = s_max, tot = s_tot;
            const uint32_t mycnt = lcnt[lane], mystart = lstart[lane];
            for (uint32_t r = wave; r < maxcnt; r += NW) {
                const bool act = r < mycnt;
                const uint64_t am = __ballot(act);
                const uint32_t oo = rs[r] + (uint32_t)__popcll(am & lt);
                if (act) particle(lperm[mystart + r], oo, true);
            }
            // ---- particles whose stencil base is outside the tile (clamped keys)
            const uint32_t nout = lcnt[64], sout = lstart[64];
            for (uint32_t t = tid; t < nout; t += TB) particle(lperm[sout + t], tot + t, false);
            __syncthreads();
            // ---- ranks in the new tiles: one global atomic per (tile, direction)
            if (G2P && P2G) {
                if (tid < 27) {
                    const uint32_t nd = dcnt[tid];
                    if (nd) {
                        const int dx_ = tid % 3 - 1, dy_ = (tid / 3) % 3 - 1, dz_ = tid / 9 - 1;
                        const uint32_t k = sc.toff + morton_tile(tc[0] + dx_, tc[1] + dy_, tc[2] + dz_);
                        dbase[tid] = atomicAdd(a.count + k, nd);
                    }
                }
                __syncthreads();
                for (uint32_t jj = tid; jj < pc; jj += TB) {
                    const uint16_t dr = ldir[jj];
                    if (dr != FAR) a.rank_out[beg + pb + jj] = dbase[dr >> 11] + (dr & 0x7FFu);
                }
                __syncthreads();
            }
        }

        if (P2G) {
            __syncthreads();
            for (int l = tid; l < WV; l += TB) {
                float4 m = pw[0][l];
#pragma unroll
                for (int w = 1; w < NW; ++w) {
                    const float4 u = pw[w][l];
                    m.x += u.x;
                    m.y += u.y;
                    m.z += u.z;
                    m.w += u.w;
                }
                if (m.x != 0.f) {
                    const int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
                    if (gi < g.nx && gj < g.ny && gk < g.nz) {
                        float* cp = reinterpret_cast<float*>(mps + node_id(g, gi, gj, gk));
                        atomic_add_f32(cp + 0, m.x);
                        atomic_add_f32(cp + 1, m.y);
                        atomic_add_f32(cp + 2, m.z);
                        atomic_add_f32(cp + 3, m.w);
                    }
                }
            }
            if (tid < 8) {
                const int bx = t0[0] / GB + (tid & 1), by = t0[1] / GB + ((tid >> 1) & 1), bz = t0[2] / GB + (tid >> 2);
                // the window spans nodes [t0, t0 + WN); flag each 8^3 block it touches
                const bool touch = bx * GB < t0[0] + WN && by * GB < t0[1] + WN && bz * GB < t0[2] + WN;
                if (touch && bx < sc.gbx && by < sc.gby && bz < sc.gbz) flags[(bz * sc.gby + by) * sc.gbx + bx] = 1u;
            }
        }
        __syncthreads();
    }
}

// Grid update over flagged 8^3 blocks: gv = (p/m + dt g, m) with sticky walls.
// Forward: the accumulator and the flag are cleared for the next P2G.  TAPE
// (segment recompute for the adjoint): both are kept for the reverse pass,
// which clears them after use.
template <bool TAPE>
__global__ void __launch_bounds__(512) k_grid_tiled(const SceneInfo* __restrict__ scenes, int nscenes,
                                                    float4* __restrict__ mp_all, float4* __restrict__ gv_all,
                                                    uint32_t* __restrict__ flags) {
    const int b = blockIdx.x;
    if (flags[b] == 0u) return;
    const SceneInfo& sc = scenes[scene_of(scenes, nscenes, (uint32_t)b, [](const SceneInfo& s) { return s.boff; })];
    const MlsGeom g = sc.g;
    float4* __restrict__ mp = mp_all + sc.goff;
    float4* __restrict__ gv = gv_all + sc.goff;
    const int lb = b - (int)sc.boff;
    const int bx = lb % sc.gbx, by = (lb / sc.gbx) % sc.gby, bz = lb / (sc.gbx * sc.gby);
    const int i = bx * GB + (threadIdx.x & 7), j = by * GB + ((threadIdx.x >> 3) & 7), k = bz * GB + (threadIdx.x >> 6);
    if (i < g.nx && j < g.ny && k < g.nz) {
        const size_t id = node_id(g, i, j, k);
        const float4 q = mp[id];
        float4 r = make_float4(0.f, 0.f, 0.f, q.x);
        if (q.x > 0.f && !is_sticky(g, i, j, k)) {
            const float im = 1.f / q.x;
            r.x = q.y * im + g.dt * g.g[0];
            r.y = q.z * im + g.dt * g.g[1];
            r.z = q.w * im + g.dt * g.g[2];
        }
        gv[id] = r;
        if (!TAPE) mp[id] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (!TAPE) {
        __syncthreads();
        if (threadIdx.x == 0) flags[b] = 0u;
    }
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
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t kl = __builtin_amdgcn_readlane(k, leader);
        const uint64_t peers = __ballot(k == kl) & todo;
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(count + kl, (uint32_t)__popcll(peers));
        b = __builtin_amdgcn_readlane(b, leader);
        if (k == kl && valid) rk = b + (uint32_t)__popcll(peers & lt);
        todo &= ~peers;
    }
    if (valid) {
        key[s] = k;
        rank[s] = rk;
    }
}

// Scan of the tile counts in two kernels over blocks of SCAN_B slots:
// block sums and active counts, then per-block exclusive scan with the
// preceding blocks' totals; writes starts, the active tile list, clears counts.
constexpr int SCAN_T = 256;
constexpr int SCAN_B = 1024;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        uint32_t t = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    const uint32_t incl = sh[tid];
    *total = sh[SCAN_T - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_a(const uint32_t* __restrict__ count, int slots,
                                                   uint32_t* __restrict__ bsum) {
    __shared__ uint32_t sh[SCAN_T];
    const int base = blockIdx.x * SCAN_B + threadIdx.x * 4;
    uint32_t s = 0, act = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = base + k < slots ? count[base + k] : 0u;
        s += v;
        act += v ? 1u : 0u;
    }
    uint32_t ts, ta;
    block_excl_scan(s, sh, &ts);
    block_excl_scan(act, sh, &ta);
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = ts;
        bsum[2 * blockIdx.x + 1] = ta;
    }
}

__global__ void __launch_bounds__(SCAN_T) k_scan_b(uint32_t* __restrict__ count, int slots,
                                                   const uint32_t* __restrict__ bsum, uint32_t* __restrict__ start,
                                                   uint4* __restrict__ alist, uint32_t* __restrict__ nactive,
                                                   const SceneInfo* __restrict__ scenes, int nscenes) {
    __shared__ uint32_t sh[SCAN_T];
    // totals of the preceding blocks
    uint32_t ps = 0, pa = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += SCAN_T) {
        ps += bsum[2 * b];
        pa += bsum[2 * b + 1];
    }
    uint32_t pre_s, pre_a;
    block_excl_scan(ps, sh, &pre_s);
    block_excl_scan(pa, sh, &pre_a);
    const int base = blockIdx.x * SCAN_B + threadIdx.x * 4;
    uint32_t v[4], s = 0, act = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < slots ? count[base + k] : 0u;
        s += v[k];
        act += v[k] ? 1u : 0u;
    }
    uint32_t ts, ta;
    uint32_t run = block_excl_scan(s, sh, &ts) + pre_s;
    uint32_t pos = block_excl_scan(act, sh, &ta) + pre_a;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < slots) {
            start[base + k] = run;
            if (v[k]) {
                const int sc = scene_of(scenes, nscenes, (uint32_t)(base + k), [](const SceneInfo& t) { return t.toff; });
                alist[pos++] = make_uint4((uint32_t)(base + k), run, v[k], (uint32_t)sc);
            }
            run += v[k];
            count[base + k] = 0u;
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *nactive = pre_a + ta;
}

__global__ void k_scatter(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ start, uint32_t* __restrict__ perm) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    perm[start[key[s]] + rank[s]] = (uint32_t)s;
}

hipError_t bin_into(TiledWork& w, int n, uint32_t* perm, uint32_t* alist, uint32_t* nactive, hipStream_t st) {
    ProfScope ps(w.prof, PK_BIN, st);
    const int nb = (w.tile_slots + SCAN_B - 1) / SCAN_B;
    k_scan_a<<<nb, SCAN_T, 0, st>>>(w.count, w.tile_slots, w.bsum);
    k_scan_b<<<nb, SCAN_T, 0, st>>>(w.count, w.tile_slots, w.bsum, w.start, reinterpret_cast<uint4*>(alist),
                                    nactive, w.scenes, w.nscenes);
    if (n > 0) k_scatter<<<(n + 255) / 256, 256, 0, st>>>(n, w.key, w.rank, w.start, perm);
    return hipGetLastError();
}

hipError_t bin(TiledWork& w, int n, hipStream_t st) { return bin_into(w, n, w.perm, w.active, w.nactive, st); }

void fill_args(TileArgs& a, const TiledWork& w, size_t cap) {
    a.key_out = w.key;
    a.rank_out = w.rank;
    a.count = w.count;
    a.src_out = nullptr;
    a.scenes = w.scenes;
    a.cap = cap;
}

}  // namespace

hipError_t tiled_init_keys(TiledWork& w, int n, const float4* S, const uint32_t* pid, hipStream_t st) {
    if (n > 0)
        k_init_keys<<<(n + 255) / 256, 256, 0, st>>>(n, S, pid, w.scenes, w.nscenes, w.key, w.rank, w.count);
    return bin(w, n, st);
}

hipError_t tiled_steps(TiledWork& w, int n, size_t cap, float4*& S, float4*& S2, uint32_t*& pid,
                       uint32_t*& pid2, float*& E, float*& E2, float*& mass, float*& mass2, float4* mp, float4* gv,
                       int nsteps, hipStream_t st) {
    if (nsteps <= 0) return hipSuccess;
    const int nblk = w.total_blocks;
    TileArgs a;
    fill_args(a, w, cap);
    a.perm = w.perm;
    a.alist = reinterpret_cast<const uint4*>(w.active);
    a.nactive = w.nactive;
    a.gv = gv;
    a.mp = mp;
    a.flags = w.flags;
    auto set_io = [&]() {
        a.Sin = S; a.Sout = S2;
        a.pid_in = pid; a.pid_out = pid2;
        a.E_in = E; a.E_out = E2;
        a.m_in = mass; a.m_out = mass2;
    };
    hipError_t e = tiled_init_keys(w, n, S, pid, st);
    if (e != hipSuccess) return e;
    set_io();
    {
        ProfScope ps(w.prof, PK_TILE_P2G, st);
        k_tile<2><<<w.nblocks_launch, TB, 0, st>>>(a);
    }
    {
        ProfScope ps(w.prof, PK_GRID, st);
        k_grid_tiled<false><<<nblk, 512, 0, st>>>(w.scenes, w.nscenes, mp, gv, w.flags);
    }
    for (int s = 0; s + 1 < nsteps; ++s) {
        set_io();
        {
            ProfScope ps(w.prof, PK_TILE_G2P2G, st);
            k_tile<3><<<w.nblocks_launch, TB, 0, st>>>(a);
        }
        {
            ProfScope ps(w.prof, PK_GRID, st);
            k_grid_tiled<false><<<nblk, 512, 0, st>>>(w.scenes, w.nscenes, mp, gv, w.flags);
        }
        e = bin(w, n, st);
        if (e != hipSuccess) return e;
        std::swap(S, S2); std::swap(pid, pid2); std::swap(E, E2); std::swap(mass, mass2);
    }
    set_io();
    {
        ProfScope ps(w.prof, PK_TILE_G2P, st);
        k_tile<1><<<w.nblocks_launch, TB, 0, st>>>(a);
    }
    std::swap(S, S2); std::swap(pid, pid2); std::swap(E, E2); std::swap(mass, mass2);
    return hipGetLastError();
}

hipError_t tiled_tape_forward(TiledWork& w, int n, size_t cap, TapeStep* tape, int L, hipStream_t st) {
    if (L <= 0) return hipSuccess;
    const int nblk = w.total_blocks;
    TileArgs a;
    fill_args(a, w, cap);
    if (n > 0)
        k_init_keys<<<(n + 255) / 256, 256, 0, st>>>(n, tape[0].S, tape[0].pid, w.scenes, w.nscenes, w.key, w.rank,
                                                     w.count);
    hipError_t e = bin_into(w, n, tape[0].perm, tape[0].alist, tape[0].nactive, st);
    if (e != hipSuccess) return e;
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
        if (s == 0) {
            a.mp = cur.mp;
            a.flags = cur.flags;
            ProfScope ps(w.prof, PK_TILE_P2G, st);
            k_tile<2><<<w.nblocks_launch, TB, 0, st>>>(a);
        }
        {
            ProfScope ps(w.prof, PK_GRID, st);
            k_grid_tiled<true><<<nblk, 512, 0, st>>>(w.scenes, w.nscenes, cur.mp, cur.gv, cur.flags);
        }
        a.Sout = nxt.S;
        a.pid_out = nxt.pid;
        a.E_out = nxt.E;
        a.m_out = nxt.mass;
        a.src_out = cur.src;
        if (s + 1 < L) {
            a.mp = nxt.mp;
            a.flags = nxt.flags;
            {
                ProfScope ps(w.prof, PK_TILE_G2P2G, st);
                k_tile<3><<<w.nblocks_launch, TB, 0, st>>>(a);
            }
            e = bin_into(w, n, nxt.perm, nxt.alist, nxt.nactive, st);
            if (e != hipSuccess) return e;
        } else {
            // last step of the segment: G2P only, to fix the end state's storage order
            ProfScope ps(w.prof, PK_TILE_G2P, st);
            k_tile<1><<<w.nblocks_launch, TB, 0, st>>>(a);
        }
    }
    return hipGetLastError();
}

}  // namespace pg
