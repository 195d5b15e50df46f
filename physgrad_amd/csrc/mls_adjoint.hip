// Reverse (adjoint) step of the MLS-MPM transfer loop on the tiled layout.
//
// Replaces the empty backward of DifferentiableMPMSystem
// (src/differentiable_mpm.h:823-878, tape lambdas that do nothing) and the
// gravity-only launch_mpm_timestep_backward (src/pytorch_autograd.cu:385-402)
// with the exact reverse of one step: G2P, grid update (incl. sticky-wall
// masks) and P2G with the elastic stress, differentiated by hand.  Works on a
// segment tape written by tiled_tape_forward; for step n of the tape:
//
//   A1  k_adj_g2p   tiles of binning n, particles in the step's output (cell)
//                   order: recompute v, C from the gv window in LDS, push the
//                   adjoint of the new (v, C) onto the grid (wave transpose-
//                   reduce into an LDS window, one global atomic per node and
//                   channel), accumulate the position adjoint from the weights
//                   and x_i - x_p; partial = (ax, aF) by storage slot.
//   A2  k_adj_grid  flagged 8^3 blocks: ap = agv/m, am = -(agv.p)/m^2 (zero
//                   on sticky or empty nodes); clears agv, the tape's
//                   accumulator and flags.
//   A3  k_adj_p2g   tiles of binning n: gather (am, ap) from the LDS window,
//                   adjoints of v, C, F (stress VJP incl. polar derivative),
//                   x, and dL/dE (accumulated in user particle order).
//
// Indexing: the adjoint of the state after step n is stored by storage slot
// of step n+1, which is the output position o of step n (src_n[o] is the
// particle's slot in step n); A3 writes the adjoint of the state before step n
// at its slot of step n, which is what A1 of step n-1 reads by its output
// position.
#include "pg_tile.h"

namespace pg {
namespace {

struct AdjArgs {
    const float4* S;        // tape state of step n
    const uint32_t* pid;    // slot -> user index (step n)
    const float* E;         // per-particle E, slot order (or NULL)
    const float* mass;      // per-particle mass, slot order (or NULL)
    const uint32_t* perm;   // sorted position of binning n -> slot of step n
    const uint32_t* src;    // output position of step n -> slot of step n
    const uint4* alist;
    const uint32_t* nactive;
    const float4* gv;       // grid (v, m) of step n
    float4* abar;           // adjoint state planes
    float4* part;           // partial adjoint planes (3*cap) by slot of step n: (ax, aF)
    float4* agv;            // grid accumulator of the velocity adjoint (x,y,z used)
    const float4* apam;     // (am, ap) per node
    float* aE_user;         // dL/dE in user order (or NULL)
    const SceneInfo* scenes;
    size_t cap;
};

PG_DEV void load_adj(const float4* __restrict__ A, size_t cap, size_t i, float* ax, float* aF, float* av, float* aC) {
    load_state(A, cap, i, ax, aF, av, aC);
}

PG_DEV void weight_grads(const Stencil& st, float dw[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float f = st.f[a];
        dw[a][0] = f - 1.5f;
        dw[a][1] = -2.f * (f - 1.f);
        dw[a][2] = f - 0.5f;
    }
}

__global__ void __launch_bounds__(TB) k_adj_g2p(AdjArgs a) {
    constexpr int NP = 27;               // 9 nodes x 3 channels per slice
    __shared__ float4 sv[WV];
    __shared__ float sa[3 * WV];         // velocity-adjoint window, channel-major
    __shared__ float scr[TB / 64][64 * NP];
    const XcdChunk ch(*a.nactive);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* myscr = scr[wave];
    for (uint32_t ti = ch.first; ti < ch.end; ti += ch.stride) {
        const uint4 te = a.alist[ti];
        const SceneInfo sc = a.scenes[te.w];   // by value: kept in registers across the stores
        const MlsGeom g = sc.g;
        const float4* __restrict__ gvs = a.gv + sc.goff;
        float4* __restrict__ agvs = a.agv + sc.goff;
        const float s4 = 4.f * g.inv_dx * g.inv_dx;
        const uint32_t beg = te.y, cnt = te.z;
        int t0[3];
        tile_origin(te.x - sc.toff, t0);
        for (int l = tid; l < WV; l += TB) {
            int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
            float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gi < g.nx && gj < g.ny && gk < g.nz) u = gvs[node_id(g, gi, gj, gk)];
            sv[l] = u;
            sa[l] = 0.f; sa[WV + l] = 0.f; sa[2 * WV + l] = 0.f;
        }
        __syncthreads();
        // output positions of the step in cell order; src gives the input slot
        for (uint32_t c0 = 0; c0 < cnt; c0 += TB) {
            const uint32_t r = c0 + tid;
            const bool act = r < cnt;
            const uint32_t o = beg + r;
            const uint32_t s = act ? a.src[o] : 0u;
            float x[3] = {0.f, 0.f, 0.f}, F[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
            float ax[3] = {0.f, 0.f, 0.f}, aF[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            float av[3] = {0.f, 0.f, 0.f}, aC[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (act) {
                load_xF(a.S, a.cap, s, x, F);
                load_adj(a.abar, a.cap, o, ax, aF, av, aC);
            }
            Stencil st;
            make_stencil(x, g.o, g.inv_dx, st);
            float dw[3][3];
            weight_grads(st, dw);
            const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
            const bool inwin = act && l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
            // recompute C' = 4/dx^2 sum w v_i d^T
            float nc[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float w = st.w[0][ii] * st.w[1][jj] * st.w[2][kk];
                        const float d[3] = {((float)ii - st.f[0]) * g.dx, ((float)jj - st.f[1]) * g.dx,
                                            ((float)kk - st.f[2]) * g.dx};
                        float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (inwin) {
                            u = sv[((l2 + kk) * WN + (l1 + jj)) * WN + (l0 + ii)];
                        } else if (act) {
                            const int gi = st.b[0] + ii, gj = st.b[1] + jj, gk = st.b[2] + kk;
                            if (in_grid(g, gi, gj, gk)) u = gvs[node_id(g, gi, gj, gk)];
                        }
                        const float uv[3] = {u.x, u.y, u.z};
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int e = 0; e < 3; ++e) nc[3 * c + e] += w * uv[c] * d[e];
                    }
#pragma unroll
            for (int c = 0; c < 9; ++c) nc[c] *= s4;
            // x' = x + dt v', F' = (I + dt C') F
            float anv[3], anc[9], aFo[9];
#pragma unroll
            for (int c = 0; c < 3; ++c) anv[c] = av[c] + g.dt * ax[c];
#pragma unroll
            for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    anc[3 * ii + jj] = aC[3 * ii + jj] + g.dt * (aF[3 * ii] * F[3 * jj] + aF[3 * ii + 1] * F[3 * jj + 1] +
                                                                 aF[3 * ii + 2] * F[3 * jj + 2]);
                    aFo[3 * ii + jj] = aF[3 * ii + jj] + g.dt * (nc[ii] * aF[jj] + nc[3 + ii] * aF[3 + jj] +
                                                                 nc[6 + ii] * aF[6 + jj]);
                }
            float af[3] = {0.f, 0.f, 0.f};
            const int base = inwin ? (l2 * WN + l1) * WN + l0 : -1;
#pragma unroll
            for (int ii = 0; ii < 3; ++ii) {
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float wx = st.w[0][ii], wy = st.w[1][jj], wz = st.w[2][kk];
                        const float w = wx * wy * wz;
                        const float d[3] = {((float)ii - st.f[0]) * g.dx, ((float)jj - st.f[1]) * g.dx,
                                            ((float)kk - st.f[2]) * g.dx};
                        float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
                        bool ok = false;
                        int gi = 0, gj = 0, gk = 0;
                        if (inwin) {
                            u = sv[((l2 + kk) * WN + (l1 + jj)) * WN + (l0 + ii)];
                            ok = true;
                        } else if (act) {
                            gi = st.b[0] + ii; gj = st.b[1] + jj; gk = st.b[2] + kk;
                            ok = in_grid(g, gi, gj, gk);
                            if (ok) u = gvs[node_id(g, gi, gj, gk)];
                        }
                        const float uv[3] = {u.x, u.y, u.z};
                        float g3[3] = {0.f, 0.f, 0.f};
                        if (ok) {
                            float Cd[3];
#pragma unroll
                            for (int c = 0; c < 3; ++c)
                                Cd[c] = anc[3 * c] * d[0] + anc[3 * c + 1] * d[1] + anc[3 * c + 2] * d[2];
                            float aw = 0.f;
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                g3[c] = w * (anv[c] + s4 * Cd[c]);
                                aw += uv[c] * (anv[c] + s4 * Cd[c]);
                            }
#pragma unroll
                            for (int e = 0; e < 3; ++e)
                                af[e] -= g.dx * s4 * w * (anc[e] * uv[0] + anc[3 + e] * uv[1] + anc[6 + e] * uv[2]);
                            af[0] += aw * dw[0][ii] * wy * wz;
                            af[1] += aw * wx * dw[1][jj] * wz;
                            af[2] += aw * wx * wy * dw[2][kk];
                            if (!inwin) {
                                float* cp = reinterpret_cast<float*>(agvs + node_id(g, gi, gj, gk));
                                atomic_add_f32(cp + 0, g3[0]);
                                atomic_add_f32(cp + 1, g3[1]);
                                atomic_add_f32(cp + 2, g3[2]);
                                g3[0] = g3[1] = g3[2] = 0.f;
                            }
                        }
                        float* row = &myscr[lane * NP + 3 * (3 * jj + kk)];
                        row[0] = g3[0]; row[1] = g3[1]; row[2] = g3[2];
                    }
                __builtin_amdgcn_wave_barrier();
                wave_slice_reduce<3>(myscr, base, ii, sa);
                __builtin_amdgcn_wave_barrier();
            }
            if (act) {
#pragma unroll
                for (int c = 0; c < 3; ++c) ax[c] += af[c] * g.inv_dx;
                a.part[s] = make_float4(ax[0], ax[1], ax[2], aFo[0]);
                a.part[a.cap + s] = make_float4(aFo[1], aFo[2], aFo[3], aFo[4]);
                a.part[2 * a.cap + s] = make_float4(aFo[5], aFo[6], aFo[7], aFo[8]);
            }
        }
        __syncthreads();
        for (int l = tid; l < WV; l += TB) {
            int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
            float s0 = sa[l], s1 = sa[WV + l], s2 = sa[2 * WV + l];
            if ((s0 != 0.f || s1 != 0.f || s2 != 0.f) && gi < g.nx && gj < g.ny && gk < g.nz) {
                float* cp = reinterpret_cast<float*>(agvs + node_id(g, gi, gj, gk));
                atomic_add_f32(cp + 0, s0);
                atomic_add_f32(cp + 1, s1);
                atomic_add_f32(cp + 2, s2);
            }
        }
        __syncthreads();
    }
}

// reverse of v_i = (p_i/m_i + dt g) * mask_i
__global__ void __launch_bounds__(512) k_adj_grid(const SceneInfo* __restrict__ scenes, int nscenes,
                                                  float4* __restrict__ agv_all, float4* __restrict__ apam_all,
                                                  float4* __restrict__ mp_all, uint32_t* __restrict__ flags) {
    const int b = blockIdx.x;
    if (flags[b] == 0u) return;
    const SceneInfo& sc = scenes[scene_of(scenes, nscenes, (uint32_t)b, [](const SceneInfo& s) { return s.boff; })];
    const MlsGeom g = sc.g;
    float4* __restrict__ agv = agv_all + sc.goff;
    float4* __restrict__ apam = apam_all + sc.goff;
    float4* __restrict__ mp = mp_all + sc.goff;
    const int lb = b - (int)sc.boff;
    const int bx = lb % sc.gbx, by = (lb / sc.gbx) % sc.gby, bz = lb / (sc.gbx * sc.gby);
    const int i = bx * GB + (threadIdx.x & 7), j = by * GB + ((threadIdx.x >> 3) & 7), k = bz * GB + (threadIdx.x >> 6);
    if (i < g.nx && j < g.ny && k < g.nz) {
        const size_t id = node_id(g, i, j, k);
        const float4 q = mp[id];
        const float4 av = agv[id];
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        if (q.x > 0.f && !is_sticky(g, i, j, k)) {
            const float im = 1.f / q.x;
            r.y = av.x * im;
            r.z = av.y * im;
            r.w = av.z * im;
            r.x = -(av.x * q.y + av.y * q.z + av.z * q.w) * im * im;
        }
        apam[id] = r;
        agv[id] = make_float4(0.f, 0.f, 0.f, 0.f);
        mp[id] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    if (threadIdx.x == 0) flags[b] = 0u;
}

__global__ void __launch_bounds__(TB) k_adj_p2g(AdjArgs a) {
    __shared__ float sp[4][WV];
    const XcdChunk ch(*a.nactive);
    const int tid = threadIdx.x;
    for (uint32_t ti = ch.first; ti < ch.end; ti += ch.stride) {
        const uint4 te = a.alist[ti];
        const SceneInfo sc = a.scenes[te.w];   // by value: kept in registers across the stores
        const MlsGeom g = sc.g;
        const float4* __restrict__ apams = a.apam + sc.goff;
        const float s4 = 4.f * g.inv_dx * g.inv_dx;
        const float kscale = -g.dt * g.p_vol * s4;
        const float emu = 1.f / (2.f * (1.f + g.nu));
        const float elam = g.nu / ((1.f + g.nu) * (1.f - 2.f * g.nu));
        const uint32_t beg = te.y, cnt = te.z;
        int t0[3];
        tile_origin(te.x - sc.toff, t0);
        for (int l = tid; l < WV; l += TB) {
            int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
            float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gi < g.nx && gj < g.ny && gk < g.nz) u = apams[node_id(g, gi, gj, gk)];
            sp[0][l] = u.x; sp[1][l] = u.y; sp[2][l] = u.z; sp[3][l] = u.w;
        }
        __syncthreads();
        for (uint32_t r = tid; r < cnt; r += TB) {
            const uint32_t i = beg + r;
            const uint32_t s = a.perm[i];
            float x[3], F[9], v[3], C[9];
            load_state(a.S, a.cap, s, x, F, v, C);
            const float Eq = a.E ? a.E[s] : g.E;
            const float mq = a.mass ? a.mass[s] : g.p_mass;
            const float mu = Eq * emu, lam = Eq * elam;
            float tau[9], A[9];
            kirchhoff(g.material, mu, lam, F, tau);
#pragma unroll
            for (int c = 0; c < 9; ++c) A[c] = kscale * tau[c] + mq * C[c];
            Stencil st;
            make_stencil(x, g.o, g.inv_dx, st);
            float dw[3][3];
            weight_grads(st, dw);
            const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
            const bool inwin = l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
            float aA[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            float af[3] = {0.f, 0.f, 0.f}, avq[3] = {0.f, 0.f, 0.f};
            for (int ii = 0; ii < 3; ++ii)
                for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float wx = st.w[0][ii], wy = st.w[1][jj], wz = st.w[2][kk];
                        const float w = wx * wy * wz;
                        float d[3] = {((float)ii - st.f[0]) * g.dx, ((float)jj - st.f[1]) * g.dx,
                                      ((float)kk - st.f[2]) * g.dx};
                        float am, ap[3];
                        if (inwin) {
                            int l = ((l2 + kk) * WN + (l1 + jj)) * WN + (l0 + ii);
                            am = sp[0][l]; ap[0] = sp[1][l]; ap[1] = sp[2][l]; ap[2] = sp[3][l];
                        } else {
                            int gi = st.b[0] + ii, gj = st.b[1] + jj, gk = st.b[2] + kk;
                            if (!in_grid(g, gi, gj, gk)) continue;
                            float4 q = apams[node_id(g, gi, gj, gk)];
                            am = q.x; ap[0] = q.y; ap[1] = q.z; ap[2] = q.w;
                        }
                        float aw = am * mq;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const float contrib = mq * v[c] + A[3 * c] * d[0] + A[3 * c + 1] * d[1] + A[3 * c + 2] * d[2];
                            aw += ap[c] * contrib;
                            avq[c] += w * mq * ap[c];
#pragma unroll
                            for (int e = 0; e < 3; ++e) aA[3 * c + e] += w * ap[c] * d[e];
                        }
#pragma unroll
                        for (int e = 0; e < 3; ++e) af[e] -= g.dx * w * (A[e] * ap[0] + A[3 + e] * ap[1] + A[6 + e] * ap[2]);
                        af[0] += aw * dw[0][ii] * wy * wz;
                        af[1] += aw * wx * dw[1][jj] * wz;
                        af[2] += aw * wx * wy * dw[2][kk];
                    }
            // partial adjoints from A1 (ax, aF after the G2P reverse)
            const float4 p0 = a.part[s], p1 = a.part[a.cap + s], p2 = a.part[2 * a.cap + s];
            float ax[3] = {p0.x + af[0] * g.inv_dx, p0.y + af[1] * g.inv_dx, p0.z + af[2] * g.inv_dx};
            float aF[9] = {p0.w, p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, p2.w};
            float at[9], aC[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                at[c] = kscale * aA[c];
                aC[c] = mq * aA[c];
            }
            float amu = 0.f, alam = 0.f;
            kirchhoff_vjp(g.material, mu, lam, F, at, aF, amu, alam);
            store_state(a.abar, a.cap, s, ax, aF, avq, aC);
            if (a.aE_user) a.aE_user[a.pid[s]] += amu * emu + alam * elam;
        }
        __syncthreads();
    }
}

// user order <-> storage order conversions of the adjoint planes
__global__ void k_gather_user(int n, size_t cap, const uint32_t* __restrict__ pid, const float* __restrict__ ax,
                              const float* __restrict__ av, const float* __restrict__ aC, const float* __restrict__ aF,
                              float4* __restrict__ abar) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t u = pid[s];
    float x[3], v[3], C[9], F[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        x[c] = ax ? ax[3 * u + c] : 0.f;
        v[c] = av ? av[3 * u + c] : 0.f;
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        C[c] = aC ? aC[9 * u + c] : 0.f;
        F[c] = aF ? aF[9 * u + c] : 0.f;
    }
    store_state(abar, cap, s, x, F, v, C);
}

__global__ void k_scatter_user(int n, size_t cap, const uint32_t* __restrict__ pid, const float4* __restrict__ abar,
                               float* __restrict__ ax, float* __restrict__ av, float* __restrict__ aC,
                               float* __restrict__ aF) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t u = pid[s];
    float x[3], v[3], C[9], F[9];
    load_state(abar, cap, s, x, F, v, C);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ax[3 * u + c] = x[c];
        av[3 * u + c] = v[c];
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        aC[9 * u + c] = C[c];
        aF[9 * u + c] = F[c];
    }
}

}  // namespace

hipError_t adjoint_segment(const TiledWork& w, int n, size_t cap, TapeStep* tape, int L, float4* abar,
                           float4* part, float4* agv, float4* apam, float* aE_user, hipStream_t st) {
    (void)n;
    const int nblk = w.total_blocks;
    for (int s = L - 1; s >= 0; --s) {
        TapeStep& t = tape[s];
        AdjArgs a;
        a.S = t.S; a.pid = t.pid; a.E = t.E; a.mass = t.mass;
        a.perm = t.perm;
        a.src = t.src;
        a.alist = reinterpret_cast<const uint4*>(t.alist);
        a.nactive = t.nactive;
        a.gv = t.gv;
        a.abar = abar; a.part = part; a.agv = agv; a.apam = apam;
        a.aE_user = aE_user;
        a.scenes = w.scenes;
        a.cap = cap;
        {
            ProfScope ps(w.prof, PK_ADJ_G2P, st);
            k_adj_g2p<<<w.nblocks_launch, TB, 0, st>>>(a);
        }
        {
            ProfScope ps(w.prof, PK_ADJ_GRID, st);
            k_adj_grid<<<nblk, 512, 0, st>>>(w.scenes, w.nscenes, agv, apam, t.mp, t.flags);
        }
        {
            ProfScope ps(w.prof, PK_ADJ_P2G, st);
            k_adj_p2g<<<w.nblocks_launch, TB, 0, st>>>(a);
        }
    }
    return hipGetLastError();
}

hipError_t adjoint_from_user(int n, size_t cap, const uint32_t* pid, const float* ax, const float* av,
                             const float* aC, const float* aF, float4* abar, hipStream_t st) {
    if (n > 0) k_gather_user<<<(n + 255) / 256, 256, 0, st>>>(n, cap, pid, ax, av, aC, aF, abar);
    return hipGetLastError();
}

hipError_t adjoint_to_user(int n, size_t cap, const uint32_t* pid, const float4* abar, float* ax, float* av,
                           float* aC, float* aF, hipStream_t st) {
    if (n > 0) k_scatter_user<<<(n + 255) / 256, 256, 0, st>>>(n, cap, pid, abar, ax, av, aC, aF);
    return hipGetLastError();
}

}  // namespace pg
