// Reverse (adjoint) step of the MLS-MPM transfer loop on the tiled layout.
//
// Replaces the empty backward of DifferentiableMPMSystem
// (src/differentiable_mpm.h:823-878, tape lambdas that do nothing) and the
// gravity-only launch_mpm_timestep_backward (src/pytorch_autograd.cu:385-402)
// with the exact reverse of one step: G2P, grid update (incl. sticky-wall
// masks) and P2G with the elastic stress, differentiated by hand.  Works on a
// segment tape written by tiled_tape_forward; for step n of the tape:
//
//   A1  k_adj_g2p   tiles of binning n, particles in the forward tile
//                   kernel's groups (contiguous output offsets): with the
//                   forward's v', C' (the new state on the tape) and the gv
//                   window in LDS, push the adjoint of the new (v, C) into the
//                   wave's LDS windows by read-modify-write (summed into the
//                   tile's window buffer), accumulate the position adjoint
//                   from the weights and x_i - x_p; partial = (ax, aF) by
//                   output position (coalesced), and dst = the storage slot ->
//                   output position map of the step.
//   A2  k_adj_grid  flagged 8^3 blocks: the node's velocity adjoint is the sum
//                   of the window buffers that cover it plus agv; ap = agv/m,
//                   am = -(agv.p)/m^2 (zero on sticky or empty nodes); clears
//                   agv, the tape's accumulator and flags.
//   A3  k_adj_p2g   tiles of binning n: gather (am, ap) from the LDS window,
//                   adjoints of v, C, F (stress VJP incl. polar derivative),
//                   x, and dL/dE (carried by storage slot like the state
//                   adjoint when the caller asks for it).
//
// Indexing: the adjoint of the state after step n is stored by storage slot
// of step n+1, which is the output position o of step n (src_n[o] is the
// particle's slot in step n); A3 writes the adjoint of the state before step n
// at its slot of step n, which is what A1 of step n-1 reads by its output
// position.  A3 reads A1's partials at dst[slot].
//
// A segment of L steps runs A1(L-1), A2(L-1), then for n = L-1 .. 1 the fused
// kernel A3(n) + A1(n-1) (k_adj_g2p<true>: per particle, in binning n-1's
// groups, the adjoint of the state between the two steps stays in registers)
// and A2(n-1), and last A3(0).  The partials, dst and the dL/dE carry
// alternate between two buffer sets, since the fused kernel reads step n's
// while it writes step n-1's.
#include <algorithm>

#include "pg_tile.h"

namespace pg {
namespace {

struct AdjArgs {
    const float4* S;        // tape state of step n
    const float4* Snext;    // tape state of step n+1 (by its slots: step n's output positions)
    const uint32_t* pid;    // slot -> user index (step n)
    const float* E;         // per-particle E, slot order (or NULL)
    const float* mass;      // per-particle mass, slot order (or NULL)
    const uint32_t* perm;   // sorted position of binning n -> slot of step n
    const uint32_t* src;    // output position of step n -> slot of step n
    const uint4* alist;
    const uint32_t* nactive;
    const float4* gv;       // grid (v, m) of step n
    float4* abar;           // adjoint state planes
    const uint32_t* in_pid; // with in_inv: A1 (not fused) reads the adjoint of output position o at slot
    const uint32_t* in_inv; // in_inv[in_pid[o]] (it is stored in another order of the same particles)
    float4* part;           // partial adjoint planes (3*cap) by slot of step n: (ax, aF)
    float4* agv;            // grid accumulator of the velocity adjoint (x,y,z used)
    const float4* apam;     // (am, ap) per node
    float* aEs;             // dL/dE accumulated over the later steps, like abar (or NULL: not wanted)
    float* aEp;             // aEs carried from A1 to A3, by output position of step n
    uint32_t* dst;          // storage slot of step n -> output position of step n (A1 -> A3)
    float4* wbuf;           // per-tile window sums of the grid adjoint (by index in alist)
    const SceneInfo* scenes;
    size_t cap;
    MatTable mt;            // per-particle materials (recs NULL: the scene's)
};

// The step after the one A1 reverses, for the fused kernel (A3 of step n+1
// followed by A1 of step n on each particle): its slots are step n's output
// positions, so the particle A1 of step n takes at output position o is the
// one in slot o of step n+1.
struct AdjNext {
    const float4* S;        // tape state of step n+1
    const uint32_t* pid;
    const float* E;
    const float* mass;
    const float4* apam;     // (am, ap) per node of step n+1
    const float4* part;     // partials of step n+1 by its output position
    const uint32_t* dst;    // slot of step n+1 -> its output position
    const float* aEp;       // dL/dE carry of step n+1 by its output position
};

PG_DEV void load_adj(const float4* __restrict__ A, size_t cap, size_t i, float* ax, float* aF, float* av, float* aC) {
    load_state(A, cap, i, ax, aF, av, aC);
}

// x, v, C of a particle from the state planes, and F
PG_DEV void load_xvC(const float4* __restrict__ S, size_t cap, size_t i, float* x, float* v, float* C) {
    const float4 q0 = S[i], q3 = S[3 * cap + i], q4 = S[4 * cap + i], q5 = S[5 * cap + i];
    x[0] = q0.x; x[1] = q0.y; x[2] = q0.z;
    v[0] = q3.x; v[1] = q3.y; v[2] = q3.z; C[0] = q3.w;
    C[1] = q4.x; C[2] = q4.y; C[3] = q4.z; C[4] = q4.w;
    C[5] = q5.x; C[6] = q5.y; C[7] = q5.z; C[8] = q5.w;
}
PG_DEV void load_F(const float4* __restrict__ S, size_t cap, size_t i, float* F) {
    const float4 q0 = S[i], q1 = S[cap + i], q2 = S[2 * cap + i];
    F[0] = q0.w; F[1] = q1.x; F[2] = q1.y; F[3] = q1.z; F[4] = q1.w;
    F[5] = q2.x; F[6] = q2.y; F[7] = q2.z; F[8] = q2.w;
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

constexpr int APASS = 1280;   // particles per LDS pass of a work item (larger work items take several)

// The reverse of one particle's G2P and state update, given its stencil, F and
// the adjoints of its new state (ax, aF, av, aC):
//   C' = 4/dx sum_n w u_n (n - f)^T,  x' = x + dt v',  F' = (I + dt C') F
// emits the velocity adjoint g_n = w_n (anv + 4/dx^2 anc d_n) of each stencil
// node (d_n = (n - f) dx) and returns the adjoints of x and F.  With VC the
// caller passes the forward's v', C' (vq, Cq: the new state holds exactly what
// the forward's G2P summed), else they are summed again from the nodes.
__device__ __forceinline__ float pick3(const float* w, int i) { return i == 0 ? w[0] : (i == 1 ? w[1] : w[2]); }

template <bool ROLLED, bool VC, class Load, class Emit>
__device__ __forceinline__ void adj_g2p_particle(Load load, Emit emit, const MlsGeom& g, const Stencil& st,
                                                 const float* F, const float* ax, const float* aF,
                                                 const float* av, const float* aC, const float* vq, const float* Cq,
                                                 float* axo, float* aFo) {
    const float s4 = 4.f * g.inv_dx * g.inv_dx;
    // adjoints of v' and C' (through x' = x + dt v' and F' = (I + dt C') F); they do
    // not depend on C' itself
    float anv[3], anc[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) anv[c] = av[c] + g.dt * ax[c];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            anc[3 * i + j] = aC[3 * i + j] + g.dt * (aF[3 * i] * F[3 * j] + aF[3 * i + 1] * F[3 * j + 1] +
                                                     aF[3 * i + 2] * F[3 * j + 2]);
    float dw[3][3];
    weight_grads(st, dw);
    float nc[9], af[3] = {0.f, 0.f, 0.f};
    if (ROLLED) {
        // rare path: two rolled passes, stencil values picked by selects
#pragma unroll
        for (int c = 0; c < 9; ++c) nc[c] = VC ? Cq[c] : 0.f;
#pragma unroll 1
        for (int n = 0; n < (VC ? 0 : 27); ++n) {
            const int ix = n % 3, iy = (n / 3) % 3, iz = n / 9;
            const float4 u = load(ix, iy, iz);
            const float w = pick3(st.w[0], ix) * pick3(st.w[1], iy) * pick3(st.w[2], iz);
            const float d[3] = {(float)ix - st.f[0], (float)iy - st.f[1], (float)iz - st.f[2]};
            const float uu[3] = {u.x, u.y, u.z};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int e = 0; e < 3; ++e) nc[3 * c + e] += 4.f * g.inv_dx * w * uu[c] * d[e];
        }
#pragma unroll 1
        for (int n = 0; n < 27; ++n) {
            const int ix = n % 3, iy = (n / 3) % 3, iz = n / 9;
            const float wx = pick3(st.w[0], ix), wy = pick3(st.w[1], iy), wz = pick3(st.w[2], iz);
            const float w = wx * wy * wz;
            const float d[3] = {((float)ix - st.f[0]) * g.dx, ((float)iy - st.f[1]) * g.dx,
                                ((float)iz - st.f[2]) * g.dx};
            const float4 u4 = load(ix, iy, iz);
            const float u[3] = {u4.x, u4.y, u4.z};
            float t[3], aw = 0.f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                t[c] = anv[c] + s4 * (anc[3 * c] * d[0] + anc[3 * c + 1] * d[1] + anc[3 * c + 2] * d[2]);
                aw += u[c] * t[c];
            }
#pragma unroll
            for (int e = 0; e < 3; ++e)
                af[e] -= g.dx * s4 * w * (anc[e] * u[0] + anc[3 + e] * u[1] + anc[6 + e] * u[2]);
            af[0] += aw * pick3(dw[0], ix) * wy * wz;
            af[1] += aw * wx * pick3(dw[1], iy) * wz;
            af[2] += aw * wx * wy * pick3(dw[2], iz);
            emit(ix, iy, iz, make_float4(w * t[0], w * t[1], w * t[2], 0.f));
        }
    } else {
        // one pass over the nodes (rows as a loop, x unrolled): the recomputed
        // v' = sum w u and sum w u (n - f)^T, the emitted w_n t_n with
        // t_n = anv + s4 anc d_n built by additions along the axes, and
        // ga[a] = sum (u_n . t_n) dw_n/df_a
        float t0[3], tdx[3], tdy[3], tdz[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            t0[c] = anv[c] - s4 * g.dx * (anc[3 * c] * st.f[0] + anc[3 * c + 1] * st.f[1] + anc[3 * c + 2] * st.f[2]);
            tdx[c] = s4 * g.dx * anc[3 * c];
            tdy[c] = s4 * g.dx * anc[3 * c + 1];
            tdz[c] = s4 * g.dx * anc[3 * c + 2];
        }
        float v[3] = {0.f, 0.f, 0.f}, cx[3] = {0.f, 0.f, 0.f}, cy[3] = {0.f, 0.f, 0.f}, cz[3] = {0.f, 0.f, 0.f};
        float ga[3] = {0.f, 0.f, 0.f};
        float tz[3] = {t0[0], t0[1], t0[2]};
#pragma unroll 1
        for (int z = 0; z < 3; ++z) {
            float R[3] = {0.f, 0.f, 0.f}, Ry[3] = {0.f, 0.f, 0.f}, Rx[3] = {0.f, 0.f, 0.f};
            float Rw = 0.f, R1 = 0.f, R0 = 0.f;
            float ty[3] = {tz[0], tz[1], tz[2]};
            const float wz = pick3(st.w[2], z);
#pragma unroll 1
            for (int y = 0; y < 3; ++y) {
                float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f};
                float Pw = 0.f, Pd = 0.f;
                float tx[3] = {ty[0], ty[1], ty[2]};
                const float wy = pick3(st.w[1], y), wyz = wy * wz;
                // the row's three velocities are read before its three emissions, whose
                // read-modify-writes are ordered after one another
                const float4 ur[3] = {load(0, y, z), load(1, y, z), load(2, y, z)};
                float4 er[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const float u[3] = {ur[x].x, ur[x].y, ur[x].z};
                    const float wx = st.w[0][x], wdx = wx * ((float)x - st.f[0]);
                    if (!VC)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            P[c] += wx * u[c];
                            Q[c] += wdx * u[c];
                        }
                    const float aw = u[0] * tx[0] + u[1] * tx[1] + u[2] * tx[2];
                    Pw += wx * aw;
                    Pd += dw[0][x] * aw;
                    const float w = wx * wyz;
                    er[x] = make_float4(w * tx[0], w * tx[1], w * tx[2], 0.f);
                    if (x < 2)
#pragma unroll
                        for (int c = 0; c < 3; ++c) tx[c] += tdx[c];
                }
#pragma unroll
                for (int x = 0; x < 3; ++x) emit(x, y, z, er[x]);
                const float wdy = wy * ((float)y - st.f[1]), dwy = pick3(dw[1], y);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (!VC) {
                        R[c] += wy * P[c];
                        Ry[c] += wdy * P[c];
                        Rx[c] += wy * Q[c];
                    }
                    ty[c] += tdy[c];
                }
                Rw += wy * Pw;
                R1 += dwy * Pw;
                R0 += wy * Pd;
            }
            const float wdz = wz * ((float)z - st.f[2]), dwz = pick3(dw[2], z);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (!VC) {
                    v[c] += wz * R[c];
                    cz[c] += wdz * R[c];
                    cy[c] += wz * Ry[c];
                    cx[c] += wz * Rx[c];
                }
                tz[c] += tdz[c];
            }
            ga[2] += dwz * Rw;
            ga[1] += wz * R1;
            ga[0] += wz * R0;
        }
        const float s = 4.f * g.inv_dx;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (VC) v[c] = vq[c];
            nc[3 * c + 0] = VC ? Cq[3 * c + 0] : s * cx[c];
            nc[3 * c + 1] = VC ? Cq[3 * c + 1] : s * cy[c];
            nc[3 * c + 2] = VC ? Cq[3 * c + 2] : s * cz[c];
        }
        // d_n = x_n - x_p also moves with x: - s4 dx anc^T sum w u
#pragma unroll
        for (int e = 0; e < 3; ++e) af[e] = ga[e] - g.dx * s4 * (anc[e] * v[0] + anc[3 + e] * v[1] + anc[6 + e] * v[2]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            aFo[3 * i + j] = aF[3 * i + j] + g.dt * (nc[i] * aF[j] + nc[3 + i] * aF[3 + j] + nc[6 + i] * aF[6 + j]);
#pragma unroll
    for (int c = 0; c < 3; ++c) axo[c] = ax[c] + af[c] * g.inv_dx;
}

// A3 for one particle: the reverse of its P2G (stress included) given the
// grid adjoints (am, ap) of its stencil nodes (load(x, y, z) -> float4, zero
// outside the grid) and its partial adjoints (ax, aF) from A1 (read by part()
// after the node loop, so they are not live across it).  The forward put
// w_n (m, b_n) on node n with b_n = m v + A d_n, d_n = (n - f) dx; with
// aw_n = m am_n + ap_n . b_n the adjoints are
//   v: m sum w ap,  A: dx sum w ap (n - f)^T,
//   x (times dx): sum aw_n dw_n/df - dx A^T sum w ap.
// The unrolled form factors the sums along the axes (x innermost) like the
// forward G2P; ROLLED is the small form for the rare particle whose stencil
// is not in the tile's window.  getF(F) reads the particle's F; it is read
// again after the node loop rather than kept live across it.
// Returns the adjoints of x, F, v, C before the step and dL/dE.
template <bool ROLLED, class Load, class Part, class GetF>
__device__ __forceinline__ void adj_p2g_particle(Load load, Part part, GetF getF, const MlsGeom& g,
                                                 const Stencil& st, const float* v, const float* C, const PMat& pm,
                                                 float mq, float* ax, float* aF, float* avq, float* aC, float& dE) {
    const float s4 = 4.f * g.inv_dx * g.inv_dx;
    const float kscale = -g.dt * pm.vol * s4;
    float A[9];
    {
        float F[9], tau[9];
        getF(F);
        stress_of(pm, F, C, tau);
#pragma unroll
        for (int c = 0; c < 9; ++c) A[c] = kscale * tau[c] + mq * C[c];
    }
    float dw[3][3];
    weight_grads(st, dw);
    // b_n = q + dx (A[:, 0] x + A[:, 1] y + A[:, 2] z)
    float q[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        q[c] = mq * v[c] - g.dx * (A[3 * c] * st.f[0] + A[3 * c + 1] * st.f[1] + A[3 * c + 2] * st.f[2]);
    // s0 = sum w ap, s1[c][e] = sum w ap[c] (n - f)[e], ga[a] = sum aw dw/df_a
    float s0[3] = {0.f, 0.f, 0.f}, ga[3] = {0.f, 0.f, 0.f};
    float s1[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (ROLLED) {
#pragma unroll 1
        for (int n = 0; n < 27; ++n) {
            const int ix = n % 3, iy = (n / 3) % 3, iz = n / 9;
            const float4 u = load(ix, iy, iz);
            const float wx = pick3(st.w[0], ix), wy = pick3(st.w[1], iy), wz = pick3(st.w[2], iz);
            const float w = wx * wy * wz;
            const float d[3] = {(float)ix - st.f[0], (float)iy - st.f[1], (float)iz - st.f[2]};
            const float ap[3] = {u.y, u.z, u.w};
            float aw = u.x * mq;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                aw += ap[c] * (q[c] + g.dx * (A[3 * c] * (float)ix + A[3 * c + 1] * (float)iy +
                                              A[3 * c + 2] * (float)iz));
                s0[c] += w * ap[c];
#pragma unroll
                for (int e = 0; e < 3; ++e) s1[c][e] += w * ap[c] * d[e];
            }
            ga[0] += aw * pick3(dw[0], ix) * wy * wz;
            ga[1] += aw * wx * pick3(dw[1], iy) * wz;
            ga[2] += aw * wx * wy * pick3(dw[2], iz);
        }
    } else {
        // rows (y, z) as a loop, nodes along x unrolled: bounds the live registers
        auto wd = [&](int a, int n) { return pick3(st.w[a], n) * ((float)n - st.f[a]); };
        auto dwf = [&](int a, int n) { return pick3(dw[a], n); };
        float bz[3] = {q[0], q[1], q[2]};
#pragma unroll 1
        for (int z = 0; z < 3; ++z) {
            float R[3] = {0.f, 0.f, 0.f}, Ry[3] = {0.f, 0.f, 0.f}, Rx[3] = {0.f, 0.f, 0.f};
            float Rw = 0.f, R1 = 0.f, R0 = 0.f;
            float by[3] = {bz[0], bz[1], bz[2]};
#pragma unroll 1
            for (int y = 0; y < 3; ++y) {
                float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f};
                float Pw = 0.f, Pd = 0.f;
                float bx[3] = {by[0], by[1], by[2]};
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const float4 u = load(x, y, z);
                    const float ap[3] = {u.y, u.z, u.w};
                    const float aw = u.x * mq + ap[0] * bx[0] + ap[1] * bx[1] + ap[2] * bx[2];
                    const float wx = st.w[0][x], wdx = wx * ((float)x - st.f[0]);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        P[c] += wx * ap[c];
                        Q[c] += wdx * ap[c];
                    }
                    Pw += wx * aw;
                    Pd += dw[0][x] * aw;
                    if (x < 2)
#pragma unroll
                        for (int c = 0; c < 3; ++c) bx[c] += g.dx * A[3 * c];
                }
                const float wy = pick3(st.w[1], y), wdy = wd(1, y), dwy = dwf(1, y);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    R[c] += wy * P[c];
                    Ry[c] += wdy * P[c];
                    Rx[c] += wy * Q[c];
                    by[c] += g.dx * A[3 * c + 1];
                }
                Rw += wy * Pw;
                R1 += dwy * Pw;
                R0 += wy * Pd;
            }
            const float wz = pick3(st.w[2], z), wdz = wd(2, z), dwz = dwf(2, z);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s0[c] += wz * R[c];
                s1[c][2] += wdz * R[c];
                s1[c][1] += wz * Ry[c];
                s1[c][0] += wz * Rx[c];
                bz[c] += g.dx * A[3 * c + 2];
            }
            ga[2] += dwz * Rw;
            ga[1] += wz * R1;
            ga[0] += wz * R0;
        }
    }
    float af[3], aA[9];
#pragma unroll
    for (int e = 0; e < 3; ++e) af[e] = ga[e] - g.dx * (A[e] * s0[0] + A[3 + e] * s0[1] + A[6 + e] * s0[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        avq[c] = mq * s0[c];
#pragma unroll
        for (int e = 0; e < 3; ++e) aA[3 * c + e] = g.dx * s1[c][e];
    }
    float4 p0, p1, p2;
    part(p0, p1, p2);
    ax[0] = p0.x + af[0] * g.inv_dx;
    ax[1] = p0.y + af[1] * g.inv_dx;
    ax[2] = p0.z + af[2] * g.inv_dx;
    aF[0] = p0.w; aF[1] = p1.x; aF[2] = p1.y; aF[3] = p1.z; aF[4] = p1.w;
    aF[5] = p2.x; aF[6] = p2.y; aF[7] = p2.z; aF[8] = p2.w;
    float at[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        at[c] = kscale * aA[c];
        aC[c] = mq * aA[c];
    }
    float F[9];
    getF(F);
    dE = 0.f;
    stress_vjp(pm, F, C, at, aF, aC, dE);
}

// A1: work items of binning n, in the forward tile kernel's groups.  The
// forward wrote the in-tile particles of a pass group by group: group grp of
// G = max(ceil(tot / 64), ceil(maxcount / 2)) holds ceil((tot - grp) / G)
// particles from output offset grp (tot div G) + min(grp, tot mod G), and the
// particles of a pass whose stencil base lies outside the tile (clamped keys)
// after them.  Counting the pass by stencil-base cell gives tot and maxcount
// again, so the reverse runs the same groups over contiguous output offsets.
// A group holds at most two particles of a cell, in adjacent lanes; the upper
// one scatters into the wave's second LDS window, so within a window the lanes
// touch distinct nodes and add with plain read-modify-writes.  The windows are
// summed into the tile's buffer; the reverse grid update gathers them.  An
// outside particle scatters with global atomics into agv.
//
// FUSED: each particle first runs A3 of step n+1 (nx) and takes the adjoints
// of its state after step n from that, in registers, instead of reading them
// from abar.  The (am, ap) of step n+1 come from an 8^3 window [t0 - 1,
// t0 + 7) of the tile of binning n, which holds the stencil of every particle
// that moved at most one cell out of the tile in step n; a particle further
// out reads them from global memory.
template <bool FUSED>
__global__ void __launch_bounds__(TB, 3) k_adj_g2p(AdjArgs a, AdjNext nx) {
    constexpr int PT = APASS / TB;
    static_assert(APASS % TB == 0, "the bulk phases cover the pass with whole thread strides");
    __shared__ float4 sv[WV];
    __shared__ float4 sa[FUSED ? PV : 1];     // (am, ap) window of step n+1
    __shared__ float4 pw[NW][2][WV];          // per-wave windows: lower and upper particle of a cell
    __shared__ uint8_t lcell[APASS];          // stencil-base cell by output offset (64: outside)
    __shared__ uint32_t lcnt[65];
    __shared__ uint32_t s_max, s_tot;
    const ItemStride ch(*a.nactive);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t ti = ch.first; ti < ch.end; ti += ch.stride) {
        const uint4 te = a.alist[ti];
        const SceneInfo sc = a.scenes[te.w];
        const MlsGeom g = sc.g;
        const float4* __restrict__ gvs = a.gv + sc.goff;
        float4* __restrict__ agvs = a.agv + sc.goff;
        const float4* __restrict__ apams = FUSED ? nx.apam + sc.goff : nullptr;
        const uint32_t beg = te.y, cnt = te.z;
        int t0[3];
        tile_origin(te.x - sc.toff, t0);
        if constexpr (FUSED) {
            for (int l = tid; l < PV; l += TB) {
                const int gi = t0[0] - 1 + l % PN, gj = t0[1] - 1 + (l / PN) % PN, gk = t0[2] - 1 + l / (PN * PN);
                sa[l] = in_grid(g, gi, gj, gk) ? apams[node_id(g, gi, gj, gk)] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        static_assert(WV <= TB, "one window node per thread");
        if (tid < WV) {
            const int gi = t0[0] + tid % WN, gj = t0[1] + (tid / WN) % WN, gk = t0[2] + tid / (WN * WN);
            float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gi < g.nx && gj < g.ny && gk < g.nz) u = gvs[node_id(g, gi, gj, gk)];
            sv[tid] = u;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                pw[w][0][tid] = make_float4(0.f, 0.f, 0.f, 0.f);
                pw[w][1][tid] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        for (uint32_t pb = 0; pb < cnt; pb += APASS) {
            const uint32_t pc = min((uint32_t)APASS, cnt - pb);
            if (tid < 65) lcnt[tid] = 0;
            __syncthreads();
            float4 k0[PT];
#pragma unroll
            for (int k = 0; k < PT; ++k) {
                const uint32_t r = tid + k * TB;
                if (r < pc) k0[k] = a.S[a.src[beg + pb + r]];
            }
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
                    atomicAdd(&lcnt[cell], 1u);
                    lcell[r] = (uint8_t)cell;
                }
            }
            __syncthreads();
            if (wave == 0) {
                uint32_t sum = lcnt[lane], mx = lcnt[lane];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    sum += (uint32_t)__shfl_xor(sum, off);
                    mx = max(mx, (uint32_t)__shfl_xor(mx, off));
                }
                if (lane == 0) {
                    s_tot = sum;
                    s_max = mx;
                }
            }
            __syncthreads();
            const uint32_t maxcnt = s_max, tot = s_tot;

            // r = output offset in the pass; own: a group particle (its base is a cell
            // of the tile); second: the upper particle of its cell in the group (the
            // wave's second window)
            auto particle = [&](uint32_t r, bool own, bool second) {
                const uint32_t o = beg + pb + r;
                float ax[3], aF[9], av[3], aC[9];
                if constexpr (FUSED) {
                    // A3 of step n+1 on the particle in its slot o
                    float xn[3], vn[3], Cn[9];
                    load_xvC(nx.S, a.cap, o, xn, vn, Cn);
                    const float mq = nx.mass ? nx.mass[o] : g.p_mass;
                    const PMat pm = particle_mat(g, a.mt, nx.pid[o], nx.E, o, mq);
                    Stencil sn;
                    make_stencil(xn, g.o, g.inv_dx, sn);
                    const int l0 = sn.b[0] - t0[0] + 1, l1 = sn.b[1] - t0[1] + 1, l2 = sn.b[2] - t0[2] + 1;
                    const bool inw = (unsigned)l0 < (unsigned)(PN - 2) && (unsigned)l1 < (unsigned)(PN - 2) &&
                                     (unsigned)l2 < (unsigned)(PN - 2);
                    const uint32_t q = nx.dst[o];
                    auto part = [&](float4& p0, float4& p1, float4& p2) {
                        p0 = nx.part[q]; p1 = nx.part[a.cap + q]; p2 = nx.part[2 * a.cap + q];
                    };
                    auto getF = [&](float* Fn) { load_F(nx.S, a.cap, o, Fn); };
                    float dE;
                    if (inw)
                        adj_p2g_particle<false>(
                            [&](int ii, int jj, int kk) { return sa[((l2 + kk) * PN + (l1 + jj)) * PN + (l0 + ii)]; },
                            part, getF, g, sn, vn, Cn, pm, mq, ax, aF, av, aC, dE);
                    else
                        adj_p2g_particle<true>(
                            [&](int ii, int jj, int kk) {
                                const int gi = sn.b[0] + ii, gj = sn.b[1] + jj, gk = sn.b[2] + kk;
                                return in_grid(g, gi, gj, gk) ? apams[node_id(g, gi, gj, gk)]
                                                              : make_float4(0.f, 0.f, 0.f, 0.f);
                            },
                            part, getF, g, sn, vn, Cn, pm, mq, ax, aF, av, aC, dE);
                    if (a.aEs) a.aEp[o] = nx.aEp[q] + dE;
                }
                const uint32_t s = a.src[o];
                float x[3];
                float F[9];
                {
                    const float4 q0 = a.S[s], q1 = a.S[a.cap + s], q2 = a.S[2 * a.cap + s];
                    x[0] = q0.x; x[1] = q0.y; x[2] = q0.z;
                    F[0] = q0.w; F[1] = q1.x; F[2] = q1.y; F[3] = q1.z; F[4] = q1.w;
                    F[5] = q2.x; F[6] = q2.y; F[7] = q2.z; F[8] = q2.w;
                }
                const uint32_t oin = (!FUSED && a.in_inv) ? a.in_inv[a.in_pid[o]] : o;
                if constexpr (!FUSED) load_adj(a.abar, a.cap, oin, ax, aF, av, aC);
                // the forward's G2P result v', C' from the new state
                float vq[3], Cq[9];
                {
                    const float4 q3 = a.Snext[3 * a.cap + o], q4 = a.Snext[4 * a.cap + o],
                                 q5 = a.Snext[5 * a.cap + o];
                    vq[0] = q3.x; vq[1] = q3.y; vq[2] = q3.z; Cq[0] = q3.w;
                    Cq[1] = q4.x; Cq[2] = q4.y; Cq[3] = q4.z; Cq[4] = q4.w;
                    Cq[5] = q5.x; Cq[6] = q5.y; Cq[7] = q5.z; Cq[8] = q5.w;
                }
                Stencil st;
                make_stencil(x, g.o, g.inv_dx, st);
                float axo[3], aFo[9];
                if (own) {
                    float4* __restrict__ win = pw[wave][second ? 1 : 0];
                    const int nb = ((st.b[2] - t0[2]) * WN + (st.b[1] - t0[1])) * WN + (st.b[0] - t0[0]);
                    adj_g2p_particle<false, true>([&](int ix, int iy, int iz) { return sv[nb + (iz * WN + iy) * WN + ix]; },
                                     [&](int ix, int iy, int iz, float4 c) {
                                         float4* t = win + nb + (iz * WN + iy) * WN + ix;
                                         float4 u = *t;
                                         u.x += c.x;
                                         u.y += c.y;
                                         u.z += c.z;
                                         *t = u;
                                         lds_order();
                                     },
                                     g, st, F, ax, aF, av, aC, vq, Cq, axo, aFo);
                } else {
                    adj_g2p_particle<true, true>(
                        [&](int ix, int iy, int iz) {
                            const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
                            return in_grid(g, gi, gj, gk) ? gvs[node_id(g, gi, gj, gk)]
                                                          : make_float4(0.f, 0.f, 0.f, 0.f);
                        },
                        [&](int ix, int iy, int iz, float4 c) {
                            const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
                            if (!in_grid(g, gi, gj, gk)) return;
                            float* cp = reinterpret_cast<float*>(agvs + node_id(g, gi, gj, gk));
                            atomic_add_f32(cp + 0, c.x);
                            atomic_add_f32(cp + 1, c.y);
                            atomic_add_f32(cp + 2, c.z);
                        },
                        g, st, F, ax, aF, av, aC, vq, Cq, axo, aFo);
                }
                // partials by output position (coalesced); A3 finds them through dst
                a.part[o] = make_float4(axo[0], axo[1], axo[2], aFo[0]);
                a.part[a.cap + o] = make_float4(aFo[1], aFo[2], aFo[3], aFo[4]);
                a.part[2 * a.cap + o] = make_float4(aFo[5], aFo[6], aFo[7], aFo[8]);
                a.dst[s] = o;
                if (!FUSED && a.aEs) a.aEp[o] = a.aEs[oin];
            };
            const uint32_t G = max((tot + 63u) / 64u, (maxcnt + 1u) / 2u);
            if (G > 0) {
                const uint32_t gq = tot / G, gr = tot % G;
                for (uint32_t grp = wave; grp < G; grp += NW) {
                    const uint32_t r = grp * gq + min(grp, gr) + (uint32_t)lane;
                    if ((uint32_t)lane < gq + (grp < gr ? 1u : 0u))
                        particle(r, true, lane > 0 && lcell[r - 1] == lcell[r]);
                }
            }
            for (uint32_t t = tid; t < pc - tot; t += TB) particle(tot + t, false, false);
            __syncthreads();
        }
        // the tile's velocity-adjoint window goes to its buffer (one per active-list entry)
        float4* __restrict__ wb = a.wbuf + (size_t)ti * PV;
        if (tid < WV) {
            float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int w = 0; w < NW; ++w)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    m.x += pw[w][h][tid].x;
                    m.y += pw[w][h][tid].y;
                    m.z += pw[w][h][tid].z;
                }
            wb[tid] = m;
        }
        __syncthreads();
    }
}

// A2: reverse of v_i = (p_i/m_i + dt g) * mask_i over the step's 8^3 blocks.
// The velocity adjoint of node i is the sum of the buffered windows of the
// tiles of binning n that cover it (the window of tile T spans nodes
// [4T, 4T + 6)) plus the global accumulator.  Clears the accumulator, the
// step's grid and flags.
__global__ void __launch_bounds__(512) k_adj_grid(const SceneInfo* __restrict__ scenes,
                                                  const uint32_t* __restrict__ bscene,
                                                  const uint32_t* __restrict__ tslot, const float4* __restrict__ wbuf,
                                                  float4* __restrict__ agv_all, float4* __restrict__ apam_all,
                                                  float4* __restrict__ mp_all, uint32_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ node_bc) {
    const int b = blockIdx.x;
    if (flags[b] == 0u) return;
    const SceneInfo sc = scenes[bscene[b]];
    const MlsGeom g = sc.g;
    float4* __restrict__ agv = agv_all + sc.goff;
    float4* __restrict__ apam = apam_all + sc.goff;
    float4* __restrict__ mp = mp_all + sc.goff;
    const int lb = b - (int)sc.boff;
    const int bx = lb % sc.gbx, by = (lb / sc.gbx) % sc.gby, bz = lb / (sc.gbx * sc.gby);
    const int n3[3] = {bx * GB + (int)(threadIdx.x & 7), by * GB + (int)((threadIdx.x >> 3) & 7),
                       bz * GB + (int)(threadIdx.x >> 6)};
    const int nt[3] = {sc.ntx, sc.nty, sc.ntz};
    if (n3[0] < g.nx && n3[1] < g.ny && n3[2] < g.nz) {
        const size_t id = node_id(g, n3[0], n3[1], n3[2]);
        float4 av = agv[id];
        // the entry ranges of the 8 tiles whose windows may cover the node, then
        // their first windows, as independent loads (two round trips); the
        // further windows of a tile split into several work items after that
        uint2 ts[8];
        int wo[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int T[3], wl[3];
            bool ok = true;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                T[ax] = n3[ax] / TT - ((e >> ax) & 1);
                wl[ax] = n3[ax] - TT * T[ax];
                ok = ok && T[ax] >= 0 && T[ax] < nt[ax] && wl[ax] < WN;
            }
            ts[e] = ok ? reinterpret_cast<const uint2*>(tslot)[sc.toff + morton_tile(T[0], T[1], T[2])]
                       : make_uint2(0u, 0u);
            wo[e] = (wl[2] * WN + wl[1]) * WN + wl[0];
        }
        float4 u[8];
#pragma unroll
        for (int e = 0; e < 8; ++e)
            u[e] = ts[e].y ? wbuf[(size_t)ts[e].x * PV + wo[e]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            av.x += u[e].x;
            av.y += u[e].y;
            av.z += u[e].z;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e)
            for (uint32_t c = 1; c < ts[e].y; ++c) {
                const float4 t = wbuf[(size_t)(ts[e].x + c) * PV + wo[e]];
                av.x += t.x;
                av.y += t.y;
                av.z += t.z;
            }
        const float4 q = mp[id];
        apam[id] = node_velocity_vjp(g, n3[0], n3[1], n3[2], node_bc ? node_bc[sc.goff + id] : 0u, q, av);
        agv[id] = make_float4(0.f, 0.f, 0.f, 0.f);
        mp[id] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    if (threadIdx.x == 0) flags[b] = 0u;
}

__global__ void __launch_bounds__(TB, 3) k_adj_p2g(AdjArgs a) {   // 3 waves per SIMD: latency-bound gather
    __shared__ float4 sp[WV];
    const ItemStride ch(*a.nactive);
    const int tid = threadIdx.x;
    for (uint32_t ti = ch.first; ti < ch.end; ti += ch.stride) {
        const uint4 te = a.alist[ti];
        const SceneInfo sc = a.scenes[te.w];   // by value: kept in registers across the stores
        const MlsGeom g = sc.g;
        const float4* __restrict__ apams = a.apam + sc.goff;
        const uint32_t beg = te.y, cnt = te.z;
        int t0[3];
        tile_origin(te.x - sc.toff, t0);
        for (int l = tid; l < WV; l += TB) {
            int gi = t0[0] + l % WN, gj = t0[1] + (l / WN) % WN, gk = t0[2] + l / (WN * WN);
            float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gi < g.nx && gj < g.ny && gk < g.nz) u = apams[node_id(g, gi, gj, gk)];
            sp[l] = u;
        }
        __syncthreads();
        // each thread loads its next particle's slot while it works on the current one
        uint32_t sn = (uint32_t)tid < cnt ? a.perm[beg + tid] : 0u;
        for (uint32_t r = tid; r < cnt; r += TB) {
            const uint32_t s = sn;
            if (r + TB < cnt) sn = a.perm[beg + r + TB];
            float x[3], v[3], C[9];
            load_xvC(a.S, a.cap, s, x, v, C);
            auto getF = [&](float* F) { load_F(a.S, a.cap, s, F); };
            const float mq = a.mass ? a.mass[s] : g.p_mass;
            const PMat pm = particle_mat(g, a.mt, a.pid[s], a.E, s, mq);
            Stencil st;
            make_stencil(x, g.o, g.inv_dx, st);
            const int l0 = st.b[0] - t0[0], l1 = st.b[1] - t0[1], l2 = st.b[2] - t0[2];
            const bool inwin = l0 >= 0 && l0 < TT && l1 >= 0 && l1 < TT && l2 >= 0 && l2 < TT;
            float ax[3], aF[9], avq[3], aC[9], dE;
            const uint32_t q = a.dst[s];
            auto part = [&](float4& p0, float4& p1, float4& p2) {
                p0 = a.part[q]; p1 = a.part[a.cap + q]; p2 = a.part[2 * a.cap + q];
            };
            if (inwin)
                adj_p2g_particle<false>(
                    [&](int ii, int jj, int kk) { return sp[((l2 + kk) * WN + (l1 + jj)) * WN + (l0 + ii)]; }, part,
                    getF, g, st, v, C, pm, mq, ax, aF, avq, aC, dE);
            else
                adj_p2g_particle<true>(
                    [&](int ii, int jj, int kk) {
                        const int gi = st.b[0] + ii, gj = st.b[1] + jj, gk = st.b[2] + kk;
                        return in_grid(g, gi, gj, gk) ? apams[node_id(g, gi, gj, gk)]
                                                      : make_float4(0.f, 0.f, 0.f, 0.f);
                    },
                    part, getF, g, st, v, C, pm, mq, ax, aF, avq, aC, dE);
            store_state(a.abar, a.cap, s, ax, aF, avq, aC);
            if (a.aEs) a.aEs[s] = a.aEp[q] + dE;
        }
        __syncthreads();
    }
}

// ---- the plain port of the reverse step: one thread per particle in user
// order, global atomics into a dense velocity-adjoint grid, dense grid pass.
// It is the baseline the tiled reverse is measured against.
__global__ void __launch_bounds__(256) k_adj_g2p_plain(int n, size_t cap, MlsGeom g, const float4* __restrict__ S,
                                                       const float4* __restrict__ gv, const float4* __restrict__ abar,
                                                       float4* __restrict__ part, float4* __restrict__ agv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[3], F[9], ax[3], aF[9], av[3], aC[9];
    load_xF(S, cap, i, x, F);
    load_adj(abar, cap, i, ax, aF, av, aC);
    Stencil st;
    make_stencil(x, g.o, g.inv_dx, st);
    float axo[3], aFo[9];
    adj_g2p_particle<false, false>(
        [&](int ix, int iy, int iz) {
            const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
            return in_grid(g, gi, gj, gk) ? gv[node_id(g, gi, gj, gk)] : make_float4(0.f, 0.f, 0.f, 0.f);
        },
        [&](int ix, int iy, int iz, float4 c) {
            const int gi = st.b[0] + ix, gj = st.b[1] + iy, gk = st.b[2] + iz;
            if (!in_grid(g, gi, gj, gk)) return;
            float* cp = reinterpret_cast<float*>(agv + node_id(g, gi, gj, gk));
            atomic_add_f32(cp + 0, c.x);
            atomic_add_f32(cp + 1, c.y);
            atomic_add_f32(cp + 2, c.z);
        },
        g, st, F, ax, aF, av, aC, nullptr, nullptr, axo, aFo);
    part[i] = make_float4(axo[0], axo[1], axo[2], aFo[0]);
    part[cap + i] = make_float4(aFo[1], aFo[2], aFo[3], aFo[4]);
    part[2 * cap + i] = make_float4(aFo[5], aFo[6], aFo[7], aFo[8]);
}

__global__ void k_adj_grid_plain(MlsGeom g, float4* __restrict__ agv, float4* __restrict__ apam,
                                 const float4* __restrict__ mp, const uint32_t* __restrict__ node_bc) {
    const size_t G = (size_t)g.nx * g.ny * g.nz;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < G; id += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(id % g.nx), j = (int)((id / g.nx) % g.ny), k = (int)(id / ((size_t)g.nx * g.ny));
        const float4 q = mp[id], av = agv[id];
        apam[id] = node_velocity_vjp(g, i, j, k, node_bc ? node_bc[id] : 0u, q, av);
        agv[id] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__global__ void __launch_bounds__(256) k_adj_p2g_plain(int n, size_t cap, MlsGeom g, const float4* __restrict__ S,
                                                       const float* __restrict__ E, const float* __restrict__ mass,
                                                       const float4* __restrict__ apam, const float4* __restrict__ part,
                                                       float4* __restrict__ abar, float* __restrict__ aE,
                                                       MatTable mt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[3], v[3], C[9];
    load_xvC(S, cap, i, x, v, C);
    const float mq = mass ? mass[i] : g.p_mass;
    const PMat pm = particle_mat(g, mt, (uint32_t)i, E, i, mq);
    Stencil st;
    make_stencil(x, g.o, g.inv_dx, st);
    float ax[3], aF[9], avq[3], aC[9], dE;
    adj_p2g_particle<false>(
        [&](int ii, int jj, int kk) {
            const int gi = st.b[0] + ii, gj = st.b[1] + jj, gk = st.b[2] + kk;
            return in_grid(g, gi, gj, gk) ? apam[node_id(g, gi, gj, gk)] : make_float4(0.f, 0.f, 0.f, 0.f);
        },
        [&](float4& p0, float4& p1, float4& p2) {
            p0 = part[i]; p1 = part[cap + i]; p2 = part[2 * cap + i];
        },
        [&](float* F) { load_F(S, cap, i, F); }, g, st, v, C, pm, mq, ax, aF, avq, aC, dE);
    store_state(abar, cap, i, ax, aF, avq, aC);
    if (aE) aE[i] += dE;
}

// user order <-> storage order conversions of the adjoint planes
__global__ void k_gather_user(int n, size_t cap, const uint32_t* __restrict__ pid, const float* __restrict__ ax,
                              const float* __restrict__ av, const float* __restrict__ aC, const float* __restrict__ aF,
                              float4* __restrict__ abar, const float* __restrict__ aE, float* __restrict__ aEs) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t u = pid ? pid[s] : (uint32_t)s;
    if (aEs) aEs[s] = aE[u];
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
                               float* __restrict__ aF, const float* __restrict__ aEs, float* __restrict__ aE) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t u = pid ? pid[s] : (uint32_t)s;
    if (aEs) aE[u] = aEs[s];
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

// caller index -> slot of a storage order
__global__ void k_slot_of(int n, const uint32_t* __restrict__ pid, uint32_t* __restrict__ inv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) {
        const uint32_t u = pid[s];
        if (u < (uint32_t)n) inv[u] = (uint32_t)s;
    }
}

}  // namespace

hipError_t slot_of_caller(int n, const uint32_t* pid, uint32_t* inv, hipStream_t st) {
    if (n > 0) k_slot_of<<<(n + 255) / 256, 256, 0, st>>>(n, pid, inv);
    return hipGetLastError();
}

hipError_t adjoint_segment(const TiledWork& w, int n, size_t cap, TapeStep* tape, int L, float4* abar,
                           const uint32_t* in_inv, float4* const part[2], float4* agv, float4* apam, float* aEs,
                           float* const aEp[2], uint32_t* const dst[2], hipStream_t st) {
    (void)n;
    const int nblk = w.total_blocks;
    // step s's A1 writes partials, map and dL/dE carry into set k, its A3 reads them there
    auto args = [&](int s, int k) {
        const TapeStep& t = tape[s];
        AdjArgs a;
        a.S = t.S; a.pid = t.pid; a.E = t.E; a.mass = t.mass;
        a.Snext = tape[s + 1].S;
        a.perm = t.perm;
        a.src = t.src;
        a.alist = reinterpret_cast<const uint4*>(t.alist);
        a.nactive = t.nactive;
        a.gv = t.gv;
        a.abar = abar; a.part = part[k]; a.agv = agv; a.apam = apam;
        a.in_pid = nullptr;
        a.in_inv = nullptr;
        a.aEs = aEs;
        a.aEp = aEp[k];
        a.dst = dst[k];
        a.wbuf = w.wbuf;
        a.scenes = w.scenes;
        a.cap = cap;
        a.mt = w.mt;
        return a;
    };
    auto grid = [&](int s) {
        ProfScope ps(w.prof, PK_ADJ_GRID, st);
        const TapeStep& t = tape[s];
        k_adj_grid<<<nblk, 512, 0, st>>>(w.scenes, w.bscene, t.tslot, w.wbuf, agv, apam, t.mp, t.flags, w.node_bc);
    };
    int k = 0;
    {
        ProfScope ps(w.prof, PK_ADJ_G2P, st);
        AdjArgs a = args(L - 1, k);
        a.in_pid = tape[L].pid;
        a.in_inv = in_inv;
        k_adj_g2p<false><<<w.nblocks_launch, TB, 0, st>>>(a, AdjNext{});
    }
    grid(L - 1);
    // A3 of step s and A1 of step s - 1 in one pass over binning s - 1
    for (int s = L - 1; s >= 1; --s) {
        const TapeStep& t = tape[s];
        const AdjNext nx{t.S, t.pid, t.E, t.mass, apam, part[k], dst[k], aEp[k]};
        {
            ProfScope ps(w.prof, PK_ADJ_FUSED, st);
            k_adj_g2p<true><<<w.nblocks_launch, TB, 0, st>>>(args(s - 1, k ^ 1), nx);
        }
        k ^= 1;
        grid(s - 1);
    }
    {
        ProfScope ps(w.prof, PK_ADJ_P2G, st);
        k_adj_p2g<<<w.nblocks_launch, TB, 0, st>>>(args(0, k));
    }
    return hipGetLastError();
}

hipError_t adjoint_segment_plain(const MlsGeom& g, int n, size_t cap, TapeStep* tape, int L, float4* abar,
                                 float4* part, float4* agv, float4* apam, float* aE_user, const float* E,
                                 const float* mass, const MatTable& mt, const uint32_t* node_bc, hipStream_t st,
                                 KernelProf* prof) {
    const size_t G = (size_t)g.nx * g.ny * g.nz;
    const int nb = (n + 255) / 256;
    const int gb = (int)std::min<size_t>((G + 255) / 256, 16384);
    for (int s = L - 1; s >= 0; --s) {
        TapeStep& t = tape[s];
        if (n > 0) {
            ProfScope ps(prof, PK_ADJ_G2P, st);
            k_adj_g2p_plain<<<nb, 256, 0, st>>>(n, cap, g, t.S, t.gv, abar, part, agv);
        }
        {
            ProfScope ps(prof, PK_ADJ_GRID, st);
            k_adj_grid_plain<<<gb, 256, 0, st>>>(g, agv, apam, t.mp, node_bc);
        }
        if (n > 0) {
            ProfScope ps(prof, PK_ADJ_P2G, st);
            k_adj_p2g_plain<<<nb, 256, 0, st>>>(n, cap, g, t.S, E, mass, apam, part, abar, aE_user, mt);
        }
    }
    return hipGetLastError();
}

hipError_t adjoint_from_user(int n, size_t cap, const uint32_t* pid, const float* ax, const float* av,
                             const float* aC, const float* aF, float4* abar, const float* aE, float* aEs,
                             hipStream_t st) {
    if (n > 0) k_gather_user<<<(n + 255) / 256, 256, 0, st>>>(n, cap, pid, ax, av, aC, aF, abar, aE, aEs);
    return hipGetLastError();
}

hipError_t adjoint_to_user(int n, size_t cap, const uint32_t* pid, const float4* abar, float* ax, float* av,
                           float* aC, float* aF, const float* aEs, float* aE, hipStream_t st) {
    if (n > 0) k_scatter_user<<<(n + 255) / 256, 256, 0, st>>>(n, cap, pid, abar, ax, av, aC, aF, aEs, aE);
    return hipGetLastError();
}

}  // namespace pg
