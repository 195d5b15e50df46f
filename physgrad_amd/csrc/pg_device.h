// Device-side math shared by the MPM kernels (gfx950, wave64).
//
// Conventions
//   3x3 matrices are row-major float[9].
//   MLS-MPM particle state lives in six float4 planes of length cap:
//     Q0 = (x0, x1, x2, F00)   Q1 = (F01, F02, F10, F11)   Q2 = (F12, F20, F21, F22)
//     Q3 = (v0, v1, v2, C00)   Q4 = (C01, C02, C10, C11)   Q5 = (C12, C20, C21, C22)
//   so that G2P, which needs only x and F, streams the first three planes.
//   Grid nodes are indexed (k*ny + j)*nx + i, as MPMGrid::getNodeIndex
//   (src/material_point_method.h:250-253).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PG_DEV __device__ __forceinline__

namespace pg {

struct Mat3 {
    float a[9];
};

PG_DEV float det3(const float* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// cofactor matrix of A (row-major, so that inv(A)^T = cof(A) / det(A)); returns det(A)
PG_DEV float cofactors3(const float* A, float* c) {
    c[0] = A[4] * A[8] - A[5] * A[7];
    c[1] = A[5] * A[6] - A[3] * A[8];
    c[2] = A[3] * A[7] - A[4] * A[6];
    c[3] = A[2] * A[7] - A[1] * A[8];
    c[4] = A[0] * A[8] - A[2] * A[6];
    c[5] = A[1] * A[6] - A[0] * A[7];
    c[6] = A[1] * A[5] - A[2] * A[4];
    c[7] = A[2] * A[3] - A[0] * A[5];
    c[8] = A[0] * A[4] - A[1] * A[3];
    return A[0] * c[0] + A[1] * c[1] + A[2] * c[2];
}

// inverse transpose via cofactors: inv(A)^T = cof(A)/det(A); returns det
PG_DEV float inv_t3(const float* A, float* out) {
    float c[9];
    const float d = cofactors3(A, c);
    const float id = 1.0f / d;
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = c[i] * id;
    return d;
}

// Orthogonal polar factor by Newton's iteration X <- (X + X^-T)/2 (Higham 1986),
// the oracle's iteration.  Once the step's relative change d is small the
// iteration is quadratic (a singular value 1 + e goes to 1 + e^2/2 + O(e^3),
// and d = e to first order), so stopping at d <= 1e-4 leaves an error of about
// 5e-9, below fp32 rounding.  ||X||_F^2 tends to 3 for a rotation, so the test
// compares the squared change with 3e-8 instead of with the norm.  X^-T uses
// the hardware reciprocal (about 1 ulp): a relative error eps in it moves the
// fixed point X = X^-T by O(eps), which is fp32 rounding again.  <= 24 steps.
PG_DEV void polar(const float* F, float* R) {
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = F[i];
    for (int it = 0; it < 24; ++it) {
        float c[9];
        const float h = 0.5f * __builtin_amdgcn_rcpf(cofactors3(R, c));
        float diff = 0.f;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const float nx = c[i] * h + 0.5f * R[i];
            const float d = nx - R[i];
            diff += d * d;
            R[i] = nx;
        }
        if (diff <= 3e-8f) break;
    }
}

enum { MAT_FIXED_COROTATED = 0, MAT_NEO_HOOKEAN = 1, MAT_FLUID = 2 };

// Constitutive data of one particle: the model and its Lame parameters,
// viscosity and initial volume (resolved from the scene or the material table).
struct PMat {
    int model;
    float mu, lam, visc, vol;
    float emu, elam;   // d mu / dE, d lam / dE (the E adjoint)
};

// One material of a context's table (device copy of pg_material_desc).
struct MatRec {
    int model;
    float E;
    float emu, elam;   // Lame parameters per unit E
    float visc;
    float inv_rho;     // 1 / density: particle volume = mass / density (0: the scene's volume)
    float pad[2];
};

// Boundary flags of a grid node (the reference's BoundaryType bits).
enum : uint32_t { BC_DIRICHLET_X = 1u, BC_DIRICHLET_Y = 2u, BC_DIRICHLET_Z = 4u, BC_STICKY = 64u, BC_SLIP = 128u };

// Velocity components a node keeps (bit c = component c): none in the sticky
// band of `band` nodes or on a STICKY node; DIRICHLET_X/Y/Z drop their
// component; SLIP drops each component that points out of the domain through
// the nearer wall of its axis (index < n/2: negative; else positive), given the
// node's velocity v before the mask.  NEUMANN bits leave it unchanged.
PG_DEV uint32_t bc_keep(const int* idx, const int* n, int band, uint32_t flags, const float* v) {
    bool wall = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) wall = wall || idx[a] < band || idx[a] >= n[a] - band;
    if (wall || (flags & BC_STICKY)) return 0u;
    uint32_t keep = 7u & ~flags;   // DIRICHLET_X/Y/Z are bits 0..2
    if (flags & BC_SLIP) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const bool lower = 2 * idx[a] < n[a];
            if (lower ? v[a] < 0.f : v[a] > 0.f) keep &= ~(1u << a);
        }
    }
    return keep;
}

// Kirchhoff stress tau = P(F) F^T.
//   fixed-corotated: 2 mu (F - R) F^T + lam (J-1) J I
//   neo-Hookean:     mu (F F^T - I) + lam ln(J) I   (J times the Cauchy stress of
//                    src/mpm_kernels.cu:97-128)
PG_DEV void kirchhoff(int material, float mu, float lam, const float* F, float* tau) {
    float J = det3(F);
    if (material == MAT_NEO_HOOKEAN) {
        float lnJ = logf(J);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float b = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
                tau[3 * i + j] = mu * (b - (i == j ? 1.f : 0.f)) + (i == j ? lam * lnJ : 0.f);
            }
    } else {
        float R[9];
        polar(F, R);
        float vol = lam * (J - 1.f) * J;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float s = (F[3 * i] - R[3 * i]) * F[3 * j] + (F[3 * i + 1] - R[3 * i + 1]) * F[3 * j + 1] +
                          (F[3 * i + 2] - R[3 * i + 2]) * F[3 * j + 2];
                tau[3 * i + j] = 2.f * mu * s + (i == j ? vol : 0.f);
            }
    }
}

// Vector-Jacobian product of tau(F; mu, lam): aF += (d tau/d F)^T : at, and the
// contributions to dL/dmu, dL/dlam.
//   neo-Hookean:  d tau = mu (dF F^T + F dF^T) + lam tr(F^-1 dF) I
//   fixed-corot.: d tau = 2mu (dF F^T + F dF^T - dR F^T - R dF^T) + lam (2J-1) dJ I,
//                 dR = R [w]x, w = (tr(S) I - S)^-1 axial(R^T dF - dF^T R), S = R^T F
//                 (differentiating F = R S with R^T dR skew and dS symmetric).
PG_DEV void kirchhoff_vjp(int material, float mu, float lam, const float* F, const float* at, float* aF,
                          float& amu, float& alam) {
    float FiT[9];
    const float J = inv_t3(F, FiT);
    const float tr = at[0] + at[4] + at[8];
    float aFm[9], aTF[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            aFm[3 * i + j] = at[3 * i] * F[j] + at[3 * i + 1] * F[3 + j] + at[3 * i + 2] * F[6 + j];  // at F
            aTF[3 * i + j] = at[i] * F[j] + at[3 + i] * F[3 + j] + at[6 + i] * F[6 + j];              // at^T F
        }
    if (material == MAT_NEO_HOOKEAN) {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float b = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
                s += at[3 * i + j] * (b - (i == j ? 1.f : 0.f));
            }
        amu += s;
        alam += tr * logf(J);
#pragma unroll
        for (int i = 0; i < 9; ++i) aF[i] += mu * (aFm[i] + aTF[i]) + lam * tr * FiT[i];
    } else {
        float R[9];
        polar(F, R);
        float S[9], M[9], aTR[9];
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                S[3 * i + j] = R[i] * F[j] + R[3 + i] * F[3 + j] + R[6 + i] * F[6 + j];            // R^T F
                M[3 * i + j] = R[i] * aFm[j] + R[3 + i] * aFm[3 + j] + R[6 + i] * aFm[6 + j];      // R^T (at F)
                aTR[3 * i + j] = at[i] * R[j] + at[3 + i] * R[3 + j] + at[6 + i] * R[6 + j];        // at^T R
                float b = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
                float rf = R[3 * i] * F[3 * j] + R[3 * i + 1] * F[3 * j + 1] + R[3 * i + 2] * F[3 * j + 2];
                s += at[3 * i + j] * 2.f * (b - rf);
            }
        amu += s;
        alam += tr * (J - 1.f) * J;
        // symmetric part of S, then H = tr(S) I - S
        const float s01 = 0.5f * (S[1] + S[3]), s02 = 0.5f * (S[2] + S[6]), s12 = 0.5f * (S[5] + S[7]);
        const float trS = S[0] + S[4] + S[8];
        float H[9] = {trS - S[0], -s01, -s02, -s01, trS - S[4], -s12, -s02, -s12, trS - S[8]};
        float Hi[9];
        inv_t3(H, Hi);  // H symmetric: inverse transpose is the inverse
        const float b0 = M[7] - M[5], b1 = M[2] - M[6], b2 = M[3] - M[1];
        const float u0 = Hi[0] * b0 + Hi[1] * b1 + Hi[2] * b2;
        const float u1 = Hi[3] * b0 + Hi[4] * b1 + Hi[5] * b2;
        const float u2 = Hi[6] * b0 + Hi[7] * b1 + Hi[8] * b2;
        // R [u]x
        float Ru[9];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Ru[3 * i + 0] = R[3 * i + 1] * u2 - R[3 * i + 2] * u1;
            Ru[3 * i + 1] = -R[3 * i] * u2 + R[3 * i + 2] * u0;
            Ru[3 * i + 2] = R[3 * i] * u1 - R[3 * i + 1] * u0;
        }
        const float vol = lam * (2.f * J - 1.f) * J * tr;
#pragma unroll
        for (int i = 0; i < 9; ++i) aF[i] += 2.f * mu * (aFm[i] + aTF[i] - aTR[i] - Ru[i]) + vol * FiT[i];
    }
}

// Kirchhoff stress of any model: the elastic models as kirchhoff(); the
// Newtonian fluid (src/mpm_kernels.cu:131-149: Cauchy sigma = 2 visc sym(grad v),
// grad v = the APIC C) as tau = J sigma = J visc (C + C^T).
PG_DEV void stress_of(const PMat& m, const float* F, const float* C, float* tau) {
    if (m.model == MAT_FLUID) {
        const float jv = det3(F) * m.visc;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) tau[3 * i + j] = jv * (C[3 * i + j] + C[3 * j + i]);
    } else {
        kirchhoff(m.model, m.mu, m.lam, F, tau);
    }
}

// VJP of stress_of: aF += (d tau / d F)^T : at, aC += (d tau / d C)^T : at (fluid
// only) and dE += (d tau / d E) : at (elastic models only).
//   fluid: d tau = visc (C + C^T) dJ + J visc (dC + dC^T), dJ = J tr(F^-1 dF)
PG_DEV void stress_vjp(const PMat& m, const float* F, const float* C, const float* at, float* aF, float* aC,
                       float& dE) {
    if (m.model == MAT_FLUID) {
        float FiT[9];
        const float J = inv_t3(F, FiT);
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s += at[3 * i + j] * (C[3 * i + j] + C[3 * j + i]);
        const float js = m.visc * J * s;
#pragma unroll
        for (int i = 0; i < 9; ++i) aF[i] += js * FiT[i];
        const float jv = J * m.visc;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) aC[3 * i + j] += jv * (at[3 * i + j] + at[3 * j + i]);
        return;
    }
    float amu = 0.f, alam = 0.f;
    kirchhoff_vjp(m.model, m.mu, m.lam, F, at, aF, amu, alam);
    dE += amu * m.emu + alam * m.elam;
}

// Quadratic B-spline stencil of MLS-MPM: base = floor(xs - 1/2), f = xs - base,
// w0 = (1.5-f)^2/2, w1 = 3/4 - (f-1)^2, w2 = (f-1/2)^2/2.
struct Stencil {
    int b[3];
    float f[3];
    float w[3][3];
};

// Position in cell units and the quadratic stencil's base cell, with explicit
// IEEE rounding (no contraction into an FMA) so that every kernel that bins or
// stencils a particle computes the same base for it.
PG_DEV float grid_coord(float x, float o, float inv_dx) {
#pragma clang fp contract(off)
    return (x - o) * inv_dx;
}
PG_DEV int stencil_base(float xs) {
#pragma clang fp contract(off)
    return (int)floorf(xs - 0.5f);
}

// Quadratic B-spline weights of the three stencil nodes at offset f in [0.5, 1.5).
PG_DEV void stencil_weights(float f, float* w) {
    w[0] = 0.5f * (1.5f - f) * (1.5f - f);
    w[1] = 0.75f - (f - 1.f) * (f - 1.f);
    w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}

PG_DEV void make_stencil(const float* x, const float* o, float inv_dx, Stencil& s) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xs = grid_coord(x[a], o[a], inv_dx);
        const int b = stencil_base(xs);
        const float f = xs - (float)b;
        s.b[a] = b;
        s.f[a] = f;
        stencil_weights(f, s.w[a]);
    }
}

PG_DEV void load_state(const float4* __restrict__ S, size_t cap, size_t i, float* x, float* F, float* v, float* C) {
    float4 q0 = S[i], q1 = S[cap + i], q2 = S[2 * cap + i];
    float4 q3 = S[3 * cap + i], q4 = S[4 * cap + i], q5 = S[5 * cap + i];
    x[0] = q0.x; x[1] = q0.y; x[2] = q0.z; F[0] = q0.w;
    F[1] = q1.x; F[2] = q1.y; F[3] = q1.z; F[4] = q1.w;
    F[5] = q2.x; F[6] = q2.y; F[7] = q2.z; F[8] = q2.w;
    v[0] = q3.x; v[1] = q3.y; v[2] = q3.z; C[0] = q3.w;
    C[1] = q4.x; C[2] = q4.y; C[3] = q4.z; C[4] = q4.w;
    C[5] = q5.x; C[6] = q5.y; C[7] = q5.z; C[8] = q5.w;
}

PG_DEV void load_xF(const float4* __restrict__ S, size_t cap, size_t i, float* x, float* F) {
    float4 q0 = S[i], q1 = S[cap + i], q2 = S[2 * cap + i];
    x[0] = q0.x; x[1] = q0.y; x[2] = q0.z; F[0] = q0.w;
    F[1] = q1.x; F[2] = q1.y; F[3] = q1.z; F[4] = q1.w;
    F[5] = q2.x; F[6] = q2.y; F[7] = q2.z; F[8] = q2.w;
}

PG_DEV void store_state(float4* __restrict__ S, size_t cap, size_t i, const float* x, const float* F, const float* v,
                        const float* C) {
    S[i] = make_float4(x[0], x[1], x[2], F[0]);
    S[cap + i] = make_float4(F[1], F[2], F[3], F[4]);
    S[2 * cap + i] = make_float4(F[5], F[6], F[7], F[8]);
    S[3 * cap + i] = make_float4(v[0], v[1], v[2], C[0]);
    S[4 * cap + i] = make_float4(C[1], C[2], C[3], C[4]);
    S[5 * cap + i] = make_float4(C[5], C[6], C[7], C[8]);
}

PG_DEV void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

}  // namespace pg
