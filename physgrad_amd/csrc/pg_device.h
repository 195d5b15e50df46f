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

// inverse transpose via cofactors: inv(A)^T = cof(A)/det(A); returns det
PG_DEV float inv_t3(const float* A, float* out) {
    float c0 = A[4] * A[8] - A[5] * A[7];
    float c1 = A[5] * A[6] - A[3] * A[8];
    float c2 = A[3] * A[7] - A[4] * A[6];
    float c3 = A[2] * A[7] - A[1] * A[8];
    float c4 = A[0] * A[8] - A[2] * A[6];
    float c5 = A[1] * A[6] - A[0] * A[7];
    float c6 = A[1] * A[5] - A[2] * A[4];
    float c7 = A[2] * A[3] - A[0] * A[5];
    float c8 = A[0] * A[4] - A[1] * A[3];
    float d = A[0] * c0 + A[1] * c1 + A[2] * c2;
    float id = 1.0f / d;
    out[0] = c0 * id; out[1] = c1 * id; out[2] = c2 * id;
    out[3] = c3 * id; out[4] = c4 * id; out[5] = c5 * id;
    out[6] = c6 * id; out[7] = c7 * id; out[8] = c8 * id;
    return d;
}

// Orthogonal polar factor by Newton's iteration X <- (X + X^-T)/2 (Higham 1986),
// the same iteration as the oracle, stopped once the relative change drops below
// 1e-6 (quadratic convergence puts the next error below fp32 rounding), <= 24 its.
PG_DEV void polar(const float* F, float* R) {
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = F[i];
    for (int it = 0; it < 24; ++it) {
        float Y[9];
        inv_t3(R, Y);
        float diff = 0.f, nrm = 0.f;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            float nx = 0.5f * (R[i] + Y[i]);
            float d = nx - R[i];
            diff += d * d;
            nrm += nx * nx;
            R[i] = nx;
        }
        if (diff <= 1e-12f * nrm) break;
    }
}

enum { MAT_FIXED_COROTATED = 0, MAT_NEO_HOOKEAN = 1 };

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

// Quadratic B-spline stencil of MLS-MPM: base = floor(xs - 1/2), f = xs - base,
// w0 = (1.5-f)^2/2, w1 = 3/4 - (f-1)^2, w2 = (f-1/2)^2/2.
struct Stencil {
    int b[3];
    float f[3];
    float w[3][3];
};

PG_DEV void make_stencil(const float* x, const float* o, float inv_dx, Stencil& s) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xs = (x[a] - o[a]) * inv_dx;
        int b = (int)floorf(xs - 0.5f);
        float f = xs - (float)b;
        s.b[a] = b;
        s.f[a] = f;
        s.w[a][0] = 0.5f * (1.5f - f) * (1.5f - f);
        s.w[a][1] = 0.75f - (f - 1.f) * (f - 1.f);
        s.w[a][2] = 0.5f * (f - 0.5f) * (f - 0.5f);
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
