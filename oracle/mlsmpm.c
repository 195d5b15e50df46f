/*
 * MLS-MPM oracle: FP64 forward (reference for error reporting and finite
 * differences), FP32 forward (the CPU port timed beside the GPU) and the FP64
 * reverse step.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference has no executable MLS-MPM (its GPU solver,
 * src/mpm_kernels.cu:247-545, does not compile); the constitutive law
 * (src/mpm_kernels.cu:97-128), F update (:384-402) and grid update (:498-545)
 * name the semantics kept here.  Pinned by tests/test_oracle_mls.py: closed
 * forms (uniform stretch stress, single-particle free fall), conservation of
 * mass/momentum/angular momentum, rigid translation keeping F = I, and central
 * finite differences of this forward for the reverse step.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define REAL double
#define SUF _d
#include "mlsmpm_impl.h"
#undef REAL
#undef SUF

#define REAL float
#define SUF _f
#include "mlsmpm_impl.h"
#undef REAL
#undef SUF

void orc_mls_run_f64(const orc_mls_params* p, int n, double* x, double* v, double* C, double* F,
                     const double* mass, const double* E, int nsteps, double* grid_m, double* grid_v) {
    const size_t G = (size_t)p->nx * p->ny * p->nz;
    double* gm = (double*)malloc(G * sizeof(double));
    double* gp = (double*)malloc(3 * G * sizeof(double));
    double* gv = (double*)malloc(3 * G * sizeof(double));
    for (int s = 0; s < nsteps; ++s) mls_step_d(p, n, x, v, C, F, mass, E, gm, gp, gv);
    if (grid_m) memcpy(grid_m, gm, G * sizeof(double));
    if (grid_v) memcpy(grid_v, gv, 3 * G * sizeof(double));
    free(gm); free(gp); free(gv);
}

void orc_mls_run_f32(const orc_mls_params* p, int n, float* x, float* v, float* C, float* F,
                     const float* mass, const float* E, int nsteps, float* grid_m, float* grid_v) {
    const size_t G = (size_t)p->nx * p->ny * p->nz;
    float* gm = (float*)malloc(G * sizeof(float));
    float* gp = (float*)malloc(3 * G * sizeof(float));
    float* gv = (float*)malloc(3 * G * sizeof(float));
    for (int s = 0; s < nsteps; ++s) mls_step_f(p, n, x, v, C, F, mass, E, gm, gp, gv);
    if (grid_m) memcpy(grid_m, gm, G * sizeof(float));
    if (grid_v) memcpy(grid_v, gv, 3 * G * sizeof(float));
    free(gm); free(gp); free(gv);
}

void orc_kirchhoff_f64(int material, double mu, double lam, const double* F, double* tau) {
    kirchhoff_d(material, mu, lam, F, tau);
}

int orc_polar_f64(const double* F, double* R, double* S) {
    int it = polar_d(F, R);
    if (S) {
        double T[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += R[3 * k + i] * F[3 * k + j];
                T[3 * i + j] = s;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) S[3 * i + j] = 0.5 * (T[3 * i + j] + T[3 * j + i]);
    }
    return it;
}

/* ---------------------------------------------------------------------------
 * Reverse step.
 *
 * Vector-Jacobian product of tau(F; mu, lam).
 *   neo-Hookean     d tau = mu (dF F^T + F dF^T) + lam tr(F^-1 dF) I
 *   fixed-corot.    d tau = 2mu (dF F^T + F dF^T - dR F^T - R dF^T) + lam (2J-1) dJ I,
 *                   dR = R [w]x with w = (tr(S) I - S)^-1 axial(R^T dF - dF^T R),
 *                   S = R^T F, which follows from differentiating F = R S with
 *                   R^T dR skew and dS symmetric.
 * ------------------------------------------------------------------------- */
static void mat_mul(const double* A, const double* B, double* Cm) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Cm[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
static void mat_tmul(const double* A, const double* B, double* Cm) { /* A^T B */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Cm[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
static void mat_mult(const double* A, const double* B, double* Cm) { /* A B^T */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Cm[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}
static void solve3(const double* H, const double* b, double* x) {
    double Hi[9];
    inv_t3_d(H, Hi); /* inverse transpose; H is symmetric */
    for (int i = 0; i < 3; ++i) x[i] = Hi[3 * i] * b[0] + Hi[3 * i + 1] * b[1] + Hi[3 * i + 2] * b[2];
}

static void kirchhoff_vjp(int material, double mu, double lam, const double* F, const double* at,
                          double* aF, double* amu, double* alam) {
    double J = det3_d(F);
    double FiT[9];
    inv_t3_d(F, FiT);
    double tr = at[0] + at[4] + at[8];
    double aFm[9], aTF[9];
    mat_mul(at, F, aFm);  /* at F */
    mat_tmul(at, F, aTF); /* at^T F */
    if (material == ORC_NEO_HOOKEAN) {
        double B[9];
        mat_mult(F, F, B);
        double s = 0;
        for (int i = 0; i < 9; ++i) s += at[i] * (B[i] - ((i % 4) == 0 ? 1.0 : 0.0));
        *amu += s;
        *alam += tr * log(J);
        for (int i = 0; i < 9; ++i) aF[i] += mu * (aFm[i] + aTF[i]) + lam * tr * FiT[i];
    } else {
        double R[9], S[9];
        orc_polar_f64(F, R, S);
        double B[9], RF[9];
        mat_mult(F, F, B);
        mat_mult(R, F, RF);
        double s = 0;
        for (int i = 0; i < 9; ++i) s += at[i] * 2.0 * (B[i] - RF[i]);
        *amu += s;
        *alam += tr * (J - 1.0) * J;
        double aTR[9];
        mat_tmul(at, R, aTR); /* at^T R */
        /* R-bar = at F; u = H^-1 axial(R^T Rbar - Rbar^T R) */
        double M[9];
        mat_tmul(R, aFm, M);
        double ax3[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
        double trS = S[0] + S[4] + S[8];
        double H[9];
        for (int i = 0; i < 9; ++i) H[i] = -S[i];
        H[0] += trS; H[4] += trS; H[8] += trS;
        double u[3];
        solve3(H, ax3, u);
        double ux[9] = {0, -u[2], u[1], u[2], 0, -u[0], -u[1], u[0], 0};
        double Rux[9];
        mat_mul(R, ux, Rux);
        for (int i = 0; i < 9; ++i)
            aF[i] += 2.0 * mu * (aFm[i] + aTF[i] - aTR[i] - Rux[i]) + lam * (2.0 * J - 1.0) * J * tr * FiT[i];
    }
}

void orc_mls_adjoint_step_f64(const orc_mls_params* p, int n,
                              const double* x, const double* v, const double* C, const double* F,
                              const double* mass, const double* E,
                              double* ax, double* av, double* aC, double* aF, double* aE) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t G = (size_t)nx * ny * nz;
    const double dx = p->dx, inv_dx = 1.0 / dx, dt = p->dt, nu = p->nu;
    const double o[3] = {p->ox, p->oy, p->oz};
    const double g[3] = {p->gx, p->gy, p->gz};
    const double s4 = 4.0 * inv_dx * inv_dx;
    const double kscale = -dt * p->p_vol * s4;
    const int dims[3] = {nx, ny, nz};
    const double emu = 1.0 / (2.0 * (1.0 + nu));
    const double elam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));

    double* gm = (double*)calloc(G, sizeof(double));
    double* gp = (double*)calloc(3 * G, sizeof(double));
    double* gv = (double*)calloc(3 * G, sizeof(double));
    double* agv = (double*)calloc(3 * G, sizeof(double));
    double* agp = (double*)calloc(3 * G, sizeof(double));
    double* agm = (double*)calloc(G, sizeof(double));
    double* Aall = (double*)malloc(9 * (size_t)n * sizeof(double));

    /* forward recompute: P2G and grid */
    for (int q = 0; q < n; ++q) {
        int b[3];
        double f[3], w[3][3];
        for (int a = 0; a < 3; ++a) stencil_d((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
        double tau[9];
        kirchhoff_d(p->material, E[q] * emu, E[q] * elam, F + 9 * q, tau);
        double* A = Aall + 9 * (size_t)q;
        for (int i = 0; i < 9; ++i) A[i] = kscale * tau[i] + mass[q] * C[9 * q + i];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                    double wt = w[0][i] * w[1][j] * w[2][k];
                    double dp[3] = {(i - f[0]) * dx, (j - f[1]) * dx, (k - f[2]) * dx};
                    size_t id = ((size_t)gk * ny + gj) * nx + gi;
                    gm[id] += wt * mass[q];
                    for (int c = 0; c < 3; ++c)
                        gp[3 * id + c] += wt * (mass[q] * v[3 * q + c] + A[3 * c] * dp[0] + A[3 * c + 1] * dp[1] + A[3 * c + 2] * dp[2]);
                }
    }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t id = ((size_t)k * ny + j) * nx + i;
                const int idx[3] = {i, j, k};
                int sticky = 0;
                for (int a = 0; a < 3; ++a)
                    if (idx[a] < p->bc || idx[a] >= dims[a] - p->bc) sticky = 1;
                for (int c = 0; c < 3; ++c)
                    gv[3 * id + c] = (!sticky && gm[id] > 0) ? gp[3 * id + c] / gm[id] + dt * g[c] : 0.0;
            }

    /* reverse G2P */
    for (int q = 0; q < n; ++q) {
        int b[3];
        double f[3], w[3][3], dw[3][3];
        for (int a = 0; a < 3; ++a) {
            stencil_d((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
            dw[a][0] = f[a] - 1.5;
            dw[a][1] = -2.0 * (f[a] - 1.0);
            dw[a][2] = f[a] - 0.5;
        }
        /* recompute new v and C */
        double nv[3] = {0, 0, 0}, nc[9] = {0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                    double wt = w[0][i] * w[1][j] * w[2][k];
                    double dp[3] = {(i - f[0]) * dx, (j - f[1]) * dx, (k - f[2]) * dx};
                    size_t id = ((size_t)gk * ny + gj) * nx + gi;
                    for (int c = 0; c < 3; ++c) {
                        nv[c] += wt * gv[3 * id + c];
                        for (int d = 0; d < 3; ++d) nc[3 * c + d] += wt * gv[3 * id + c] * dp[d] * s4;
                    }
                }
        const double* Fq = F + 9 * q;
        double* aFq = aF + 9 * q;
        /* x' = x + dt v', F' = (I + dt C') F */
        double anv[3], anc[9], aFold[9];
        for (int c = 0; c < 3; ++c) anv[c] = av[3 * q + c] + dt * ax[3 * q + c];
        double tmp[9];
        mat_mult(aFq, Fq, tmp); /* aF' F^T */
        for (int i = 0; i < 9; ++i) anc[i] = aC[9 * q + i] + dt * tmp[i];
        double Mx[9];
        for (int i = 0; i < 9; ++i) Mx[i] = ((i % 4) == 0 ? 1.0 : 0.0) + dt * nc[i];
        mat_tmul(Mx, aFq, aFold); /* (I + dt C')^T aF' */
        double axq[3] = {ax[3 * q], ax[3 * q + 1], ax[3 * q + 2]};
        double af[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                    const int d3[3] = {i, j, k};
                    double wt = w[0][i] * w[1][j] * w[2][k];
                    double dp[3] = {(i - f[0]) * dx, (j - f[1]) * dx, (k - f[2]) * dx};
                    size_t id = ((size_t)gk * ny + gj) * nx + gi;
                    const double* gvi = gv + 3 * id;
                    double Cdp[3];
                    for (int c = 0; c < 3; ++c) Cdp[c] = anc[3 * c] * dp[0] + anc[3 * c + 1] * dp[1] + anc[3 * c + 2] * dp[2];
                    double aw = 0;
                    for (int c = 0; c < 3; ++c) {
                        agv[3 * id + c] += wt * anv[c] + s4 * wt * Cdp[c];
                        aw += anv[c] * gvi[c] + s4 * gvi[c] * Cdp[c];
                    }
                    for (int d = 0; d < 3; ++d) {
                        double adp = s4 * wt * (anc[d] * gvi[0] + anc[3 + d] * gvi[1] + anc[6 + d] * gvi[2]);
                        af[d] += -dx * adp;
                    }
                    af[0] += aw * dw[0][d3[0]] * w[1][d3[1]] * w[2][d3[2]];
                    af[1] += aw * w[0][d3[0]] * dw[1][d3[1]] * w[2][d3[2]];
                    af[2] += aw * w[0][d3[0]] * w[1][d3[1]] * dw[2][d3[2]];
                }
        for (int c = 0; c < 3; ++c) ax[3 * q + c] = axq[c] + af[c] * inv_dx;
        for (int i = 0; i < 9; ++i) aFq[i] = aFold[i];
        /* old v and C do not feed G2P */
        for (int c = 0; c < 3; ++c) av[3 * q + c] = 0;
        for (int i = 0; i < 9; ++i) aC[9 * q + i] = 0;
    }

    /* reverse grid update */
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t id = ((size_t)k * ny + j) * nx + i;
                const int idx[3] = {i, j, k};
                int sticky = 0;
                for (int a = 0; a < 3; ++a)
                    if (idx[a] < p->bc || idx[a] >= dims[a] - p->bc) sticky = 1;
                if (sticky || !(gm[id] > 0)) continue;
                double im = 1.0 / gm[id], s = 0;
                for (int c = 0; c < 3; ++c) {
                    agp[3 * id + c] = agv[3 * id + c] * im;
                    s += agv[3 * id + c] * gp[3 * id + c];
                }
                agm[id] = -s * im * im;
            }

    /* reverse P2G */
    for (int q = 0; q < n; ++q) {
        int b[3];
        double f[3], w[3][3], dw[3][3];
        for (int a = 0; a < 3; ++a) {
            stencil_d((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
            dw[a][0] = f[a] - 1.5;
            dw[a][1] = -2.0 * (f[a] - 1.0);
            dw[a][2] = f[a] - 0.5;
        }
        const double* A = Aall + 9 * (size_t)q;
        const double mq = mass[q];
        double aA[9] = {0}, af[3] = {0, 0, 0}, avq[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                    const int d3[3] = {i, j, k};
                    double wt = w[0][i] * w[1][j] * w[2][k];
                    double dp[3] = {(i - f[0]) * dx, (j - f[1]) * dx, (k - f[2]) * dx};
                    size_t id = ((size_t)gk * ny + gj) * nx + gi;
                    const double* ap = agp + 3 * id;
                    double aw = agm[id] * mq;
                    for (int c = 0; c < 3; ++c) {
                        double contrib = mq * v[3 * q + c] + A[3 * c] * dp[0] + A[3 * c + 1] * dp[1] + A[3 * c + 2] * dp[2];
                        aw += ap[c] * contrib;
                        avq[c] += wt * mq * ap[c];
                        for (int d = 0; d < 3; ++d) aA[3 * c + d] += wt * ap[c] * dp[d];
                    }
                    for (int d = 0; d < 3; ++d) {
                        double adp = wt * (A[d] * ap[0] + A[3 + d] * ap[1] + A[6 + d] * ap[2]);
                        af[d] += -dx * adp;
                    }
                    af[0] += aw * dw[0][d3[0]] * w[1][d3[1]] * w[2][d3[2]];
                    af[1] += aw * w[0][d3[0]] * dw[1][d3[1]] * w[2][d3[2]];
                    af[2] += aw * w[0][d3[0]] * w[1][d3[1]] * dw[2][d3[2]];
                }
        for (int c = 0; c < 3; ++c) {
            ax[3 * q + c] += af[c] * inv_dx;
            av[3 * q + c] += avq[c];
        }
        for (int i = 0; i < 9; ++i) aC[9 * q + i] += mq * aA[i];
        double at[9];
        for (int i = 0; i < 9; ++i) at[i] = kscale * aA[i];
        double amu = 0, alam = 0;
        kirchhoff_vjp(p->material, E[q] * emu, E[q] * elam, F + 9 * q, at, aF + 9 * q, &amu, &alam);
        if (aE) aE[q] += amu * emu + alam * elam;
    }

    free(gm); free(gp); free(gv); free(agv); free(agp); free(agm); free(Aall);
}
