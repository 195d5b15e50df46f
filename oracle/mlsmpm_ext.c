/*
 * MLS-MPM FP64 oracle with per-particle materials and per-node boundary
 * flags.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The same step as mlsmpm_impl.h (and bit-identical to it when every particle
 * has the scene's material, Poisson ratio and volume and no node is flagged),
 * generalised to what the reference's GPU solver describes:
 *   - a material per particle (src/mpm_kernels.cu:295, :407-421): fixed-
 *     corotated, neo-Hookean (:97-128) or the Newtonian fluid of :131-149,
 *     Cauchy stress 2 mu_visc sym(grad v), with grad v the APIC C of the
 *     particle; like the elastic models it enters P2G as the Kirchhoff stress
 *     tau = J sigma;
 *   - a Poisson ratio and an initial volume V0 per particle (the reference's
 *     volume = mass / material density, src/mpm_solver.h:490-493);
 *   - a flag word per grid node (BoundaryType, src/mpm_data_structures.h:
 *     617-627): DIRICHLET_X/Y/Z zero that velocity component
 *     (src/mpm_kernels.cu:537-539), STICKY zeroes the whole velocity, SLIP
 *     removes each component that points out of the domain through the
 *     nearer wall of its axis (index < n/2: negative components; otherwise
 *     positive ones), NEUMANN_* leave the velocity unchanged (the reference
 *     never applies them).  The sticky band of p->bc nodes still applies.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define REAL double
#define SUF _x
#include "mlsmpm_impl.h"
#undef REAL
#undef SUF

enum { BC_DX = 1, BC_DY = 2, BC_DZ = 4, BC_STICKY = 64, BC_SLIP = 128 };

/* velocity components kept at a node (bit c = component c) */
static int keep_mask(const orc_mls_params* p, const int* idx, unsigned flags, const double* v) {
    const int dims[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; ++a)
        if (idx[a] < p->bc || idx[a] >= dims[a] - p->bc) return 0;
    if (flags & BC_STICKY) return 0;
    int keep = 7;
    if (flags & BC_DX) keep &= ~1;
    if (flags & BC_DY) keep &= ~2;
    if (flags & BC_DZ) keep &= ~4;
    if (flags & BC_SLIP)
        for (int a = 0; a < 3; ++a) {
            const int lower = 2 * idx[a] < dims[a];
            if ((lower && v[a] < 0.0) || (!lower && v[a] > 0.0)) keep &= ~(1 << a);
        }
    return keep;
}

void orc_kirchhoff_ext_f64(int model, double mu, double lam, double visc, const double* F, const double* C,
                           double* tau) {
    if (model == ORC_FLUID) {
        const double J = det3_x(F);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) tau[3 * i + j] = J * visc * (C[3 * i + j] + C[3 * j + i]);
    } else {
        kirchhoff_x(model, mu, lam, F, tau);
    }
}

void orc_mls_run_ext_f64(const orc_mls_params* p, int n, double* x, double* v, double* C, double* F,
                         const double* mass, const double* E, const int* model, const double* nu,
                         const double* visc, const double* vol, const unsigned* flags, int nsteps,
                         double* grid_m, double* grid_v) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t G = (size_t)nx * ny * nz;
    const double dx = p->dx, inv_dx = 1.0 / dx, dt = p->dt;
    const double o[3] = {p->ox, p->oy, p->oz};
    const double g[3] = {p->gx, p->gy, p->gz};
    const double s4 = 4.0 * inv_dx * inv_dx;
    double* gm = (double*)malloc(G * sizeof(double));
    double* gp = (double*)malloc(3 * G * sizeof(double));
    double* gv = (double*)malloc(3 * G * sizeof(double));
    for (int s = 0; s < nsteps; ++s) {
        memset(gm, 0, G * sizeof(double));
        memset(gp, 0, 3 * G * sizeof(double));
        for (int q = 0; q < n; ++q) {
            int b[3];
            double f[3], w[3][3];
            for (int a = 0; a < 3; ++a) stencil_x((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
            const double nq = nu ? nu[q] : p->nu;
            const double mu = E[q] / (2.0 * (1.0 + nq));
            const double lam = E[q] * nq / ((1.0 + nq) * (1.0 - 2.0 * nq));
            double tau[9], A[9];
            orc_kirchhoff_ext_f64(model ? model[q] : p->material, mu, lam, visc ? visc[q] : 0.0, F + 9 * q, C + 9 * q,
                                  tau);
            const double kscale = -dt * (vol ? vol[q] : p->p_vol) * s4;
            const double mq = mass[q];
            for (int i = 0; i < 9; ++i) A[i] = kscale * tau[i] + mq * C[9 * q + i];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        const int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                        if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                        const double wt = w[0][i] * w[1][j] * w[2][k];
                        const double dp[3] = {(i - f[0]) * dx, (j - f[1]) * dx, (k - f[2]) * dx};
                        const size_t id = ((size_t)gk * ny + gj) * nx + gi;
                        gm[id] += wt * mq;
                        for (int c = 0; c < 3; ++c)
                            gp[3 * id + c] += wt * (mq * v[3 * q + c] + A[3 * c] * dp[0] + A[3 * c + 1] * dp[1] +
                                                    A[3 * c + 2] * dp[2]);
                    }
        }
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t id = ((size_t)k * ny + j) * nx + i;
                    const int idx[3] = {i, j, k};
                    const double m = gm[id];
                    double vv[3];
                    for (int c = 0; c < 3; ++c) vv[c] = m > 0.0 ? gp[3 * id + c] / m + dt * g[c] : 0.0;
                    const int keep = keep_mask(p, idx, flags ? flags[id] : 0u, vv);
                    for (int c = 0; c < 3; ++c) gv[3 * id + c] = (keep >> c) & 1 ? vv[c] : 0.0;
                }
        for (int q = 0; q < n; ++q) {
            int b[3];
            double f[3], w[3][3];
            for (int a = 0; a < 3; ++a) stencil_x((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
            double nvel[3] = {0, 0, 0}, nc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        const int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                        if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                        const double wt = w[0][i] * w[1][j] * w[2][k];
                        const double dp[3] = {(i - f[0]) * dx, (j - f[1]) * dx, (k - f[2]) * dx};
                        const size_t id = ((size_t)gk * ny + gj) * nx + gi;
                        for (int c = 0; c < 3; ++c) {
                            nvel[c] += wt * gv[3 * id + c];
                            for (int d = 0; d < 3; ++d) nc[3 * c + d] += wt * gv[3 * id + c] * dp[d];
                        }
                    }
            for (int i = 0; i < 9; ++i) nc[i] *= s4;
            double Fn[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double acc = 0;
                    for (int k = 0; k < 3; ++k) acc += ((i == k ? 1.0 : 0.0) + dt * nc[3 * i + k]) * F[9 * q + 3 * k + j];
                    Fn[3 * i + j] = acc;
                }
            for (int i = 0; i < 9; ++i) {
                F[9 * q + i] = Fn[i];
                C[9 * q + i] = nc[i];
            }
            for (int c = 0; c < 3; ++c) {
                v[3 * q + c] = nvel[c];
                x[3 * q + c] += dt * nvel[c];
            }
        }
    }
    if (grid_m) memcpy(grid_m, gm, G * sizeof(double));
    if (grid_v) memcpy(grid_v, gv, 3 * G * sizeof(double));
    free(gm);
    free(gp);
    free(gv);
}
