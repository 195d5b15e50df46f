/*
 * Multi-threaded FP32 port of the MLS-MPM step (OpenMP), the CPU baseline on
 * all host cores.  TEST INFRASTRUCTURE ONLY (see oracle.h): it is timed beside
 * the GPU path and never used as the product.
 *
 * Same arithmetic per particle and node as mls_step_f in mlsmpm_impl.h (which
 * restates src/mpm_kernels.cu:97-128, :384-402, :477-489, :498-545 in the
 * MLS-MPM form).  Parallel structure:
 *   order the particles are sorted by stencil-base cell once (z-major), so
 *         that neighbouring particles share cache lines of the grid;
 *   P2G   particles are bucketed by z-slabs of the stencil base (slab width
 *         >= 3 cells); even slabs, then odd slabs, scatter in parallel, so no
 *         two threads ever add into the same node (a particle of slab t only
 *         touches nodes [lo_t, hi_t + 2]);
 *   grid  parallel over z-planes;
 *   G2P   parallel over particles.
 * Summation order inside a node differs from the sequential step (bucket
 * order), so results agree with it to float rounding, not bit for bit.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* the per-particle and per-node arithmetic of the sequential FP32 port */
#define REAL float
#define SUF _m
#include "mlsmpm_impl.h"
#undef REAL
#undef SUF

static void run_sorted(const orc_mls_params* p, int n, float* x, float* v, float* C, float* F,
                       const float* mass, const float* E, int nsteps) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t G = (size_t)nx * ny * nz;
    const float dx = (float)p->dx, inv_dx = 1.f / dx, dt = (float)p->dt;
    const float o[3] = {(float)p->ox, (float)p->oy, (float)p->oz};
    const float gr[3] = {(float)p->gx, (float)p->gy, (float)p->gz};
    const float nu = (float)p->nu;
    const float s4 = 4.f * inv_dx * inv_dx;
    const float kscale = -dt * (float)p->p_vol * s4;
    const int dims[3] = {nx, ny, nz};
    const int T = omp_get_max_threads();
    int width = nz / (4 * T);
    if (width < 3) width = 3;
    const int nslab = (nz + width - 1) / width + 1;  /* + 1: bases below 0 or at nz land at the ends */
    float* gm = (float*)malloc(G * sizeof(float));
    float* gp = (float*)malloc(3 * G * sizeof(float));
    float* gv = (float*)malloc(3 * G * sizeof(float));
    int* slab = (int*)malloc((size_t)n * sizeof(int));
    int* order = (int*)malloc((size_t)n * sizeof(int));
    int* start = (int*)calloc((size_t)nslab + 1, sizeof(int));

    for (int step = 0; step < nsteps; ++step) {
        /* bucket particles by z-slab of the stencil base */
        memset(start, 0, ((size_t)nslab + 1) * sizeof(int));
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n; ++q) {
            int b;
            float f, w[3];
            stencil_m((x[3 * q + 2] - o[2]) * inv_dx, &b, &f, w);
            int s = b < 0 ? 0 : b / width;
            if (s >= nslab) s = nslab - 1;
            slab[q] = s;
        }
        for (int q = 0; q < n; ++q) start[slab[q] + 1]++;
        for (int s = 0; s < nslab; ++s) start[s + 1] += start[s];
        {
            int* fill = (int*)malloc((size_t)nslab * sizeof(int));
            memcpy(fill, start, (size_t)nslab * sizeof(int));
            for (int q = 0; q < n; ++q) order[fill[slab[q]]++] = q;
            free(fill);
        }
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < G; ++i) {
            gm[i] = 0.f;
            gp[3 * i] = gp[3 * i + 1] = gp[3 * i + 2] = 0.f;
        }
        for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
            for (int s = phase; s < nslab; s += 2) {
                for (int r = start[s]; r < start[s + 1]; ++r) {
                    const int q = order[r];
                    int b[3];
                    float f[3], w[3][3];
                    for (int a = 0; a < 3; ++a) stencil_m((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
                    const float Eq = E[q];
                    const float mu = Eq / (2.f * (1.f + nu));
                    const float lam = Eq * nu / ((1.f + nu) * (1.f - 2.f * nu));
                    float tau[9], A[9];
                    kirchhoff_m(p->material, mu, lam, F + 9 * (size_t)q, tau);
                    const float mq = mass[q];
                    for (int i = 0; i < 9; ++i) A[i] = kscale * tau[i] + mq * C[9 * (size_t)q + i];
                    const float mv[3] = {mq * v[3 * q], mq * v[3 * q + 1], mq * v[3 * q + 2]};
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            for (int k = 0; k < 3; ++k) {
                                const int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                                if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                                const float wt = w[0][i] * w[1][j] * w[2][k];
                                const float dp[3] = {((float)i - f[0]) * dx, ((float)j - f[1]) * dx,
                                                     ((float)k - f[2]) * dx};
                                const size_t id = ((size_t)gk * ny + gj) * nx + gi;
                                gm[id] += wt * mq;
                                for (int c = 0; c < 3; ++c)
                                    gp[3 * id + c] += wt * (mv[c] + A[3 * c] * dp[0] + A[3 * c + 1] * dp[1] +
                                                            A[3 * c + 2] * dp[2]);
                            }
                }
            }
        }
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t id = ((size_t)k * ny + j) * nx + i;
                    const int idx[3] = {i, j, k};
                    const float m = gm[id];
                    int sticky = 0;
                    for (int a = 0; a < 3; ++a)
                        if (idx[a] < p->bc || idx[a] >= dims[a] - p->bc) sticky = 1;
                    for (int c = 0; c < 3; ++c) {
                        const float vv = m > 0.f ? gp[3 * id + c] / m + dt * gr[c] : 0.f;
                        gv[3 * id + c] = sticky ? 0.f : vv;
                    }
                }
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n; ++q) {
            int b[3];
            float f[3], w[3][3];
            for (int a = 0; a < 3; ++a) stencil_m((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
            float nv[3] = {0, 0, 0}, nc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        const int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                        if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                        const float wt = w[0][i] * w[1][j] * w[2][k];
                        const float dp[3] = {((float)i - f[0]) * dx, ((float)j - f[1]) * dx, ((float)k - f[2]) * dx};
                        const size_t id = ((size_t)gk * ny + gj) * nx + gi;
                        for (int c = 0; c < 3; ++c) {
                            const float gvc = gv[3 * id + c];
                            nv[c] += wt * gvc;
                            for (int d = 0; d < 3; ++d) nc[3 * c + d] += wt * gvc * dp[d];
                        }
                    }
            for (int i = 0; i < 9; ++i) nc[i] *= s4;
            float Fn[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    float s = 0;
                    for (int k = 0; k < 3; ++k)
                        s += ((i == k ? 1.f : 0.f) + dt * nc[3 * i + k]) * F[9 * (size_t)q + 3 * k + j];
                    Fn[3 * i + j] = s;
                }
            for (int i = 0; i < 9; ++i) {
                F[9 * (size_t)q + i] = Fn[i];
                C[9 * (size_t)q + i] = nc[i];
            }
            for (int c = 0; c < 3; ++c) {
                v[3 * q + c] = nv[c];
                x[3 * q + c] += dt * nv[c];
            }
        }
    }
    free(gm); free(gp); free(gv); free(slab); free(order); free(start);
}

void orc_mls_run_f32_mt(const orc_mls_params* p, int n, float* x, float* v, float* C, float* F,
                        const float* mass, const float* E, int nsteps, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const float inv_dx = 1.f / (float)p->dx;
    const float o[3] = {(float)p->ox, (float)p->oy, (float)p->oz};
    const size_t G = (size_t)nx * ny * nz;
    /* counting sort by stencil-base cell (clamped into the grid) */
    unsigned* key = (unsigned*)malloc((size_t)n * sizeof(unsigned));
    size_t* cstart = (size_t*)calloc(G + 1, sizeof(size_t));
    int* perm = (int*)malloc((size_t)n * sizeof(int));
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) {
        int b[3];
        for (int a = 0; a < 3; ++a) {
            float f, w[3];
            stencil_m((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f, w);
        }
        const int d[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) b[a] = b[a] < 0 ? 0 : (b[a] >= d[a] ? d[a] - 1 : b[a]);
        key[q] = (unsigned)(((size_t)b[2] * ny + b[1]) * nx + b[0]);
    }
    for (int q = 0; q < n; ++q) cstart[key[q] + 1]++;
    for (size_t c = 0; c < G; ++c) cstart[c + 1] += cstart[c];
    for (int q = 0; q < n; ++q) perm[cstart[key[q]]++] = q;
    free(cstart);
    free(key);
    float* xs = (float*)malloc((size_t)n * 3 * sizeof(float));
    float* vs = (float*)malloc((size_t)n * 3 * sizeof(float));
    float* Cs = (float*)malloc((size_t)n * 9 * sizeof(float));
    float* Fs = (float*)malloc((size_t)n * 9 * sizeof(float));
    float* ms = (float*)malloc((size_t)n * sizeof(float));
    float* Es = (float*)malloc((size_t)n * sizeof(float));
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r) {
        const int q = perm[r];
        for (int c = 0; c < 3; ++c) { xs[3 * (size_t)r + c] = x[3 * (size_t)q + c]; vs[3 * (size_t)r + c] = v[3 * (size_t)q + c]; }
        for (int c = 0; c < 9; ++c) { Cs[9 * (size_t)r + c] = C[9 * (size_t)q + c]; Fs[9 * (size_t)r + c] = F[9 * (size_t)q + c]; }
        ms[r] = mass[q];
        Es[r] = E[q];
    }
    run_sorted(p, n, xs, vs, Cs, Fs, ms, Es, nsteps);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r) {
        const int q = perm[r];
        for (int c = 0; c < 3; ++c) { x[3 * (size_t)q + c] = xs[3 * (size_t)r + c]; v[3 * (size_t)q + c] = vs[3 * (size_t)r + c]; }
        for (int c = 0; c < 9; ++c) { C[9 * (size_t)q + c] = Cs[9 * (size_t)r + c]; F[9 * (size_t)q + c] = Fs[9 * (size_t)r + c]; }
    }
    free(xs); free(vs); free(Cs); free(Fs); free(ms); free(Es); free(perm);
}

int orc_num_threads(void) { return omp_get_max_threads(); }
