/*
 * MLS-MPM forward step, written once and instantiated for double (the FP64
 * oracle) and float (the FP32 CPU port timed beside the GPU).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Included by mlsmpm.c with REAL and SUF defined.  Semantics:
 *   stencil      base = floor((x-o)/dx - 0.5), f = (x-o)/dx - base in [0.5,1.5)
 *   weights      w0 = (1.5-f)^2/2, w1 = 3/4-(f-1)^2, w2 = (f-0.5)^2/2 (quadratic B-spline,
 *                the same kernel as src/material_point_method.h:288-293 on the
 *                correct three nodes)
 *   P2G          m_i += w m;  p_i += w (m v + A (x_i - x_p)),
 *                A = -dt V0 (4/dx^2) tau(F) + m C   (Hu et al. 2018, MLS-MPM)
 *   grid         v_i = p_i/m_i + dt g for m_i > 0, else 0; sticky walls: v_i = 0
 *                when any node index < bc or >= n - bc
 *   G2P          v = sum w v_i;  C = (4/dx^2) sum w v_i (x_i - x_p)^T;
 *                F <- (I + dt C) F;  x <- x + dt v
 *   elasticity   tau = P F^T.  Fixed-corotated (Stomakhin et al. 2012):
 *                tau = 2 mu (F - R) F^T + lam (J-1) J I, R the polar factor.
 *                Neo-Hookean: tau = J sigma with sigma the Cauchy stress of
 *                src/mpm_kernels.cu:97-128, i.e. tau = mu (F F^T - I) + lam ln(J) I.
 *   mu = E / (2(1+nu)), lam = E nu / ((1+nu)(1-2nu))   (src/mpm_kernels.cu:112-114)
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

static REAL FN(det3)(const REAL* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

/* inverse transpose via the cofactor matrix: inv(A)^T = cof(A) / det(A) */
static void FN(inv_t3)(const REAL* A, REAL* out) {
    REAL c[9];
    c[0] = A[4] * A[8] - A[5] * A[7];
    c[1] = A[5] * A[6] - A[3] * A[8];
    c[2] = A[3] * A[7] - A[4] * A[6];
    c[3] = A[2] * A[7] - A[1] * A[8];
    c[4] = A[0] * A[8] - A[2] * A[6];
    c[5] = A[1] * A[6] - A[0] * A[7];
    c[6] = A[1] * A[5] - A[2] * A[4];
    c[7] = A[2] * A[3] - A[0] * A[5];
    c[8] = A[0] * A[4] - A[1] * A[3];
    REAL d = A[0] * c[0] + A[1] * c[1] + A[2] * c[2];
    REAL id = (REAL)1 / d;
    for (int i = 0; i < 9; ++i) out[i] = c[i] * id;
}

/* Orthogonal polar factor by Newton's iteration X <- (X + X^-T)/2 (Higham 1986).
 * Converges quadratically to R for det F > 0. */
static int FN(polar)(const REAL* F, REAL* R) {
    REAL X[9], Y[9];
    for (int i = 0; i < 9; ++i) X[i] = F[i];
    const REAL tol = sizeof(REAL) == 8 ? (REAL)1e-15 : (REAL)1e-6;
    int it = 0;
    for (; it < 64; ++it) {
        FN(inv_t3)(X, Y);
        REAL diff = 0, nrm = 0;
        for (int i = 0; i < 9; ++i) {
            REAL nx = (REAL)0.5 * (X[i] + Y[i]);
            diff += (nx - X[i]) * (nx - X[i]);
            nrm += nx * nx;
            X[i] = nx;
        }
        if (diff <= tol * tol * nrm) { ++it; break; }
    }
    for (int i = 0; i < 9; ++i) R[i] = X[i];
    return it;
}

static void FN(kirchhoff)(int material, REAL mu, REAL lam, const REAL* F, REAL* tau) {
    REAL J = FN(det3)(F);
    if (material == ORC_NEO_HOOKEAN) {
        REAL lnJ = (REAL)log((double)J);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                REAL b = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
                tau[3 * i + j] = mu * (b - (i == j ? (REAL)1 : (REAL)0)) + (i == j ? lam * lnJ : (REAL)0);
            }
    } else {
        REAL R[9];
        FN(polar)(F, R);
        REAL vol = lam * (J - (REAL)1) * J;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                REAL s = 0;
                for (int k = 0; k < 3; ++k) s += (F[3 * i + k] - R[3 * i + k]) * F[3 * j + k];
                tau[3 * i + j] = (REAL)2 * mu * s + (i == j ? vol : (REAL)0);
            }
    }
}

static void FN(stencil)(REAL xs, int* base, REAL* f, REAL* w) {
    int b = (int)floor((double)(xs - (REAL)0.5));
    REAL fx = xs - (REAL)b;
    *base = b;
    *f = fx;
    w[0] = (REAL)0.5 * ((REAL)1.5 - fx) * ((REAL)1.5 - fx);
    w[1] = (REAL)0.75 - (fx - (REAL)1) * (fx - (REAL)1);
    w[2] = (REAL)0.5 * (fx - (REAL)0.5) * (fx - (REAL)0.5);
}

static void FN(mls_step)(const orc_mls_params* p, int n, REAL* x, REAL* v, REAL* C, REAL* F,
                         const REAL* mass, const REAL* E, REAL* gm, REAL* gp, REAL* gv) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t G = (size_t)nx * ny * nz;
    const REAL dx = (REAL)p->dx, inv_dx = (REAL)1 / dx, dt = (REAL)p->dt;
    const REAL o[3] = {(REAL)p->ox, (REAL)p->oy, (REAL)p->oz};
    const REAL g[3] = {(REAL)p->gx, (REAL)p->gy, (REAL)p->gz};
    const REAL nu = (REAL)p->nu;
    const REAL s4 = (REAL)4 * inv_dx * inv_dx;
    const REAL kscale = -dt * (REAL)p->p_vol * s4;
    const int dims[3] = {nx, ny, nz};

    memset(gm, 0, G * sizeof(REAL));
    memset(gp, 0, 3 * G * sizeof(REAL));

    for (int q = 0; q < n; ++q) {
        int b[3];
        REAL f[3], w[3][3];
        for (int a = 0; a < 3; ++a) FN(stencil)((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
        const REAL Eq = E[q];
        const REAL mu = Eq / ((REAL)2 * ((REAL)1 + nu));
        const REAL lam = Eq * nu / (((REAL)1 + nu) * ((REAL)1 - (REAL)2 * nu));
        REAL tau[9], A[9];
        FN(kirchhoff)(p->material, mu, lam, F + 9 * q, tau);
        const REAL mq = mass[q];
        for (int i = 0; i < 9; ++i) A[i] = kscale * tau[i] + mq * C[9 * q + i];
        const REAL mv[3] = {mq * v[3 * q], mq * v[3 * q + 1], mq * v[3 * q + 2]};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                    REAL wt = w[0][i] * w[1][j] * w[2][k];
                    REAL dp[3] = {((REAL)i - f[0]) * dx, ((REAL)j - f[1]) * dx, ((REAL)k - f[2]) * dx};
                    size_t id = ((size_t)gk * ny + gj) * nx + gi;
                    gm[id] += wt * mq;
                    for (int c = 0; c < 3; ++c)
                        gp[3 * id + c] += wt * (mv[c] + A[3 * c] * dp[0] + A[3 * c + 1] * dp[1] + A[3 * c + 2] * dp[2]);
                }
    }

    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t id = ((size_t)k * ny + j) * nx + i;
                const int idx[3] = {i, j, k};
                REAL m = gm[id];
                int sticky = 0;
                for (int a = 0; a < 3; ++a)
                    if (idx[a] < p->bc || idx[a] >= dims[a] - p->bc) sticky = 1;
                for (int c = 0; c < 3; ++c) {
                    REAL vv = m > (REAL)0 ? gp[3 * id + c] / m + dt * g[c] : (REAL)0;
                    gv[3 * id + c] = sticky ? (REAL)0 : vv;
                }
            }

    for (int q = 0; q < n; ++q) {
        int b[3];
        REAL f[3], w[3][3];
        for (int a = 0; a < 3; ++a) FN(stencil)((x[3 * q + a] - o[a]) * inv_dx, &b[a], &f[a], w[a]);
        REAL nv[3] = {0, 0, 0}, nc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
                    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || gk < 0 || gk >= nz) continue;
                    REAL wt = w[0][i] * w[1][j] * w[2][k];
                    REAL dp[3] = {((REAL)i - f[0]) * dx, ((REAL)j - f[1]) * dx, ((REAL)k - f[2]) * dx};
                    size_t id = ((size_t)gk * ny + gj) * nx + gi;
                    for (int c = 0; c < 3; ++c) {
                        REAL gvc = gv[3 * id + c];
                        nv[c] += wt * gvc;
                        for (int d = 0; d < 3; ++d) nc[3 * c + d] += wt * gvc * dp[d];
                    }
                }
        for (int i = 0; i < 9; ++i) nc[i] *= s4;
        REAL Fn[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                REAL s = 0;
                for (int k = 0; k < 3; ++k) s += ((i == k ? (REAL)1 : (REAL)0) + dt * nc[3 * i + k]) * F[9 * q + 3 * k + j];
                Fn[3 * i + j] = s;
            }
        for (int i = 0; i < 9; ++i) { F[9 * q + i] = Fn[i]; C[9 * q + i] = nc[i]; }
        for (int c = 0; c < 3; ++c) {
            v[3 * q + c] = nv[c];
            x[3 * q + c] += dt * nv[c];
        }
    }
}

#undef FN
#undef CAT
#undef CAT2
