/*
 * ref-compat oracle: restatement of the reference CPU MPM step.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows src/material_point_method.h:
 *   MPMGrid::clear                    :213-218
 *   MPMShapeFunctions linear/quadratic/cubic :277-320
 *   MPMSolver::step                   :479-500
 *   particleToGrid                    :515-570
 *   updateGridMomentum                :572-585
 *   calculateGridVelocity             :587-597
 *   gridToParticle                    :599-649
 *   updateParticlePositions           :651-666
 *   computeShapeFunction              :668-675
 * Every expression keeps the reference's operand order so that, built with
 * -ffp-contract=off, the results are bit-identical to the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* material_point_method.h:277-281 */
static float sf_linear(float x) {
    float a = x < 0.0f ? -x : x;
    if (a >= 1.0f) return 0.0f;
    return 1.0f - a;
}

/* material_point_method.h:288-293 */
static float sf_quadratic(float x) {
    float a = x < 0.0f ? -x : x;
    if (a >= 1.5f) return 0.0f;
    if (a < 0.5f) return 0.75f - x * x;
    return 0.5f * (1.5f - a) * (1.5f - a);
}

/* material_point_method.h:303-311; constants are the float roundings of 2/3 and 1/6 */
static float sf_cubic(float x) {
    float a = x < 0.0f ? -x : x;
    if (a >= 2.0f) return 0.0f;
    if (a < 1.0f) return (float)(2.0 / 3.0) - a * a + 0.5f * a * a * a;
    float t = 2.0f - a;
    return (float)(1.0 / 6.0) * t * t * t;
}

/* material_point_method.h:668-675 */
static float shape(int order, float x) {
    switch (order) {
        case 1: return sf_linear(x);
        case 2: return sf_quadratic(x);
        case 3: return sf_cubic(x);
        default: return sf_quadratic(x);
    }
}

static void ref_step(const orc_ref_params* p, int n, float* x, float* v, const float* m,
                     float* gm, float* gp, float* gv) {
    const size_t G = (size_t)p->nx * p->ny * p->nz;
    const float dt = p->dt;
    const float g[3] = {p->gx, p->gy, p->gz};
    const float o[3] = {p->ox, p->oy, p->oz};
    const float h[3] = {p->dx, p->dy, p->dz};
    const int ks = (p->order == 1) ? 2 : (p->order == 2) ? 3 : 4;
    const int off = ks / 2;
    const float thr = (float)1e-10;

    /* grid_.clear() */
    memset(gm, 0, G * sizeof(float));
    memset(gp, 0, 3 * G * sizeof(float));
    memset(gv, 0, 3 * G * sizeof(float));

    /* particleToGrid :515-570 */
    for (int q = 0; q < n; ++q) {
        const float* pos = x + 3 * (size_t)q;
        const float* vel = v + 3 * (size_t)q;
        const float mass = m[q];
        int b[3];
        float f[3];
        for (int a = 0; a < 3; ++a) b[a] = (int)floorf((pos[a] - o[a]) / h[a]);
        for (int a = 0; a < 3; ++a) f[a] = (pos[a] - o[a]) / h[a] - (float)b[a];
        for (int di = -off; di < ks - off; ++di)
            for (int dj = -off; dj < ks - off; ++dj)
                for (int dk = -off; dk < ks - off; ++dk) {
                    int gi = b[0] + di, gj = b[1] + dj, gk = b[2] + dk;
                    if (gi < 0 || gi >= p->nx || gj < 0 || gj >= p->ny || gk < 0 || gk >= p->nz) continue;
                    float wx = shape(p->order, f[0] - (float)di);
                    float wy = shape(p->order, f[1] - (float)dj);
                    float wz = shape(p->order, f[2] - (float)dk);
                    float w = wx * wy * wz;
                    if (w > thr) {
                        size_t id = (size_t)gk * p->nx * p->ny + (size_t)gj * p->nx + gi;
                        gm[id] += w * mass;
                        gp[3 * id + 0] = gp[3 * id + 0] + w * mass * vel[0];
                        gp[3 * id + 1] = gp[3 * id + 1] + w * mass * vel[1];
                        gp[3 * id + 2] = gp[3 * id + 2] + w * mass * vel[2];
                    }
                }
    }
    /* updateGridMomentum :572-585 */
    for (size_t i = 0; i < G; ++i) {
        if (gm[i] > thr) {
            gp[3 * i + 0] = gp[3 * i + 0] + gm[i] * g[0] * dt;
            gp[3 * i + 1] = gp[3 * i + 1] + gm[i] * g[1] * dt;
            gp[3 * i + 2] = gp[3 * i + 2] + gm[i] * g[2] * dt;
        }
    }
    /* calculateGridVelocity :587-597 */
    for (size_t i = 0; i < G; ++i) {
        if (gm[i] > thr) {
            gv[3 * i + 0] = gp[3 * i + 0] / gm[i];
            gv[3 * i + 1] = gp[3 * i + 1] / gm[i];
            gv[3 * i + 2] = gp[3 * i + 2] / gm[i];
        }
    }
    /* gridToParticle :599-649 */
    for (int q = 0; q < n; ++q) {
        const float* pos = x + 3 * (size_t)q;
        float nv[3] = {0.0f, 0.0f, 0.0f};
        int b[3];
        float f[3];
        for (int a = 0; a < 3; ++a) b[a] = (int)floorf((pos[a] - o[a]) / h[a]);
        for (int a = 0; a < 3; ++a) f[a] = (pos[a] - o[a]) / h[a] - (float)b[a];
        for (int di = -off; di < ks - off; ++di)
            for (int dj = -off; dj < ks - off; ++dj)
                for (int dk = -off; dk < ks - off; ++dk) {
                    int gi = b[0] + di, gj = b[1] + dj, gk = b[2] + dk;
                    if (gi < 0 || gi >= p->nx || gj < 0 || gj >= p->ny || gk < 0 || gk >= p->nz) continue;
                    float wx = shape(p->order, f[0] - (float)di);
                    float wy = shape(p->order, f[1] - (float)dj);
                    float wz = shape(p->order, f[2] - (float)dk);
                    float w = wx * wy * wz;
                    if (w > thr) {
                        size_t id = (size_t)gk * p->nx * p->ny + (size_t)gj * p->nx + gi;
                        nv[0] = nv[0] + w * gv[3 * id + 0];
                        nv[1] = nv[1] + w * gv[3 * id + 1];
                        nv[2] = nv[2] + w * gv[3 * id + 2];
                    }
                }
        v[3 * (size_t)q + 0] = nv[0];
        v[3 * (size_t)q + 1] = nv[1];
        v[3 * (size_t)q + 2] = nv[2];
    }
    /* updateParticlePositions :651-666 */
    for (int q = 0; q < n; ++q) {
        for (int a = 0; a < 3; ++a) x[3 * (size_t)q + a] = x[3 * (size_t)q + a] + v[3 * (size_t)q + a] * dt;
    }
}

void orc_ref_run(const orc_ref_params* p, int n, float* x, float* v, const float* m,
                 int nsteps, float* grid_mass, float* grid_mom, float* grid_vel) {
    const size_t G = (size_t)p->nx * p->ny * p->nz;
    float* gm = grid_mass ? grid_mass : (float*)malloc(G * sizeof(float));
    float* gp = grid_mom ? grid_mom : (float*)malloc(3 * G * sizeof(float));
    float* gv = grid_vel ? grid_vel : (float*)malloc(3 * G * sizeof(float));
    for (int s = 0; s < nsteps; ++s) ref_step(p, n, x, v, m, gm, gp, gv);
    if (!grid_mass) free(gm);
    if (!grid_mom) free(gp);
    if (!grid_vel) free(gv);
}
