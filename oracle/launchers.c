/*
 * Oracle of the reference's standalone transfer launchers.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/mpm_g2p2g_kernels.cu: evalQuadraticGPU/evalCubicGPU (:27-66),
 * optimizedP2GKernel (:72-157), optimizedG2PKernel (:163-271) and the phases
 * of fullyFusedG2P2GKernel (:277-500), particle by particle in index order,
 * and the header launcher G2P2GKernelLauncher<T>::launchG2P2G
 * (src/mpm_g2p2g_kernels.h:494-547).
 * The CUDA file does not compile and no NVIDIA toolchain exists here, so
 * these are pinned by closed forms (tests/test_oracle_launchers.py), not by
 * an executable reference: parity with the CUDA source is "unpinned" beyond
 * those checks.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static float lw(int order, float x) {
    float a = fabsf(x);
    if (order == 2) {
        if (a < 0.5f) return 0.75f - x * x;
        if (a < 1.5f) {
            float t = 1.5f - a;
            return 0.5f * t * t;
        }
        return 0.0f;
    }
    if (a < 1.0f) {
        float a2 = a * a, a3 = a2 * a;
        return (float)(2.0 / 3.0) - a2 + 0.5f * a3;
    }
    if (a < 2.0f) {
        float t = 2.0f - a, t2 = t * t, t3 = t2 * t;
        return (float)(1.0 / 6.0) * t3;
    }
    return 0.0f;
}

static void lbase(const float* p, const float* o, const float* h, int* b, float* f) {
    for (int a = 0; a < 3; ++a) b[a] = (int)floorf((p[a] - o[a]) / h[a]);
    for (int a = 0; a < 3; ++a) f[a] = (p[a] - o[a]) / h[a] - (float)b[a];
}

static void p2g(const float* pos, const float* vel, float* gm, float* gp, const float* o, const float* h,
                const int* d, int n, int order, const float* gdt) {
    const int ks = order == 2 ? 3 : 4, off = ks / 2;
    for (int q = 0; q < n; ++q) {
        const float* pm = pos + 4 * (size_t)q;
        const float* vp = vel + 4 * (size_t)q;
        float mass = pm[3];
        float mom[3] = {(vp[0] + gdt[0]) * mass, (vp[1] + gdt[1]) * mass, (vp[2] + gdt[2]) * mass};
        int b[3];
        float f[3];
        lbase(pm, o, h, b, f);
        for (int di = -off; di < ks - off; ++di)
            for (int dj = -off; dj < ks - off; ++dj)
                for (int dk = -off; dk < ks - off; ++dk) {
                    int i = b[0] + di, j = b[1] + dj, k = b[2] + dk;
                    if (i < 0 || i >= d[0] || j < 0 || j >= d[1] || k < 0 || k >= d[2]) continue;
                    float w = lw(order, f[0] - di) * lw(order, f[1] - dj) * lw(order, f[2] - dk);
                    if (w > 1e-10f) {
                        size_t id = (size_t)i + (size_t)j * d[0] + (size_t)k * d[0] * d[1];
                        gm[id] += mass * w;
                        for (int c = 0; c < 3; ++c) gp[3 * id + c] += mom[c] * w;
                    }
                }
    }
}

static void g2p(float* pos, float* vel, const float* gm, const float* gv, const float* o, const float* h,
                const int* d, int n, float dt, float flip, int order, int normalise) {
    const int ks = order == 2 ? 3 : 4, off = ks / 2;
    for (int q = 0; q < n; ++q) {
        float* pm = pos + 4 * (size_t)q;
        float* vo = vel + 4 * (size_t)q;
        int b[3];
        float f[3];
        lbase(pm, o, h, b, f);
        float nv[3] = {0, 0, 0}, tw = 0.0f;
        for (int di = -off; di < ks - off; ++di)
            for (int dj = -off; dj < ks - off; ++dj)
                for (int dk = -off; dk < ks - off; ++dk) {
                    int i = b[0] + di, j = b[1] + dj, k = b[2] + dk;
                    if (i < 0 || i >= d[0] || j < 0 || j >= d[1] || k < 0 || k >= d[2]) continue;
                    float w = lw(order, f[0] - di) * lw(order, f[1] - dj) * lw(order, f[2] - dk);
                    if (w > 1e-10f) {
                        size_t id = (size_t)i + (size_t)j * d[0] + (size_t)k * d[0] * d[1];
                        if (gm[id] > 1e-10f) {
                            for (int c = 0; c < 3; ++c) nv[c] += w * gv[3 * id + c];
                            tw += w;
                        }
                    }
                }
        if (normalise && tw > 1e-10f)
            for (int c = 0; c < 3; ++c) nv[c] /= tw;
        float bv[3];
        for (int c = 0; c < 3; ++c) bv[c] = vo[c] * flip + nv[c] * (1.0f - flip);
        for (int c = 0; c < 3; ++c) pm[c] += bv[c] * dt;
        for (int c = 0; c < 3; ++c) vo[c] = bv[c];
        vo[3] = 0.0f;
    }
}

void orc_launch_p2g(const float* pos, const float* vel, float* gm, float* gp, const float* o, const float* h,
                    const int* d, int n, int order) {
    const float z[3] = {0, 0, 0};
    p2g(pos, vel, gm, gp, o, h, d, n, order, z);
}

void orc_launch_g2p(float* pos, float* vel, const float* gm, const float* gv, const float* o, const float* h,
                    const int* d, int n, float dt, float flip, int order) {
    g2p(pos, vel, gm, gv, o, h, d, n, dt, flip, order, 1);
}

void orc_launch_fused(float* pos, float* vel, float* gm, float* gp, float* gv, const float* o, const float* h,
                      const int* d, int n, float dt, const float* grav, float flip, int order) {
    size_t G = (size_t)d[0] * d[1] * d[2];
    memset(gm, 0, G * sizeof(float));
    memset(gp, 0, 3 * G * sizeof(float));
    memset(gv, 0, 3 * G * sizeof(float));
    float gdt[3] = {grav[0] * dt, grav[1] * dt, grav[2] * dt};
    p2g(pos, vel, gm, gp, o, h, d, n, order, gdt);
    for (size_t i = 0; i < G; ++i)
        if (gm[i] > 1e-10f)
            for (int c = 0; c < 3; ++c) gv[3 * i + c] = gp[3 * i + c] / gm[i];
    g2p(pos, vel, gm, gv, o, h, d, n, dt, flip, order, 1);
}

/* G2P2GKernelLauncher<T>::launchG2P2G, src/mpm_g2p2g_kernels.h:494-547:
 * clear (:520-521), optimizedP2GKernel with gravity on the particle velocity
 * (:343-413), calculateGridVelocityKernel (:419-437), fusedG2P2GKernel's
 * un-normalised gather and PIC / blended update (:211-335). */
void orc_launch_g2p2g(float* pos, float* vel, float* gm, float* gp, float* gv, const float* o, const float* h,
                      const int* d, int n, float dt, const float* grav, int use_affine_pic, float flip, int order) {
    size_t G = (size_t)d[0] * d[1] * d[2];
    memset(gm, 0, G * sizeof(float));
    memset(gp, 0, 3 * G * sizeof(float));
    memset(gv, 0, 3 * G * sizeof(float));
    float gdt[3] = {grav[0] * dt, grav[1] * dt, grav[2] * dt};
    p2g(pos, vel, gm, gp, o, h, d, n, order, gdt);
    for (size_t i = 0; i < G; ++i)
        if (gm[i] > 1e-10f)
            for (int c = 0; c < 3; ++c) gv[3 * i + c] = gp[3 * i + c] / gm[i];
    g2p(pos, vel, gm, gv, o, h, d, n, dt, use_affine_pic ? flip : 0.0f, order, 0);
}
