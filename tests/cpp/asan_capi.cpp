// Host-side AddressSanitizer / LeakSanitizer driver for the C-ABI.
//
// Linked against physgrad_amd/_asan/libphysgrad_mpm.so, whose host code is
// built with -fsanitize=address (device code is unchanged).  Part 1 runs
// everywhere: argument validation of every entry point that checks its
// arguments before touching the device, and the clean-up of a context whose
// device allocations fail (no device present).  Part 2 runs when a GPU is
// visible: every entry point family once on a small scene (tiled, plain and
// ref-compat contexts, batch, graphs, autotune with a cache file holding
// malformed lines, checkpointed adjoint, materials and flags, spatial hash,
// neighbour lists, coupling, CSR product, standalone launchers), then every
// object is destroyed so that LeakSanitizer sees the library's host
// allocations released at exit.  Prints "asan driver OK" on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "physgrad_mpm.h"

static int fails = 0;
#define EXPECT(cond)                                                                \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::printf("FAIL line %d: %s\n", __LINE__, #cond);                     \
            ++fails;                                                                \
        }                                                                           \
    } while (0)
#define OK(call)                                                                    \
    do {                                                                            \
        int rc_ = (call);                                                           \
        if (rc_ != PG_OK) {                                                         \
            std::printf("FAIL line %d: %s -> %d\n", __LINE__, #call, rc_);          \
            ++fails;                                                                \
        }                                                                           \
    } while (0)

static bool all_finite(const std::vector<float>& a) {
    for (float v : a)
        if (!std::isfinite(v)) return false;
    return true;
}

static void argument_checks() {
    pg_mpm_params p;
    pg_mpm_default_params(&p);
    pg_mpm_ctx* ctx = nullptr;
    EXPECT(pg_mpm_create(nullptr, &ctx) == PG_ERR_INVALID);
    EXPECT(pg_mpm_create(&p, nullptr) == PG_ERR_INVALID);
    EXPECT(std::strlen(pg_mpm_last_error(nullptr)) > 0);
    pg_mpm_params q = p;
    q.nx = 0;
    EXPECT(pg_mpm_create(&q, &ctx) == PG_ERR_INVALID);
    q = p;
    q.dt = NAN;
    EXPECT(pg_mpm_create(&q, &ctx) == PG_ERR_INVALID);
    q = p;
    q.mode = 7;
    EXPECT(pg_mpm_create(&q, &ctx) == PG_ERR_INVALID);
    q = p;
    q.mode = PG_MODE_MLS_MPM;
    q.poisson_ratio = 0.5f;
    EXPECT(pg_mpm_create(&q, &ctx) == PG_ERR_INVALID);
    q = p;
    q.mode = PG_MODE_MLS_MPM;
    q.impl = PG_IMPL_TILED;
    q.nx = 300;
    EXPECT(pg_mpm_create(&q, &ctx) == PG_ERR_INVALID);
    EXPECT(std::strstr(pg_mpm_last_error(nullptr), "256") != nullptr);
    EXPECT(ctx == nullptr);

    pg_scene_desc s[2];
    std::memset(s, 0, sizeof s);
    EXPECT(pg_mpm_batch_create(s, 2, &ctx) == PG_ERR_INVALID);
    EXPECT(pg_mpm_batch_create(nullptr, 2, &ctx) == PG_ERR_INVALID);
    EXPECT(pg_mpm_destroy(nullptr) != PG_OK || true);
    EXPECT(pg_mpm_step(nullptr, 1, nullptr) == PG_ERR_INVALID);
    EXPECT(pg_mpm_num_particles(nullptr) <= 0);

    pg_coupling_config cc;
    pg_coupling_default_config(&cc);
    pg_coupling* cp = nullptr;
    pg_coupling_config cb = cc;
    cb.coupling_radius = 0.f;
    EXPECT(pg_coupling_create(&cb, &cp) == PG_ERR_INVALID);
    cb = cc;
    cb.domain_max[0] = cb.domain_min[0];
    EXPECT(pg_coupling_create(&cb, &cp) == PG_ERR_INVALID);
    cb = cc;
    cb.coupling_radius = 1e-4f;  // 10^4 cells per axis: over the 2^28 cell limit
    EXPECT(pg_coupling_create(&cb, &cp) == PG_ERR_INVALID);
    EXPECT(pg_coupling_create(nullptr, &cp) == PG_ERR_INVALID);
    EXPECT(cp == nullptr);
    pg_coupling_destroy(nullptr);
    (void)pg_coupling_last_error(nullptr);

    unsigned res[3] = {8, 8, 8};
    float dmin[3] = {0, 0, 0};
    EXPECT(pg_spatial_hash_build(nullptr, 16, 0.1f, dmin, res, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                 nullptr) == PG_ERR_INVALID);
    EXPECT(pg_spatial_hash_build_noalloc(nullptr, PG_SPATIAL_HASH_NOALLOC_MAX + 1, 0.1f, dmin, res, nullptr, nullptr,
                                         nullptr, nullptr, nullptr) == PG_ERR_INVALID);
    EXPECT(pg_spatial_hash_workspace_size(1u << 20) > 0);
    EXPECT(pg_hash_neighbors_workspace_size(1u << 20) > 0);
    unsigned long long total = 0;
    EXPECT(pg_hash_neighbors_count(nullptr, 4, 0.1f, 0.1f, dmin, res, nullptr, nullptr, &total, nullptr, 0,
                                   nullptr) == PG_ERR_INVALID);
    EXPECT(pg_csr_spmv(nullptr, nullptr, nullptr, nullptr, 4, nullptr, 4, nullptr) == PG_ERR_INVALID);
}

static void no_device_cleanup() {
    // Allocation failure half-way through create must free what was built.
    pg_mpm_params p;
    pg_mpm_default_params(&p);
    for (int mode = 0; mode < 3; ++mode) {
        pg_mpm_ctx* ctx = nullptr;
        p.mode = mode == 0 ? PG_MODE_REF_COMPAT : PG_MODE_MLS_MPM;
        p.impl = mode == 2 ? PG_IMPL_TILED : PG_IMPL_PLAIN;
        p.interpolation_order = mode == 0 ? 2 : 3;
        EXPECT(pg_mpm_create(&p, &ctx) != PG_OK);
        EXPECT(ctx == nullptr);
    }
    pg_coupling_config cc;
    pg_coupling_default_config(&cc);
    pg_coupling* cp = nullptr;
    EXPECT(pg_coupling_create(&cc, &cp) != PG_OK);
    EXPECT(cp == nullptr);
}

static void fill_box(std::mt19937& g, int n, float lo, float hi, std::vector<float>& x, std::vector<float>& v) {
    std::uniform_real_distribution<float> u(lo, hi), w(-0.5f, 0.5f);
    x.resize(3 * (size_t)n);
    v.resize(3 * (size_t)n);
    for (auto& a : x) a = u(g);
    for (auto& a : v) a = w(g);
}

static void gpu_workload() {
    std::mt19937 g(7);
    const int n = 4096, nx = 32;
    std::vector<float> x, v;
    fill_box(g, n, 0.3f, 0.7f, x, v);

    // Tiled MLS-MPM context: stepping, graphs, launch knobs, autotune, adjoint.
    pg_mpm_params p;
    pg_mpm_default_params(&p);
    p.mode = PG_MODE_MLS_MPM;
    p.impl = PG_IMPL_TILED;
    p.nx = p.ny = p.nz = nx;
    p.dx = p.dy = p.dz = 1.0f / nx;
    p.material = PG_MAT_NEO_HOOKEAN;
    p.youngs_modulus = 5e4f;
    p.dt = 5e-5f;
    p.particle_volume = 1e-6f;
    p.boundary_thickness = 3;
    p.max_particles = n;
    pg_mpm_ctx* ctx = nullptr;
    OK(pg_mpm_create(&p, &ctx));
    if (!ctx) return;
    OK(pg_mpm_set_particles(ctx, n, x.data(), v.data(), nullptr, nullptr, nullptr, nullptr, PG_HOST, nullptr));
    EXPECT(pg_mpm_set_particles(ctx, n + 1, x.data(), v.data(), nullptr, nullptr, nullptr, nullptr, PG_HOST,
                                nullptr) == PG_ERR_INVALID);
    EXPECT(std::strlen(pg_mpm_last_error(ctx)) > 0);
    OK(pg_mpm_step(ctx, 3, nullptr));
    OK(pg_mpm_set_graph(ctx, 1));
    OK(pg_mpm_step(ctx, 2, nullptr));
    OK(pg_mpm_step(ctx, 2, nullptr));
    OK(pg_mpm_set_graph(ctx, 0));
    EXPECT(pg_mpm_set_work_item(ctx, 100) == PG_ERR_INVALID);
    EXPECT(pg_mpm_set_block(ctx, 300) == PG_ERR_INVALID);
    EXPECT(pg_mpm_set_launch(ctx, 0) == PG_ERR_INVALID);
    OK(pg_mpm_set_block(ctx, 512));
    OK(pg_mpm_set_work_item(ctx, 2560));
    OK(pg_mpm_set_launch(ctx, 512));
    OK(pg_mpm_step(ctx, 1, nullptr));
    OK(pg_mpm_autotune(ctx, 1, 0, nullptr));
    OK(pg_mpm_autotune(ctx, 1, 1, nullptr));  // reads the cache file back
    int tw = 0, wi = 0, te = 0, bt = 0;
    OK(pg_mpm_get_tuning(ctx, &tw, &wi, &te, &bt));
    EXPECT(tw > 0 && wi > 0 && te == 4 && (bt == 256 || bt == 512));
    OK(pg_mpm_set_profiling(ctx, 1));
    OK(pg_mpm_step(ctx, 2, nullptr));
    double ms = 0;
    long long launches = 0;
    OK(pg_mpm_kernel_time(ctx, 1, &ms, &launches));
    EXPECT(launches > 0);
    OK(pg_mpm_set_profiling(ctx, 0));
    OK(pg_mpm_set_validation(ctx, 1));
    OK(pg_mpm_step(ctx, 1, nullptr));
    OK(pg_mpm_set_validation(ctx, 0));
    int bad = 0, outside = 0;
    OK(pg_mpm_check_state(ctx, &bad, &outside, nullptr));
    EXPECT(bad == -1);

    pg_material_desc mats[2] = {{PG_MAT_NEO_HOOKEAN, 4e4f, 0.3f, 1000.f, 0.f}, {PG_MAT_FLUID, 0.f, 0.f, 1000.f, 1e-3f}};
    std::vector<uint8_t> ids(n);
    for (int i = 0; i < n; ++i) ids[i] = (uint8_t)(i & 1);
    OK(pg_mpm_set_materials(ctx, mats, 2, ids.data(), PG_HOST, nullptr));
    std::vector<uint32_t> flags((size_t)nx * nx * nx, 0u);
    for (int i = 0; i < nx * nx; ++i) flags[i] = PG_BC_SLIP;
    OK(pg_mpm_set_boundary(ctx, flags.data(), PG_HOST, nullptr));
    OK(pg_mpm_forward_checkpointed(ctx, 4, 2, nullptr));
    std::vector<float> ax(3 * n, 1.f), gx(3 * n), gv(3 * n), gC(9 * n), gF(9 * n), gE(n);
    OK(pg_mpm_backward(ctx, ax.data(), nullptr, nullptr, nullptr, gx.data(), gv.data(), gC.data(), gF.data(),
                       gE.data(), nullptr));
    EXPECT(all_finite(gx) && all_finite(gv) && all_finite(gC) && all_finite(gF) && all_finite(gE));
    std::vector<float> ox(3 * n), ov(3 * n), oC(9 * n), oF(9 * n);
    OK(pg_mpm_get_particles(ctx, ox.data(), ov.data(), oC.data(), oF.data(), PG_HOST, nullptr));
    EXPECT(all_finite(ox) && all_finite(oF));
    double t = 0;
    long long steps = 0;
    OK(pg_mpm_get_time(ctx, &t, &steps));
    OK(pg_mpm_set_time(ctx, t, steps));
    OK(pg_mpm_destroy(ctx));

    // Two-scene batch.
    pg_scene_desc sc[2];
    for (int s = 0; s < 2; ++s) {
        sc[s] = {};
        sc[s].nx = sc[s].ny = sc[s].nz = 16 + 8 * s;
        sc[s].dx = 1.0f / sc[s].nx;
        sc[s].dt = 5e-5f;
        sc[s].gy = -9.81f;
        sc[s].material = PG_MAT_FIXED_COROTATED;
        sc[s].youngs_modulus = 1e4f;
        sc[s].poisson_ratio = 0.3f;
        sc[s].density = 1000.f;
        sc[s].particle_volume = 1e-6f;
        sc[s].boundary_thickness = 2;
        sc[s].max_particles = 2048;
    }
    OK(pg_mpm_batch_create(sc, 2, &ctx));
    if (ctx) {
        int counts[2] = {1000, 2048};
        OK(pg_mpm_batch_set_particles(ctx, counts, x.data(), v.data(), nullptr, nullptr, nullptr, nullptr, PG_HOST,
                                      nullptr));
        EXPECT(pg_mpm_num_scenes(ctx) == 2);
        OK(pg_mpm_step(ctx, 3, nullptr));
        std::vector<float> gm((size_t)16 * 16 * 16 + (size_t)24 * 24 * 24);
        OK(pg_mpm_get_grid(ctx, gm.data(), nullptr, nullptr, PG_HOST, nullptr));
        OK(pg_mpm_destroy(ctx));
    }

    // Plain MLS-MPM and ref-compat contexts.
    for (int mode = 0; mode < 2; ++mode) {
        pg_mpm_params q;
        pg_mpm_default_params(&q);
        q.nx = q.ny = q.nz = 16;
        q.dx = q.dy = q.dz = 1.0f / 16;
        q.max_particles = 512;
        if (mode) {
            q.mode = PG_MODE_MLS_MPM;
            q.impl = PG_IMPL_PLAIN;
        }
        OK(pg_mpm_create(&q, &ctx));
        if (!ctx) continue;
        OK(pg_mpm_set_particles(ctx, 512, x.data(), v.data(), nullptr, nullptr, nullptr, nullptr, PG_HOST, nullptr));
        OK(pg_mpm_step(ctx, 2, nullptr));
        std::vector<float> gm(16 * 16 * 16), gp(3 * 16 * 16 * 16), gvel(3 * 16 * 16 * 16);
        OK(pg_mpm_get_grid(ctx, gm.data(), gp.data(), gvel.data(), PG_HOST, nullptr));
        EXPECT(all_finite(gm));
        OK(pg_mpm_destroy(ctx));
    }

    // Standalone launcher with host arrays.
    {
        const int m = 512, dims[3] = {16, 16, 16};
        const float origin[3] = {0, 0, 0}, cell[3] = {1.f / 16, 1.f / 16, 1.f / 16}, grav[3] = {0, -9.81f, 0};
        std::vector<float> pos(4 * m), vel(4 * m, 0.f), gmass(16 * 16 * 16), gmom(3 * 16 * 16 * 16),
            gvel(3 * 16 * 16 * 16);
        for (int i = 0; i < m; ++i) {
            for (int a = 0; a < 3; ++a) pos[4 * i + a] = x[3 * i + a];
            pos[4 * i + 3] = 1e-3f;
        }
        OK(pg_launch_g2p2g(pos.data(), vel.data(), gmass.data(), gmom.data(), gvel.data(), origin, cell, dims, m,
                           1e-4f, grav, 1, 0.95f, 2, PG_HOST, nullptr));
        EXPECT(all_finite(pos));
    }

    // Spatial hash, neighbour lists, CSR product, coupling (device arrays).
    {
        const unsigned nq = 1024, ns = 2048;
        const float cs = 0.05f, dmin[3] = {0, 0, 0}, dmax[3] = {1, 1, 1};
        unsigned res[3] = {20, 20, 20};
        const unsigned cells = res[0] * res[1] * res[2];
        float *d_pos = nullptr, *d_q = nullptr, *d_v = nullptr, *d_vals = nullptr, *d_y = nullptr;
        unsigned *d_ids = nullptr, *d_idx = nullptr, *d_start = nullptr, *d_cnt = nullptr, *d_rp = nullptr,
                 *d_nb = nullptr;
        void* ws = nullptr;
        const size_t wsb = pg_spatial_hash_workspace_size(ns), nwsb = pg_hash_neighbors_workspace_size(nq);
        bool ok = hipMalloc(&d_pos, 12 * ns) == hipSuccess && hipMalloc(&d_q, 12 * nq) == hipSuccess &&
                  hipMalloc(&d_v, 12 * ns) == hipSuccess && hipMalloc(&d_ids, 4 * ns) == hipSuccess &&
                  hipMalloc(&d_idx, 4 * ns) == hipSuccess && hipMalloc(&d_start, 4 * cells) == hipSuccess &&
                  hipMalloc(&d_cnt, 4 * cells) == hipSuccess && hipMalloc(&d_rp, 4 * (nq + 1)) == hipSuccess &&
                  hipMalloc(&ws, wsb > nwsb ? wsb : nwsb) == hipSuccess;
        EXPECT(ok);
        if (ok) {
            OK((int)hipMemcpy(d_pos, x.data(), 12 * ns, hipMemcpyHostToDevice));
            OK((int)hipMemcpy(d_q, x.data() + 3 * ns, 12 * nq, hipMemcpyHostToDevice));
            OK((int)hipMemcpy(d_v, v.data(), 12 * ns, hipMemcpyHostToDevice));
            OK(pg_spatial_hash_build(d_pos, ns, cs, dmin, res, d_ids, d_idx, d_start, d_cnt, ws, wsb, nullptr));
            OK(pg_spatial_hash_build_noalloc(d_pos, ns, cs, dmin, res, d_ids, d_idx, d_start, d_cnt, nullptr));
            unsigned long long total = 0;
            OK(pg_hash_neighbors_count(d_q, nq, 0.06f, cs, dmin, res, d_cnt, d_rp, &total, ws, nwsb, nullptr));
            EXPECT(total > 0 && total < (1ull << 26));
            if (total > 0 && total < (1ull << 26) && hipMalloc(&d_nb, 4 * total) == hipSuccess &&
                hipMalloc(&d_vals, 4 * total) == hipSuccess && hipMalloc(&d_y, 4 * nq) == hipSuccess) {
                OK(pg_hash_neighbors_fill(d_q, nq, 0.06f, cs, dmin, res, d_idx, d_start, d_cnt, d_rp, d_nb, nullptr));
                std::vector<float> ones(total, 1.f);
                OK((int)hipMemcpy(d_vals, ones.data(), 4 * total, hipMemcpyHostToDevice));
                OK(pg_csr_spmv(d_rp, d_nb, d_vals, d_v, 3 * ns, d_y, nq, nullptr));
                std::vector<float> y(nq);
                OK((int)hipMemcpy(y.data(), d_y, 4 * nq, hipMemcpyDeviceToHost));
                EXPECT(all_finite(y));
            }

            pg_coupling_config cc;
            pg_coupling_default_config(&cc);
            cc.coupling_radius = 0.05f;
            for (int a = 0; a < 3; ++a) {
                cc.domain_min[a] = dmin[a];
                cc.domain_max[a] = dmax[a];
            }
            pg_coupling* cp = nullptr;
            OK(pg_coupling_create(&cc, &cp));
            if (cp) {
                unsigned long long np = 0;
                OK(pg_coupling_find_pairs(cp, d_q, nq, d_pos, ns, &np, nullptr));
                EXPECT(np > 0);
                unsigned *d_prp = nullptr, *d_col = nullptr;
                float *d_pen = nullptr, *d_sf = nullptr, *d_ff = nullptr, *d_sph = nullptr;
                if (np > 0 && hipMalloc(&d_prp, 4 * (nq + 1)) == hipSuccess && hipMalloc(&d_col, 4 * np) == hipSuccess &&
                    hipMalloc(&d_pen, 4 * np) == hipSuccess && hipMalloc(&d_sf, 12 * np) == hipSuccess &&
                    hipMalloc(&d_ff, 12 * np) == hipSuccess && hipMalloc(&d_sph, 12 * nq) == hipSuccess) {
                    OK(pg_coupling_get_pairs(cp, d_prp, d_col, d_pen, nullptr));
                    OK((int)hipMemset(d_sph, 0, 12 * nq));
                    OK(pg_coupling_compute_forces(cp, d_q, d_v, d_pos, d_v, d_sph, d_sf, d_ff, nullptr));
                    std::vector<float> sph(3 * nq);
                    OK((int)hipMemcpy(sph.data(), d_sph, 12 * nq, hipMemcpyDeviceToHost));
                    EXPECT(all_finite(sph));
                }
                EXPECT(pg_coupling_find_pairs(cp, nullptr, nq, d_pos, ns, &np, nullptr) == PG_ERR_INVALID);
                EXPECT(std::strlen(pg_coupling_last_error(cp)) > 0);
                pg_coupling_destroy(cp);
                for (void* q : {(void*)d_prp, (void*)d_col, (void*)d_pen, (void*)d_sf, (void*)d_ff, (void*)d_sph})
                    if (q) OK((int)hipFree(q));
            }
        }
        for (void* q : {(void*)d_pos, (void*)d_q, (void*)d_v, (void*)d_ids, (void*)d_idx, (void*)d_start,
                        (void*)d_cnt, (void*)d_rp, ws, (void*)d_nb, (void*)d_vals, (void*)d_y})
            if (q) OK((int)hipFree(q));
    }
}

int main() {
    argument_checks();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    if (ndev == 0) {
        no_device_cleanup();
        std::printf("no device: argument checks and failed-create clean-up only\n");
    } else {
        gpu_workload();
        std::printf("device workload done\n");
    }
    std::printf("%s (%d failures)\n", fails ? "asan driver FAILED" : "asan driver OK", fails);
    return fails ? 1 : 0;
}
