// C-ABI over the MPM kernels: context lifetime, particle I/O, stepping.
// Entry points and the reference interfaces they replace are listed in
// include/physgrad_mpm.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/physgrad_mpm.h"
#include "pg_kernels.h"

namespace {

thread_local std::string g_thread_error;

}  // namespace

struct pg_mpm_ctx {
    pg_mpm_params p{};
    std::string err;
    int n = 0;
    size_t cap = 0;
    int device = 0;
    double time = 0.0;
    long long steps = 0;
    bool have_mass = false, have_E = false;
    // ref-compat buffers (reference layouts)
    float *rx = nullptr, *rv = nullptr, *rm = nullptr, *rgm = nullptr, *rgp = nullptr, *rgv = nullptr;
    // MLS-MPM buffers
    float4 *S = nullptr, *S2 = nullptr;
    uint32_t *pid = nullptr, *pid2 = nullptr;
    float *E = nullptr, *E2 = nullptr, *mass = nullptr, *mass2 = nullptr;
    float4 *mp = nullptr, *gv = nullptr;
    pg::TiledWork tw{};
    pg::KernelProf prof;
    std::vector<void*> allocs;
    // differentiation: checkpoints every ck_every steps, a segment tape, adjoint buffers
    int ck_every = 0, ck_steps = 0, ck_cap = 0, tape_len = 0;
    long long version = 0, ck_version = -1;
    std::vector<pg::TapeStep> ckpt;   // S, pid, E, mass used
    std::vector<pg::TapeStep> tape;
    std::vector<void*> ck_allocs, tape_allocs;
    float4 *abar = nullptr, *part = nullptr, *agv = nullptr, *apam = nullptr;
    float *aE = nullptr, *ux = nullptr, *uv = nullptr, *uC = nullptr, *uF = nullptr;
};

namespace {

int fail(pg_mpm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    g_thread_error = msg;
    return code;
}

int hip_fail(pg_mpm_ctx* c, hipError_t e, const char* where) {
    return fail(c, PG_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

template <typename T>
int dalloc(pg_mpm_ctx* c, T** out, size_t count) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, PG_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    c->allocs.push_back(p);
    *out = static_cast<T*>(p);
    return PG_OK;
}

hipMemcpyKind kind_to_dev(int memory) { return memory == PG_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice; }
hipMemcpyKind kind_from_dev(int memory) { return memory == PG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice; }

pg::MlsGeom mls_geom(const pg_mpm_ctx* c) {
    pg::MlsGeom g;
    const pg_mpm_params& p = c->p;
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
    g.dx = p.dx; g.inv_dx = 1.0f / p.dx;
    g.o[0] = p.ox; g.o[1] = p.oy; g.o[2] = p.oz;
    g.dt = p.dt;
    g.g[0] = p.gx; g.g[1] = p.gy; g.g[2] = p.gz;
    g.bc = p.boundary_thickness;
    g.material = p.material;
    g.nu = p.poisson_ratio;
    g.E = p.youngs_modulus;
    g.p_vol = p.particle_volume;
    g.p_mass = p.density * p.particle_volume;
    return g;
}

int ceil_log2(int v) {
    int k = 0;
    while ((1 << k) < v) ++k;
    return k;
}

// Copy a caller array (host or device) into a device buffer on the stream.
int put(pg_mpm_ctx* c, void* dst, const void* src, size_t bytes, int memory, hipStream_t st) {
    if (!bytes) return PG_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind_to_dev(memory), st);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpyAsync");
    if (memory == PG_HOST) {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamSynchronize");
    }
    return PG_OK;
}

}  // namespace

extern "C" {

void pg_mpm_default_params(pg_mpm_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->mode = PG_MODE_REF_COMPAT;
    p->impl = PG_IMPL_TILED;
    p->nx = p->ny = p->nz = 64;
    p->dx = p->dy = p->dz = 1.0f / 64.0f;
    p->dt = 1e-4f;
    p->gx = 0.f; p->gy = -9.81f; p->gz = 0.f;
    p->interpolation_order = 2;
    p->material = PG_MAT_FIXED_COROTATED;
    p->youngs_modulus = 1e6f;
    p->poisson_ratio = 0.3f;
    p->density = 1000.f;
    p->particle_volume = 1.0f / (64.f * 64.f * 64.f * 8.f);
    p->boundary_thickness = 3;
    p->max_particles = 1 << 20;
}

const char* pg_mpm_last_error(const pg_mpm_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_thread_error.c_str();
}

int pg_mpm_create(const pg_mpm_params* p, pg_mpm_ctx** out) {
    if (!p || !out) return fail(nullptr, PG_ERR_INVALID, "pg_mpm_create: NULL argument");
    *out = nullptr;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(nullptr, PG_ERR_INVALID, "grid dims must be >= 1");
    if ((size_t)p->nx * p->ny * p->nz > (size_t)1 << 31) return fail(nullptr, PG_ERR_INVALID, "grid too large");
    if (!(p->dx > 0.f) || !(p->dy > 0.f) || !(p->dz > 0.f)) return fail(nullptr, PG_ERR_INVALID, "cell size must be > 0");
    if (!(p->dt > 0.f) || !std::isfinite(p->dt)) return fail(nullptr, PG_ERR_INVALID, "dt must be finite and > 0");
    if (p->max_particles < 0 || p->max_particles > (1 << 30))
        return fail(nullptr, PG_ERR_INVALID, "max_particles out of range");
    if (p->mode != PG_MODE_REF_COMPAT && p->mode != PG_MODE_MLS_MPM) return fail(nullptr, PG_ERR_INVALID, "bad mode");
    if (p->mode == PG_MODE_REF_COMPAT && (p->interpolation_order < 1 || p->interpolation_order > 3))
        return fail(nullptr, PG_ERR_INVALID, "interpolation_order must be 1, 2 or 3");
    if (p->mode == PG_MODE_MLS_MPM) {
        if (p->dx != p->dy || p->dx != p->dz) return fail(nullptr, PG_ERR_INVALID, "MLS-MPM needs a cubic cell");
        if (p->material != PG_MAT_FIXED_COROTATED && p->material != PG_MAT_NEO_HOOKEAN)
            return fail(nullptr, PG_ERR_INVALID, "bad material");
        if (p->impl != PG_IMPL_PLAIN && p->impl != PG_IMPL_TILED) return fail(nullptr, PG_ERR_INVALID, "bad impl");
        if (!(p->poisson_ratio > -1.f && p->poisson_ratio < 0.5f))
            return fail(nullptr, PG_ERR_INVALID, "poisson_ratio must be in (-1, 0.5)");
        if (!(p->particle_volume > 0.f)) return fail(nullptr, PG_ERR_INVALID, "particle_volume must be > 0");
        if (p->boundary_thickness < 0) return fail(nullptr, PG_ERR_INVALID, "boundary_thickness must be >= 0");
    }
    pg_mpm_ctx* c = new (std::nothrow) pg_mpm_ctx();
    if (!c) return fail(nullptr, PG_ERR_NOMEM, "host allocation failed");
    c->p = *p;
    c->cap = (size_t)std::max(p->max_particles, 1);
    hipGetDevice(&c->device);
    const size_t G = (size_t)p->nx * p->ny * p->nz;
    int rc = PG_OK;
    if (p->mode == PG_MODE_REF_COMPAT) {
        if ((rc = dalloc(c, &c->rx, 3 * c->cap)) || (rc = dalloc(c, &c->rv, 3 * c->cap)) ||
            (rc = dalloc(c, &c->rm, c->cap)) || (rc = dalloc(c, &c->rgm, G)) || (rc = dalloc(c, &c->rgp, 3 * G)) ||
            (rc = dalloc(c, &c->rgv, 3 * G))) {
            std::string m = c->err;
            pg_mpm_destroy(c);
            return fail(nullptr, rc, m);
        }
        hipMemset(c->rgm, 0, G * 4);
        hipMemset(c->rgp, 0, 3 * G * 4);
        hipMemset(c->rgv, 0, 3 * G * 4);
    } else {
        const bool tiled = p->impl == PG_IMPL_TILED;
        pg::TiledWork& w = c->tw;
        w.T = 4;
        w.ntx = (p->nx + w.T - 1) / w.T;
        w.nty = (p->ny + w.T - 1) / w.T;
        w.ntz = (p->nz + w.T - 1) / w.T;
        int bits = ceil_log2(std::max(w.ntx, std::max(w.nty, w.ntz)));
        w.tile_slots = 1 << (3 * bits);
        w.gbx = (p->nx + 7) / 8;
        w.gby = (p->ny + 7) / 8;
        w.gbz = (p->nz + 7) / 8;
        w.nblocks_launch = 2048;
        w.prof = &c->prof;
        if ((rc = dalloc(c, &c->S, 6 * c->cap)) || (rc = dalloc(c, &c->pid, c->cap)) ||
            (rc = dalloc(c, &c->mp, G)) || (rc = dalloc(c, &c->gv, G)) ||
            (tiled && ((rc = dalloc(c, &c->S2, 6 * c->cap)) || (rc = dalloc(c, &c->pid2, c->cap)) ||
                       (rc = dalloc(c, &w.count, (size_t)w.tile_slots)) ||
                       (rc = dalloc(c, &w.start, (size_t)w.tile_slots)) ||
                       (rc = dalloc(c, &w.active, 4 * (size_t)w.tile_slots)) || (rc = dalloc(c, &w.nactive, 1)) ||
                       (rc = dalloc(c, &w.key, c->cap)) || (rc = dalloc(c, &w.rank, c->cap)) ||
                       (rc = dalloc(c, &w.perm, c->cap)) ||
                       (rc = dalloc(c, &w.flags, (size_t)w.gbx * w.gby * w.gbz)) ||
                       (rc = dalloc(c, &w.bsum, 2 * (size_t)((w.tile_slots + 1023) / 1024)))))) {
            std::string m = c->err;
            pg_mpm_destroy(c);
            return fail(nullptr, rc, m);
        }
        hipMemset(c->mp, 0, G * sizeof(float4));
        hipMemset(c->gv, 0, G * sizeof(float4));
        if (tiled) {
            hipMemset(w.count, 0, (size_t)w.tile_slots * 4);
            hipMemset(w.flags, 0, (size_t)w.gbx * w.gby * w.gbz * 4);
            hipMemset(w.nactive, 0, 4);
        }
    }
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        pg_mpm_destroy(c);
        return hip_fail(nullptr, e, "pg_mpm_create");
    }
    *out = c;
    return PG_OK;
}

int pg_mpm_destroy(pg_mpm_ctx* c) {
    if (!c) return PG_OK;
    hipDeviceSynchronize();
    c->prof.flush();
    for (int i = 0; i < c->prof.npool; ++i) hipEventDestroy(c->prof.pool[i]);
    for (void* p : c->allocs) hipFree(p);
    for (void* p : c->ck_allocs) hipFree(p);
    for (void* p : c->tape_allocs) hipFree(p);
    delete c;
    return PG_OK;
}

int pg_mpm_num_particles(const pg_mpm_ctx* c) { return c ? c->n : -1; }

int pg_mpm_get_time(const pg_mpm_ctx* c, double* t, long long* s) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (t) *t = c->time;
    if (s) *s = c->steps;
    return PG_OK;
}

int pg_mpm_set_particles(pg_mpm_ctx* c, int n, const float* x, const float* v, const float* C, const float* F,
                         const float* mass, const float* E, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (n < 0 || (size_t)n > c->cap)
        return fail(c, PG_ERR_INVALID, "particle count " + std::to_string(n) + " exceeds max_particles " +
                                           std::to_string(c->cap));
    if (n > 0 && (!x || !v)) return fail(c, PG_ERR_INVALID, "x and v are required");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    c->n = n;
    c->time = 0.0;
    c->steps = 0;
    c->version++;
    if (c->p.mode == PG_MODE_REF_COMPAT) {
        if ((rc = put(c, c->rx, x, 12 * (size_t)n, memory, st)) || (rc = put(c, c->rv, v, 12 * (size_t)n, memory, st)))
            return rc;
        if (mass) {
            if ((rc = put(c, c->rm, mass, 4 * (size_t)n, memory, st))) return rc;
        } else {
            float m = c->p.density * c->p.particle_volume;
            hipError_t e = pg::fill_u32(reinterpret_cast<uint32_t*>(c->rm), n, *reinterpret_cast<uint32_t*>(&m), st);
            if (e != hipSuccess) return hip_fail(c, e, "fill");
        }
        return PG_OK;
    }
    // MLS-MPM: stage user arrays on the device, then pack into the float4 planes.
    float *tx = nullptr, *tv = nullptr, *tC = nullptr, *tF = nullptr;
    const float *dx_ = x, *dv_ = v, *dC_ = C, *dF_ = F;
    auto cleanup = [&]() {
        hipStreamSynchronize(st);
        hipFree(tx); hipFree(tv); hipFree(tC); hipFree(tF);
    };
    if (memory == PG_HOST && n > 0) {
        if (hipMalloc(&tx, 12 * (size_t)n) != hipSuccess || hipMalloc(&tv, 12 * (size_t)n) != hipSuccess ||
            (C && hipMalloc(&tC, 36 * (size_t)n) != hipSuccess) || (F && hipMalloc(&tF, 36 * (size_t)n) != hipSuccess)) {
            cleanup();
            return fail(c, PG_ERR_NOMEM, "staging allocation failed");
        }
        hipMemcpyAsync(tx, x, 12 * (size_t)n, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(tv, v, 12 * (size_t)n, hipMemcpyHostToDevice, st);
        if (C) hipMemcpyAsync(tC, C, 36 * (size_t)n, hipMemcpyHostToDevice, st);
        if (F) hipMemcpyAsync(tF, F, 36 * (size_t)n, hipMemcpyHostToDevice, st);
        dx_ = tx; dv_ = tv; dC_ = C ? tC : nullptr; dF_ = F ? tF : nullptr;
    }
    hipError_t e = pg::pack_state(n, c->cap, c->S, dx_, dv_, dC_, dF_, st);
    if (e == hipSuccess) e = pg::iota_u32(c->pid, c->cap, st);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(c, e, "pack");
    }
    c->have_mass = mass != nullptr;
    c->have_E = E != nullptr;
    const bool tiled = c->p.impl == PG_IMPL_TILED;
    if (mass) {
        if (!c->mass && ((rc = dalloc(c, &c->mass, c->cap)) || (tiled && (rc = dalloc(c, &c->mass2, c->cap))))) {
            cleanup();
            return rc;
        }
        if ((rc = put(c, c->mass, mass, 4 * (size_t)n, memory, st))) { cleanup(); return rc; }
    }
    if (E) {
        if (!c->E && ((rc = dalloc(c, &c->E, c->cap)) || (tiled && (rc = dalloc(c, &c->E2, c->cap))))) {
            cleanup();
            return rc;
        }
        if ((rc = put(c, c->E, E, 4 * (size_t)n, memory, st))) { cleanup(); return rc; }
    }
    cleanup();
    return PG_OK;
}

int pg_mpm_get_particles(pg_mpm_ctx* c, float* x, float* v, float* C, float* F, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)c->n;
    hipError_t e;
    if (c->p.mode == PG_MODE_REF_COMPAT) {
        if (x && (e = hipMemcpyAsync(x, c->rx, 12 * n, kind_from_dev(memory), st)) != hipSuccess) return hip_fail(c, e, "copy x");
        if (v && (e = hipMemcpyAsync(v, c->rv, 12 * n, kind_from_dev(memory), st)) != hipSuccess) return hip_fail(c, e, "copy v");
        if (C && memory == PG_HOST) std::memset(C, 0, 36 * n);
        if (F && memory == PG_HOST)
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < 9; ++k) F[9 * i + k] = (k % 4) == 0 ? 1.f : 0.f;
        if (memory == PG_HOST && (e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(c, e, "sync");
        return PG_OK;
    }
    float *tx = nullptr, *tv = nullptr, *tC = nullptr, *tF = nullptr;
    if (memory == PG_HOST && n > 0) {
        if ((x && hipMalloc(&tx, 12 * n) != hipSuccess) || (v && hipMalloc(&tv, 12 * n) != hipSuccess) ||
            (C && hipMalloc(&tC, 36 * n) != hipSuccess) || (F && hipMalloc(&tF, 36 * n) != hipSuccess)) {
            hipFree(tx); hipFree(tv); hipFree(tC); hipFree(tF);
            return fail(c, PG_ERR_NOMEM, "staging allocation failed");
        }
    }
    float* ox = memory == PG_HOST ? tx : x;
    float* ov = memory == PG_HOST ? tv : v;
    float* oC = memory == PG_HOST ? tC : C;
    float* oF = memory == PG_HOST ? tF : F;
    e = pg::unpack_state((int)n, c->cap, c->S, c->p.impl == PG_IMPL_TILED ? c->pid : nullptr, ox, ov, oC, oF, st);
    if (e == hipSuccess && memory == PG_HOST && n > 0) {
        if (x) hipMemcpyAsync(x, tx, 12 * n, hipMemcpyDeviceToHost, st);
        if (v) hipMemcpyAsync(v, tv, 12 * n, hipMemcpyDeviceToHost, st);
        if (C) hipMemcpyAsync(C, tC, 36 * n, hipMemcpyDeviceToHost, st);
        if (F) hipMemcpyAsync(F, tF, 36 * n, hipMemcpyDeviceToHost, st);
        e = hipStreamSynchronize(st);
    }
    hipFree(tx); hipFree(tv); hipFree(tC); hipFree(tF);
    if (e != hipSuccess) return hip_fail(c, e, "get_particles");
    return PG_OK;
}

int pg_mpm_get_grid(pg_mpm_ctx* c, float* mass, float* momentum, float* velocity, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    const size_t G = (size_t)c->p.nx * c->p.ny * c->p.nz;
    hipError_t e = hipSuccess;
    if (c->p.mode == PG_MODE_REF_COMPAT) {
        if (mass) e = hipMemcpyAsync(mass, c->rgm, 4 * G, kind_from_dev(memory), st);
        if (e == hipSuccess && momentum) e = hipMemcpyAsync(momentum, c->rgp, 12 * G, kind_from_dev(memory), st);
        if (e == hipSuccess && velocity) e = hipMemcpyAsync(velocity, c->rgv, 12 * G, kind_from_dev(memory), st);
        if (e == hipSuccess && memory == PG_HOST) e = hipStreamSynchronize(st);
        return e == hipSuccess ? PG_OK : hip_fail(c, e, "get_grid");
    }
    if (momentum && c->p.impl == PG_IMPL_TILED)
        return fail(c, PG_ERR_INVALID, "the tiled path clears the momentum accumulator after the grid update; "
                                       "momentum is available from the plain path only");
    float *tm = nullptr, *tp = nullptr, *tv = nullptr;
    if (memory == PG_HOST) {
        if ((mass && hipMalloc(&tm, 4 * G) != hipSuccess) || (momentum && hipMalloc(&tp, 12 * G) != hipSuccess) ||
            (velocity && hipMalloc(&tv, 12 * G) != hipSuccess)) {
            hipFree(tm); hipFree(tp); hipFree(tv);
            return fail(c, PG_ERR_NOMEM, "staging allocation failed");
        }
    }
    e = pg::grid_split(G, c->mp, c->gv, memory == PG_HOST ? tm : mass, memory == PG_HOST ? tp : momentum,
                       memory == PG_HOST ? tv : velocity, st);
    if (e == hipSuccess && memory == PG_HOST) {
        if (mass) hipMemcpyAsync(mass, tm, 4 * G, hipMemcpyDeviceToHost, st);
        if (momentum) hipMemcpyAsync(momentum, tp, 12 * G, hipMemcpyDeviceToHost, st);
        if (velocity) hipMemcpyAsync(velocity, tv, 12 * G, hipMemcpyDeviceToHost, st);
        e = hipStreamSynchronize(st);
    }
    hipFree(tm); hipFree(tp); hipFree(tv);
    return e == hipSuccess ? PG_OK : hip_fail(c, e, "get_grid");
}

int pg_mpm_step(pg_mpm_ctx* c, int nsteps, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (nsteps < 0) return fail(c, PG_ERR_INVALID, "nsteps must be >= 0");
    if (nsteps == 0) return PG_OK;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (c->p.mode == PG_MODE_REF_COMPAT) {
        pg::RefGeom g;
        g.nx = c->p.nx; g.ny = c->p.ny; g.nz = c->p.nz;
        g.o[0] = c->p.ox; g.o[1] = c->p.oy; g.o[2] = c->p.oz;
        g.h[0] = c->p.dx; g.h[1] = c->p.dy; g.h[2] = c->p.dz;
        g.order = c->p.interpolation_order;
        const float grav[3] = {c->p.gx, c->p.gy, c->p.gz};
        for (int s = 0; s < nsteps && e == hipSuccess; ++s)
            e = pg::refcompat_step(g, c->n, c->rx, c->rv, c->rm, c->rgm, c->rgp, c->rgv, c->p.dt, grav, st, &c->prof);
    } else {
        pg::MlsGeom g = mls_geom(c);
        if (c->p.impl == PG_IMPL_PLAIN) {
            for (int s = 0; s < nsteps && e == hipSuccess; ++s)
                e = pg::mls_plain_step(g, c->n, c->cap, c->S, c->have_mass ? c->mass : nullptr,
                                       c->have_E ? c->E : nullptr, c->mp, c->gv, st, &c->prof);
        } else {
            float* E = c->have_E ? c->E : nullptr;
            float* E2 = c->have_E ? c->E2 : nullptr;
            float* m = c->have_mass ? c->mass : nullptr;
            float* m2 = c->have_mass ? c->mass2 : nullptr;
            e = pg::tiled_steps(g, c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv,
                                nsteps, st);
            if (c->have_E) { c->E = E; c->E2 = E2; }
            if (c->have_mass) { c->mass = m; c->mass2 = m2; }
        }
    }
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_step");
    c->time += (double)c->p.dt * nsteps;
    c->steps += nsteps;
    c->version++;
    return PG_OK;
}

int pg_mpm_set_profiling(pg_mpm_ctx* c, int on) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    hipError_t e = c->prof.flush();
    for (int k = 0; k < pg::PK_COUNT; ++k) { c->prof.total_ms[k] = 0.0; c->prof.count[k] = 0; }
    c->prof.on = on != 0;
    return e == hipSuccess ? PG_OK : hip_fail(c, e, "profiling flush");
}

int pg_mpm_kernel_time(pg_mpm_ctx* c, int kernel, double* total_ms, long long* launches) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (kernel < 0 || kernel >= pg::PK_COUNT) return fail(c, PG_ERR_INVALID, "bad kernel id");
    hipError_t e = c->prof.flush();
    if (e != hipSuccess) return hip_fail(c, e, "profiling flush");
    if (total_ms) *total_ms = c->prof.total_ms[kernel];
    if (launches) *launches = c->prof.count[kernel];
    return PG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- differentiation
namespace {

template <typename T>
int valloc(pg_mpm_ctx* c, std::vector<void*>& owner, T** out, size_t count, bool zero) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, PG_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes for the differentiation "
                                     "buffers failed: " + hipGetErrorString(e));
    }
    if (zero && hipMemset(p, 0, bytes) != hipSuccess) return fail(c, PG_ERR_HIP, "hipMemset failed");
    owner.push_back(p);
    *out = static_cast<T*>(p);
    return PG_OK;
}

void free_all(std::vector<void*>& v) {
    for (void* p : v) hipFree(p);
    v.clear();
}

hipError_t copy_planes(float4* dst, const float4* src, size_t cap, int n, int planes, hipStream_t st) {
    for (int p = 0; p < planes; ++p) {
        hipError_t e = hipMemcpyAsync(dst + p * cap, src + p * cap, (size_t)n * sizeof(float4),
                                      hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t copy_u32(uint32_t* dst, const uint32_t* src, int n, hipStream_t st) {
    return hipMemcpyAsync(dst, src, (size_t)n * 4, hipMemcpyDeviceToDevice, st);
}

int ensure_checkpoints(pg_mpm_ctx* c, int ncp) {
    if ((int)c->ckpt.size() >= ncp && (!c->have_E || c->ckpt[0].E) && (!c->have_mass || c->ckpt[0].mass))
        return PG_OK;
    hipDeviceSynchronize();
    free_all(c->ck_allocs);
    c->ckpt.assign(ncp, pg::TapeStep{});
    int rc;
    for (auto& k : c->ckpt) {
        if ((rc = valloc(c, c->ck_allocs, &k.S, 6 * c->cap, false)) ||
            (rc = valloc(c, c->ck_allocs, &k.pid, c->cap, false)) ||
            (c->have_E && (rc = valloc(c, c->ck_allocs, &k.E, c->cap, false))) ||
            (c->have_mass && (rc = valloc(c, c->ck_allocs, &k.mass, c->cap, false)))) {
            c->ckpt.clear();
            free_all(c->ck_allocs);
            return rc;
        }
    }
    return PG_OK;
}

int ensure_tape(pg_mpm_ctx* c, int K) {
    if (c->tape_len >= K + 1 && c->abar && (!c->have_E || c->tape[0].E) && (!c->have_mass || c->tape[0].mass))
        return PG_OK;
    hipDeviceSynchronize();
    free_all(c->tape_allocs);
    c->tape.assign(K + 1, pg::TapeStep{});
    c->tape_len = 0;
    const size_t G = (size_t)c->p.nx * c->p.ny * c->p.nz;
    const pg::TiledWork& w = c->tw;
    const size_t nblk = (size_t)w.gbx * w.gby * w.gbz;
    auto& A = c->tape_allocs;
    int rc;
    for (auto& t : c->tape) {
        if ((rc = valloc(c, A, &t.S, 6 * c->cap, false)) || (rc = valloc(c, A, &t.pid, c->cap, false)) ||
            (c->have_E && (rc = valloc(c, A, &t.E, c->cap, false))) ||
            (c->have_mass && (rc = valloc(c, A, &t.mass, c->cap, false))) ||
            (rc = valloc(c, A, &t.perm, c->cap, false)) || (rc = valloc(c, A, &t.src, c->cap, false)) ||
            (rc = valloc(c, A, &t.alist, 4 * (size_t)w.tile_slots, false)) ||
            (rc = valloc(c, A, &t.nactive, 1, true)) || (rc = valloc(c, A, &t.mp, G, true)) ||
            (rc = valloc(c, A, &t.gv, G, true)) || (rc = valloc(c, A, &t.flags, nblk, true)))
            goto fail;
    }
    if ((rc = valloc(c, A, &c->abar, 6 * c->cap, false)) || (rc = valloc(c, A, &c->part, 3 * c->cap, false)) ||
        (rc = valloc(c, A, &c->agv, G, true)) || (rc = valloc(c, A, &c->apam, G, true)) ||
        (rc = valloc(c, A, &c->aE, c->cap, false)) || (rc = valloc(c, A, &c->ux, 3 * c->cap, false)) ||
        (rc = valloc(c, A, &c->uv, 3 * c->cap, false)) || (rc = valloc(c, A, &c->uC, 9 * c->cap, false)) ||
        (rc = valloc(c, A, &c->uF, 9 * c->cap, false)))
        goto fail;
    c->tape_len = K + 1;
    return PG_OK;
fail:
    c->tape.clear();
    free_all(c->tape_allocs);
    c->abar = c->part = c->agv = c->apam = nullptr;
    c->aE = c->ux = c->uv = c->uC = c->uF = nullptr;
    return rc;
}

}  // namespace

extern "C" {

int pg_mpm_forward_checkpointed(pg_mpm_ctx* c, int nsteps, int checkpoint_every, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM || c->p.impl != PG_IMPL_TILED)
        return fail(c, PG_ERR_INVALID, "checkpointed rollouts need mode MLS_MPM with the tiled implementation");
    if (nsteps <= 0 || checkpoint_every <= 0) return fail(c, PG_ERR_INVALID, "nsteps and checkpoint_every must be > 0");
    const int K = checkpoint_every;
    const int ncp = (nsteps + K - 1) / K;
    int rc;
    if ((rc = ensure_checkpoints(c, ncp)) || (rc = ensure_tape(c, K))) return rc;
    hipStream_t st = (hipStream_t)stream;
    pg::MlsGeom g = mls_geom(c);
    hipError_t e = hipSuccess;
    for (int k = 0; k < ncp && e == hipSuccess; ++k) {
        pg::TapeStep& ck = c->ckpt[k];
        e = copy_planes(ck.S, c->S, c->cap, c->n, 6, st);
        if (e == hipSuccess) e = copy_u32(ck.pid, c->pid, c->n, st);
        if (e == hipSuccess && c->have_E) e = hipMemcpyAsync(ck.E, c->E, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && c->have_mass)
            e = hipMemcpyAsync(ck.mass, c->mass, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, st);
        const int L = std::min(K, nsteps - k * K);
        float* E = c->have_E ? c->E : nullptr;
        float* E2 = c->have_E ? c->E2 : nullptr;
        float* m = c->have_mass ? c->mass : nullptr;
        float* m2 = c->have_mass ? c->mass2 : nullptr;
        if (e == hipSuccess)
            e = pg::tiled_steps(g, c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv, L, st);
        if (c->have_E) { c->E = E; c->E2 = E2; }
        if (c->have_mass) { c->mass = m; c->mass2 = m2; }
    }
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_forward_checkpointed");
    c->time += (double)c->p.dt * nsteps;
    c->steps += nsteps;
    c->version++;
    c->ck_version = c->version;
    c->ck_steps = nsteps;
    c->ck_every = K;
    return PG_OK;
}

int pg_mpm_backward(pg_mpm_ctx* c, const float* ax, const float* av, const float* aC, const float* aF, float* gx,
                    float* gv, float* gC, float* gF, float* gE, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->ck_version != c->version || c->ck_steps <= 0)
        return fail(c, PG_ERR_STATE, "pg_mpm_backward needs the state of the last pg_mpm_forward_checkpointed "
                                     "(the particles were changed or stepped since)");
    hipStream_t st = (hipStream_t)stream;
    const int n = c->n, K = c->ck_every;
    const int ncp = (c->ck_steps + K - 1) / K;
    pg::MlsGeom g = mls_geom(c);
    hipError_t e = hipSuccess;
    auto seed = [&](float* dst, const float* src, size_t count) {
        if (e != hipSuccess) return;
        if (src) e = hipMemcpyAsync(dst, src, count * 4, hipMemcpyDeviceToDevice, st);
        else e = pg::fill_u32(reinterpret_cast<uint32_t*>(dst), count, 0u, st);
    };
    seed(c->ux, ax, 3 * (size_t)n);
    seed(c->uv, av, 3 * (size_t)n);
    seed(c->uC, aC, 9 * (size_t)n);
    seed(c->uF, aF, 9 * (size_t)n);
    seed(c->aE, nullptr, (size_t)n);
    for (int k = ncp - 1; k >= 0 && e == hipSuccess; --k) {
        const int L = std::min(K, c->ck_steps - k * K);
        pg::TapeStep* T = c->tape.data();
        const pg::TapeStep& ck = c->ckpt[k];
        e = copy_planes(T[0].S, ck.S, c->cap, n, 6, st);
        if (e == hipSuccess) e = copy_u32(T[0].pid, ck.pid, n, st);
        if (e == hipSuccess && c->have_E) e = hipMemcpyAsync(T[0].E, ck.E, 4 * (size_t)n, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && c->have_mass)
            e = hipMemcpyAsync(T[0].mass, ck.mass, 4 * (size_t)n, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = pg::tiled_tape_forward(g, c->tw, n, c->cap, T, L, st);
        if (e == hipSuccess) e = pg::adjoint_from_user(n, c->cap, T[L].pid, c->ux, c->uv, c->uC, c->uF, c->abar, st);
        if (e == hipSuccess)
            e = pg::adjoint_segment(g, c->tw, n, c->cap, T, L, c->abar, c->part, c->agv, c->apam, c->aE, st);
        if (e == hipSuccess) e = pg::adjoint_to_user(n, c->cap, T[0].pid, c->abar, c->ux, c->uv, c->uC, c->uF, st);
    }
    auto out = [&](float* dst, const float* src, size_t count) {
        if (e == hipSuccess && dst) e = hipMemcpyAsync(dst, src, count * 4, hipMemcpyDeviceToDevice, st);
    };
    out(gx, c->ux, 3 * (size_t)n);
    out(gv, c->uv, 3 * (size_t)n);
    out(gC, c->uC, 9 * (size_t)n);
    out(gF, c->uF, 9 * (size_t)n);
    out(gE, c->aE, (size_t)n);
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_backward");
    return PG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- kernel timing
namespace pg {

hipEvent_t KernelProf::get() {
    if (npool > 0) return pool[--npool];
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

void KernelProf::begin(int k, hipStream_t st, hipEvent_t* out) {
    (void)k;
    if (nrec >= (int)(sizeof(recs) / sizeof(recs[0]))) flush();
    *out = get();
    hipEventRecord(*out, st);
}

void KernelProf::end(int k, hipEvent_t a, hipStream_t st) {
    hipEvent_t b = get();
    hipEventRecord(b, st);
    recs[nrec++] = Rec{k, a, b};
}

hipError_t KernelProf::flush() {
    hipError_t err = hipSuccess;
    for (int i = 0; i < nrec; ++i) {
        hipError_t e = hipEventSynchronize(recs[i].b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, recs[i].a, recs[i].b);
        if (e != hipSuccess) err = e;
        total_ms[recs[i].k] += ms;
        count[recs[i].k] += 1;
        if (npool + 2 <= (int)(sizeof(pool) / sizeof(pool[0]))) {
            pool[npool++] = recs[i].a;
            pool[npool++] = recs[i].b;
        } else {
            hipEventDestroy(recs[i].a);
            hipEventDestroy(recs[i].b);
        }
    }
    nrec = 0;
    return err;
}

}  // namespace pg
