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

#include "physgrad_mpm.h"
#include "pg_device.h"
#include "pg_kernels.h"
#include "pg_tile.h"

namespace {

thread_local std::string g_thread_error;

}  // namespace

void pg::set_thread_error(const char* msg) { g_thread_error = msg; }

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
    // scenes (one for pg_mpm_create, many for pg_mpm_batch_create)
    std::vector<pg::SceneInfo> scenes_h;
    pg::SceneInfo* scenes_d = nullptr;
    int nscenes = 1;
    size_t G = 0;   // grid nodes over all scenes
    pg::TiledWork tw{};
    pg::KernelProf prof;
    std::vector<void*> allocs;
    // differentiation: checkpoints every ck_every steps, a segment tape, adjoint buffers
    int ck_every = 0, ck_steps = 0, ck_cap = 0, tape_len = 0;
    long long version = 0, ck_version = -1;
    // the tape holds the last segment of the last checkpointed forward (tiled)
    int tape_seg = -1;
    long long tape_version = -1;
    int tape_filled = 0;   // steps of the tape holding unreversed grids
    std::vector<pg::TapeStep> ckpt;   // S, pid, E, mass used
    std::vector<pg::TapeStep> tape;
    // tiled: a second segment tape, so that the backward recomputes one segment
    // on its own stream while it reverses the other, with the recompute's own
    // binning scratch and window buffers (rw: those fields only)
    std::vector<pg::TapeStep> tape2;
    pg::TiledWork rw{};
    hipStream_t bstream = nullptr;
    hipEvent_t bev[7] = {};     // backward start, segment recomputed (per tape), segment reversed (per tape);
                                // forward: state ready for its checkpoint, checkpoint copied
    std::vector<void*> ck_allocs, tape_allocs;
    // hipGraph replay of pg_mpm_step sequences: key (nsteps, state buffer) -> executable graph
    bool use_graph = false;
    // the last tiled step left the current state's P2G and binning counts pending
    // for the next step (consumed by it, or by settle() before anything else)
    bool primed = false;
    struct GraphEntry { int nsteps; const void* S; bool primed; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs;
    hipStream_t gstream = nullptr;   // capture stream for callers on the NULL stream
    float4 *abar = nullptr, *part = nullptr, *agv = nullptr, *apam = nullptr;
    float *aE = nullptr, *ux = nullptr, *uv = nullptr, *uC = nullptr, *uF = nullptr;
    float *aEs = nullptr, *aEp = nullptr;   // dL/dE by storage slot through a segment, and its carry
    uint32_t* adst = nullptr;               // storage slot -> output position of a reverse step
    // second set of per-step reverse buffers (the tiled reverse alternates between the two)
    float4* part2 = nullptr;
    float* aEp2 = nullptr;
    uint32_t* adst2 = nullptr;
    uint32_t* ainv = nullptr;               // caller index -> slot of the order the adjoint is carried in
    uint32_t* check = nullptr;              // [4] results of the state validation kernel
    bool validate = false;                  // pg_mpm_step validates the state after stepping
    // per-particle materials (pg_mpm_set_materials) and per-node boundary flags (pg_mpm_set_boundary)
    pg::MatRec* mats_d = nullptr;           // [MAX_MATERIALS]
    uint8_t* mat_ids = nullptr;             // [cap] material per caller index
    int nmat = 0;                           // 0: every particle has its scene's material
    uint32_t* node_bc = nullptr;            // [G]
    bool have_bc = false;
    pg::MatTable mat_table() const {
        return pg::MatTable{nmat > 0 ? mats_d : nullptr, nmat > 0 ? mat_ids : nullptr};
    }
    const uint32_t* bc_flags() const { return have_bc ? node_bc : nullptr; }
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

pg::MlsGeom mls_geom(const pg_mpm_ctx* c) { return c->scenes_h[0].g; }

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

pg_scene_desc scene_from_params(const pg_mpm_params& p) {
    pg_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
    d.dx = p.dx;
    d.ox = p.ox; d.oy = p.oy; d.oz = p.oz;
    d.dt = p.dt;
    d.gx = p.gx; d.gy = p.gy; d.gz = p.gz;
    d.material = p.material;
    d.youngs_modulus = p.youngs_modulus;
    d.poisson_ratio = p.poisson_ratio;
    d.density = p.density;
    d.particle_volume = p.particle_volume;
    d.boundary_thickness = p.boundary_thickness;
    d.max_particles = std::max(p.max_particles, 1);
    return d;
}

const char* check_scene(const pg_scene_desc& d) {
    if (d.nx < 1 || d.ny < 1 || d.nz < 1) return "grid dims must be >= 1";
    if ((size_t)d.nx * d.ny * d.nz > (size_t)1 << 31) return "grid too large";
    if (!(d.dx > 0.f)) return "cell size must be > 0";
    if (!(d.dt > 0.f) || !std::isfinite(d.dt)) return "dt must be finite and > 0";
    if (d.material != PG_MAT_FIXED_COROTATED && d.material != PG_MAT_NEO_HOOKEAN) return "bad material";
    if (!(d.poisson_ratio > -1.f && d.poisson_ratio < 0.5f)) return "poisson_ratio must be in (-1, 0.5)";
    if (!(d.particle_volume > 0.f)) return "particle_volume must be > 0";
    if (d.boundary_thickness < 0) return "boundary_thickness must be >= 0";
    if (d.max_particles < 0 || d.max_particles > (1 << 30)) return "max_particles out of range";
    if (std::max(d.nx, std::max(d.ny, d.nz)) > pg::MAX_TILED_NODES) return "more than 256 nodes per axis";
    return nullptr;
}

pg::MlsGeom geom_of(const pg_scene_desc& d) {
    pg::MlsGeom g;
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz;
    g.dx = d.dx; g.inv_dx = 1.0f / d.dx;
    g.o[0] = d.ox; g.o[1] = d.oy; g.o[2] = d.oz;
    g.dt = d.dt;
    g.g[0] = d.gx; g.g[1] = d.gy; g.g[2] = d.gz;
    g.bc = d.boundary_thickness;
    g.material = d.material;
    g.nu = d.poisson_ratio;
    g.E = d.youngs_modulus;
    g.p_vol = d.particle_volume;
    g.p_mass = d.density * d.particle_volume;
    return g;
}

// Builds the scene table (tile ids, grid nodes and 8^3 blocks concatenated in
// scene order) and allocates the MLS-MPM buffers of a context.
int init_mls(pg_mpm_ctx* c, const pg_scene_desc* d, int ns) {
    const bool tiled = c->p.impl == PG_IMPL_TILED;
    std::vector<pg::SceneInfo>& S = c->scenes_h;
    S.assign(ns, pg::SceneInfo{});
    size_t goff = 0, cap = 0;
    uint64_t toff = 0, boff = 0;
    for (int s = 0; s < ns; ++s) {
        pg::SceneInfo& e = S[s];
        e.g = geom_of(d[s]);
        e.goff = goff;
        e.ntx = (d[s].nx + 3) / 4;
        e.nty = (d[s].ny + 3) / 4;
        e.ntz = (d[s].nz + 3) / 4;
        const int bits = ceil_log2(std::max(e.ntx, std::max(e.nty, e.ntz)));
        e.toff = (uint32_t)toff;
        e.tslots = 1u << (3 * bits);
        e.gbx = (d[s].nx + 7) / 8;
        e.gby = (d[s].ny + 7) / 8;
        e.gbz = (d[s].nz + 7) / 8;
        e.boff = (uint32_t)boff;
        e.nblk = (uint32_t)e.gbx * e.gby * e.gbz;
        e.poff = (uint32_t)cap;
        e.np = 0;
        goff += (size_t)d[s].nx * d[s].ny * d[s].nz;
        toff += e.tslots;
        boff += e.nblk;
        cap += (size_t)d[s].max_particles;
    }
    if (toff > (1ull << 31) || boff > (1ull << 31) || cap > (1ull << 31))
        return fail(c, PG_ERR_INVALID, "scene table too large");
    c->cap = std::max<size_t>(cap, 1);
    c->G = goff;
    c->nscenes = ns;
    pg::TiledWork& w = c->tw;
    w.nscenes = ns;
    w.tile_slots = (int)toff;
    w.total_blocks = (int)boff;
    // persistent tile workgroups: 16 per CU (3 resident per CU); 4,096 ran 1-3%
    // faster than 2,048 in back-to-back runs on the 1M-particle scene and the
    // 64-scene batch, and a workgroup past the active list exits at once
    w.nblocks_launch = 4096;
    w.chunk = pg::CHUNK;
    w.block = 256;
    // active-list entries: one per chunk of at most `chunk` particles of a
    // non-empty tile.  A tile with c particles has ceil(c / chunk) =
    // 1 + floor((c - 1) / chunk) entries, so a binning has at most
    // min(tiles, particles) + particles / chunk; sized for the smallest chunk.
    uint64_t tiles = 0;
    for (int s = 0; s < ns; ++s) tiles += (uint64_t)S[s].ntx * S[s].nty * S[s].ntz;
    const uint64_t max_entries = std::min<uint64_t>(tiles, c->cap) + c->cap / pg::MIN_CHUNK + 1;
    if (max_entries >= (1ull << 31)) return fail(c, PG_ERR_INVALID, "too many tiles and particles");
    w.max_entries = (int)max_entries;
    w.prof = &c->prof;
    int rc;
    if ((rc = dalloc(c, &c->scenes_d, (size_t)ns)) || (rc = dalloc(c, &c->S, 6 * c->cap)) ||
        (rc = dalloc(c, &c->pid, c->cap)) || (rc = dalloc(c, &c->mp, goff)) || (rc = dalloc(c, &c->gv, goff)) ||
        (tiled && ((rc = dalloc(c, &c->S2, 6 * c->cap)) || (rc = dalloc(c, &c->pid2, c->cap)) ||
                   (rc = dalloc(c, &w.count, toff)) || (rc = dalloc(c, &w.start, toff)) ||
                   (rc = dalloc(c, &w.active, 4 * max_entries)) || (rc = dalloc(c, &w.nactive, 1)) ||
                   (rc = dalloc(c, &w.key, c->cap)) || (rc = dalloc(c, &w.rank, c->cap)) ||
                   (rc = dalloc(c, &w.perm, c->cap)) || (rc = dalloc(c, &w.flags, boff)) ||
                   (rc = dalloc(c, &w.bsum, 5 * ((toff + 1023) / 1024))) || (rc = dalloc(c, &w.tslot, 2 * toff)) ||
                   (rc = dalloc(c, &w.tscene, toff)) || (rc = dalloc(c, &w.bscene, boff)) ||
                   (rc = dalloc(c, &w.bflag, toff)) || (rc = dalloc(c, &w.blist, toff)) ||
                   (rc = dalloc(c, &w.nblist, 1)) || (rc = dalloc(c, &w.err, 1)))))
        return rc;
    if (tiled) {
        // P2G window buffers: 8^3 nodes (8 KiB) for every possible active-list entry
        if ((rc = dalloc(c, &w.wbuf, (size_t)max_entries * 512))) return rc;
    }
    if ((rc = dalloc(c, &c->check, 4))) return rc;
    w.scenes = c->scenes_d;
    hipError_t e = hipSuccess;
    auto zero = [&](void* p, size_t bytes) {
        if (e == hipSuccess) e = hipMemset(p, 0, bytes);
    };
    zero(c->mp, goff * sizeof(float4));
    zero(c->gv, goff * sizeof(float4));
    if (tiled) {
        zero(w.count, toff * 4);
        zero(w.flags, boff * 4);
        zero(w.nactive, 4);
        zero(w.bflag, toff * 4);
        zero(w.nblist, 4);
        zero(w.err, 4);
        zero(w.tslot, toff * 8);
        // scene of every tile slot and grid block (the kernels look scenes up by these)
        std::vector<uint32_t> ts(toff), bs(boff);
        for (int s = 0; s < ns; ++s) {
            std::fill(ts.begin() + S[s].toff, ts.begin() + S[s].toff + S[s].tslots, (uint32_t)s);
            std::fill(bs.begin() + S[s].boff, bs.begin() + S[s].boff + S[s].nblk, (uint32_t)s);
        }
        if (e == hipSuccess) e = hipMemcpy(w.tscene, ts.data(), toff * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(w.bscene, bs.data(), boff * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return hip_fail(c, e, "context initialisation");
    e = hipMemcpy(c->scenes_d, S.data(), ns * sizeof(pg::SceneInfo), hipMemcpyHostToDevice);
    return e == hipSuccess ? PG_OK : hip_fail(c, e, "scene table upload");
}

// per-scene particle counts: poff/np of the table, uploaded on the stream
int set_scene_counts(pg_mpm_ctx* c, const int* counts, hipStream_t st) {
    uint32_t off = 0;
    for (int s = 0; s < c->nscenes; ++s) {
        c->scenes_h[s].poff = off;
        c->scenes_h[s].np = (uint32_t)counts[s];
        off += (uint32_t)counts[s];
    }
    hipError_t e = hipMemcpyAsync(c->scenes_d, c->scenes_h.data(), c->nscenes * sizeof(pg::SceneInfo),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? PG_OK : hip_fail(c, e, "scene table upload");
}

void drop_graphs(pg_mpm_ctx* c) {
    if (c->graphs.empty()) return;
    hipDeviceSynchronize();
    for (auto& ge : c->graphs) hipGraphExecDestroy(ge.exec);
    c->graphs.clear();
}

// Completes the work a tiled step left pending for the next one (its P2G
// windows, global-accumulator adds, block flags and tile counts: the grid
// update and the binning) before the particles, the binning buffers or the
// grid accumulators are used by anything but the next step.  The grid
// velocities then belong to the step after the last one taken.
hipError_t settle(pg_mpm_ctx* c, hipStream_t st) {
    if (!c->primed) return hipSuccess;
    c->primed = false;
    return pg::tiled_settle(c->tw, c->n, c->mp, c->gv, st);
}

// Runs nsteps tiled steps as one hipGraph launch on the caller's stream.  The
// step sequence of a given length starting from a given state buffer (and
// from pending work or not) is captured once and then replayed; the host-side
// ping-pong of the state buffers is applied after each launch exactly as the
// captured sequence did it.  A NULL (legacy) stream cannot be captured, so
// the capture then records on an internal stream (capture runs nothing); the
// graph is launched on the caller's stream either way.
hipError_t graph_steps(pg_mpm_ctx* c, int nsteps, hipStream_t st) {
    hipError_t e = hipSuccess;
    hipStream_t cap_st = st;
    if (!cap_st) {
        if (!c->gstream && (e = hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking)) != hipSuccess) return e;
        cap_st = c->gstream;
    }
    hipGraphExec_t exec = nullptr;
    for (auto& ge : c->graphs)
        if (ge.nsteps == nsteps && ge.S == c->S && ge.primed == c->primed) exec = ge.exec;
    if (!exec) {
        const bool prof = c->prof.on;
        c->prof.on = false;  // no event records inside the captured sequence
        float4 *S = c->S, *S2 = c->S2;
        uint32_t *pid = c->pid, *pid2 = c->pid2;
        float* E = c->have_E ? c->E : nullptr;
        float* E2 = c->have_E ? c->E2 : nullptr;
        float* m = c->have_mass ? c->mass : nullptr;
        float* m2 = c->have_mass ? c->mass2 : nullptr;
        hipGraph_t graph = nullptr;
        e = hipStreamBeginCapture(cap_st, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            hipError_t ek = pg::tiled_steps(c->tw, c->n, c->cap, S, S2, pid, pid2, E, E2, m, m2, c->mp, c->gv,
                                            nsteps, cap_st, c->primed, true);
            e = hipStreamEndCapture(cap_st, &graph);
            if (ek != hipSuccess) e = ek;
        }
        c->prof.on = prof;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) hipGraphDestroy(graph);
        if (e != hipSuccess) return e;
        c->graphs.push_back({nsteps, c->S, c->primed, exec});
    }
    if ((e = hipGraphLaunch(exec, st)) != hipSuccess) return e;
    c->primed = true;
    if (nsteps & 1) {
        std::swap(c->S, c->S2);
        std::swap(c->pid, c->pid2);
        if (c->have_E) std::swap(c->E, c->E2);
        if (c->have_mass) std::swap(c->mass, c->mass2);
    }
    return e;
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
        if (p->impl == PG_IMPL_TILED && std::max(p->nx, std::max(p->ny, p->nz)) > pg::MAX_TILED_NODES)
            return fail(nullptr, PG_ERR_INVALID, "the tiled path takes at most " +
                                                     std::to_string(pg::MAX_TILED_NODES) + " nodes per axis");
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
        hipError_t e = hipMemset(c->rgm, 0, G * 4);
        if (e == hipSuccess) e = hipMemset(c->rgp, 0, 3 * G * 4);
        if (e == hipSuccess) e = hipMemset(c->rgv, 0, 3 * G * 4);
        if (e == hipSuccess && (rc = dalloc(c, &c->check, 4))) {
            std::string m = c->err;
            pg_mpm_destroy(c);
            return fail(nullptr, rc, m);
        }
        if (e != hipSuccess) {
            pg_mpm_destroy(c);
            return hip_fail(nullptr, e, "grid initialisation");
        }
    } else {
        pg_scene_desc d = scene_from_params(*p);
        if ((rc = init_mls(c, &d, 1))) {
            std::string m = c->err;
            pg_mpm_destroy(c);
            return fail(nullptr, rc, m);
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
    for (auto& ge : c->graphs) hipGraphExecDestroy(ge.exec);
    if (c->gstream) hipStreamDestroy(c->gstream);
    if (c->bstream) hipStreamDestroy(c->bstream);
    for (hipEvent_t ev : c->bev)
        if (ev) hipEventDestroy(ev);
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

int pg_mpm_set_time(pg_mpm_ctx* c, double t, long long s) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (!std::isfinite(t) || s < 0) return fail(c, PG_ERR_INVALID, "time must be finite and steps >= 0");
    c->time = t;
    c->steps = s;
    return PG_OK;
}

}  // extern "C"

namespace {

// counts: per-scene particle counts of an MLS-MPM context (NULL: one scene of n)
int set_particles_impl(pg_mpm_ctx* c, int n, const int* counts, const float* x, const float* v, const float* C,
                       const float* F, const float* mass, const float* E, int memory, void* stream) {
    if (n < 0 || (size_t)n > c->cap)
        return fail(c, PG_ERR_INVALID, "particle count " + std::to_string(n) + " exceeds max_particles " +
                                           std::to_string(c->cap));
    if (n > 0 && (!x || !v)) return fail(c, PG_ERR_INVALID, "x and v are required");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (hipError_t es = settle(c, st); es != hipSuccess) return hip_fail(c, es, "settle");
    c->n = n;
    c->time = 0.0;
    c->steps = 0;
    c->version++;
    // material ids are per particle: a new particle set starts from the scenes' materials
    c->nmat = 0;
    c->tw.mt = c->mat_table();
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
    if ((rc = set_scene_counts(c, counts ? counts : &n, st))) return rc;
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
    drop_graphs(c);
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

}  // namespace

extern "C" {

int pg_mpm_set_particles(pg_mpm_ctx* c, int n, const float* x, const float* v, const float* C, const float* F,
                         const float* mass, const float* E, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->nscenes > 1) return fail(c, PG_ERR_INVALID, "a batch context takes pg_mpm_batch_set_particles");
    return set_particles_impl(c, n, nullptr, x, v, C, F, mass, E, memory, stream);
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

}  // extern "C"

namespace {
int check_state_impl(pg_mpm_ctx* c, int* first_nonfinite, int* first_outside, hipStream_t st);
}  // namespace

extern "C" {

int pg_mpm_get_grid(pg_mpm_ctx* c, float* mass, float* momentum, float* velocity, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    const size_t G = c->p.mode == PG_MODE_REF_COMPAT ? (size_t)c->p.nx * c->p.ny * c->p.nz : c->G;
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
                                       c->have_E ? c->E : nullptr, c->mat_table(), c->bc_flags(), c->mp, c->gv, st,
                                       &c->prof);
        } else {
            float* E = c->have_E ? c->E : nullptr;
            float* E2 = c->have_E ? c->E2 : nullptr;
            float* m = c->have_mass ? c->mass : nullptr;
            float* m2 = c->have_mass ? c->mass2 : nullptr;
            if (c->use_graph) {
                e = graph_steps(c, nsteps, st);
            } else {
                e = pg::tiled_steps(c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv,
                                    nsteps, st, c->primed, true);
                c->primed = true;
                if (c->have_E) { c->E = E; c->E2 = E2; }
                if (c->have_mass) { c->mass = m; c->mass2 = m2; }
            }
        }
    }
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_step");
    c->time += (double)c->p.dt * nsteps;
    c->steps += nsteps;
    c->version++;
    if (c->validate) return check_state_impl(c, nullptr, nullptr, st);
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

namespace {

// Runs the validation kernel and reads the binning error word; fills the
// message and returns PG_ERR_STATE when the state holds non-finite values or
// a binning index was out of bounds.  Synchronises the stream.
int check_state_impl(pg_mpm_ctx* c, int* first_nonfinite, int* first_outside, hipStream_t st) {
    if (first_nonfinite) *first_nonfinite = -1;
    if (first_outside) *first_outside = -1;
    // reads the state and the error word only: work a tiled step left pending
    // for the next one stays pending
    hipError_t e;
    if (c->p.mode == PG_MODE_REF_COMPAT) {
        const float lo[3] = {c->p.ox, c->p.oy, c->p.oz};
        const float hi[3] = {c->p.ox + (c->p.nx - 1) * c->p.dx, c->p.oy + (c->p.ny - 1) * c->p.dy,
                             c->p.oz + (c->p.nz - 1) * c->p.dz};
        e = pg::check_state_ref(c->n, c->rx, c->rv, lo, hi, c->check, st);
    } else {
        e = pg::check_state_mls(c->n, c->cap, c->S, c->p.impl == PG_IMPL_TILED ? c->pid : nullptr, c->scenes_d,
                                c->nscenes, c->check, st);
    }
    uint32_t out[4] = {0, 0, 0, 0}, berr = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(out, c->check, sizeof(out), hipMemcpyDeviceToHost, st);
    const bool tiled = c->p.mode == PG_MODE_MLS_MPM && c->p.impl == PG_IMPL_TILED;
    if (e == hipSuccess && tiled) e = hipMemcpyAsync(&berr, c->tw.err, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && tiled && berr) e = hipMemsetAsync(c->tw.err, 0, 4, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_check_state");
    if (first_nonfinite && out[1]) *first_nonfinite = (int)out[0];
    if (first_outside && out[3]) *first_outside = (int)out[2];
    std::string msg;
    if (out[1])
        msg = std::to_string(out[1]) + " particle(s) with a non-finite state; the first is particle " +
              std::to_string(out[0]);
    if (berr) {
        if (!msg.empty()) msg += "; ";
        msg += "binning index out of bounds (error bits " + std::to_string(berr) + ": 1 active list, "
               "2 sorted positions, 4 node blocks)";
    }
    if (!msg.empty()) return fail(c, PG_ERR_STATE, msg);
    return PG_OK;
}

}  // namespace

extern "C" {

int pg_mpm_check_state(pg_mpm_ctx* c, int* first_nonfinite, int* first_outside, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    return check_state_impl(c, first_nonfinite, first_outside, (hipStream_t)stream);
}

int pg_mpm_set_validation(pg_mpm_ctx* c, int on) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    c->validate = on != 0;
    return PG_OK;
}

int pg_mpm_set_materials(pg_mpm_ctx* c, const pg_material_desc* table, int nmat, const uint8_t* ids, int memory,
                         void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM) return fail(c, PG_ERR_INVALID, "materials need mode MLS_MPM");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    if (hipError_t es = settle(c, st); es != hipSuccess) return hip_fail(c, es, "settle");
    drop_graphs(c);
    if (!table || nmat == 0) {   // back to the scenes' materials
        c->nmat = 0;
        c->tw.mt = c->mat_table();
        return PG_OK;
    }
    if (nmat < 0 || nmat > pg::MAX_MATERIALS)
        return fail(c, PG_ERR_INVALID, "nmat must be in [1, " + std::to_string(pg::MAX_MATERIALS) + "]");
    if (c->n > 0 && !ids) return fail(c, PG_ERR_INVALID, "ids is required when the context has particles");
    std::vector<pg::MatRec> recs(pg::MAX_MATERIALS);
    for (int m = 0; m < nmat; ++m) {
        const pg_material_desc& d = table[m];
        pg::MatRec& r = recs[m];
        std::memset(&r, 0, sizeof(r));
        if (d.model != PG_MAT_FIXED_COROTATED && d.model != PG_MAT_NEO_HOOKEAN && d.model != PG_MAT_FLUID)
            return fail(c, PG_ERR_INVALID, "material " + std::to_string(m) + ": bad model");
        if (!(d.density >= 0.f) || !std::isfinite(d.density))
            return fail(c, PG_ERR_INVALID, "material " + std::to_string(m) + ": density must be >= 0");
        r.model = d.model;
        r.inv_rho = d.density > 0.f ? 1.f / d.density : 0.f;
        if (d.model == PG_MAT_FLUID) {
            if (!(d.viscosity >= 0.f) || !std::isfinite(d.viscosity))
                return fail(c, PG_ERR_INVALID, "material " + std::to_string(m) + ": viscosity must be >= 0");
            r.visc = d.viscosity;
        } else {
            if (!(d.poisson_ratio > -1.f && d.poisson_ratio < 0.5f))
                return fail(c, PG_ERR_INVALID, "material " + std::to_string(m) + ": poisson_ratio must be in (-1, 0.5)");
            if (!(d.youngs_modulus >= 0.f) || !std::isfinite(d.youngs_modulus))
                return fail(c, PG_ERR_INVALID, "material " + std::to_string(m) + ": youngs_modulus must be >= 0");
            const float nu = d.poisson_ratio;
            r.E = d.youngs_modulus;
            r.emu = 1.f / (2.f * (1.f + nu));
            r.elam = nu / ((1.f + nu) * (1.f - 2.f * nu));
        }
    }
    int rc;
    if (!c->mats_d && ((rc = dalloc(c, &c->mats_d, (size_t)pg::MAX_MATERIALS)) || (rc = dalloc(c, &c->mat_ids, c->cap))))
        return rc;
    hipError_t e = hipMemcpyAsync(c->mats_d, recs.data(), recs.size() * sizeof(pg::MatRec), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && c->n > 0) {
        // ids are checked on the host: copy device ids back first
        std::vector<uint8_t> h((size_t)c->n);
        if (memory == PG_HOST) {
            std::memcpy(h.data(), ids, h.size());
        } else {
            e = hipMemcpyAsync(h.data(), ids, h.size(), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
        }
        if (e == hipSuccess) {
            for (size_t i = 0; i < h.size(); ++i)
                if (h[i] >= nmat) return fail(c, PG_ERR_INVALID, "material id " + std::to_string(h[i]) +
                                                                     " of particle " + std::to_string(i) + " >= nmat");
            e = hipMemcpyAsync(c->mat_ids, h.data(), h.size(), hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
        }
    } else if (e == hipSuccess) {
        e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_set_materials");
    c->nmat = nmat;
    c->tw.mt = c->mat_table();
    c->version++;
    return PG_OK;
}

int pg_mpm_set_boundary(pg_mpm_ctx* c, const uint32_t* flags, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM) return fail(c, PG_ERR_INVALID, "boundary flags need mode MLS_MPM");
    if (memory != PG_HOST && memory != PG_DEVICE) return fail(c, PG_ERR_INVALID, "bad memory kind");
    hipStream_t st = (hipStream_t)stream;
    // the pending grid update of a tiled step uses the flags in force when it was issued
    if (hipError_t es = settle(c, st); es != hipSuccess) return hip_fail(c, es, "settle");
    drop_graphs(c);
    if (!flags) {
        c->have_bc = false;
        c->tw.node_bc = c->bc_flags();
        return PG_OK;
    }
    int rc;
    if (!c->node_bc && (rc = dalloc(c, &c->node_bc, c->G))) return rc;
    hipError_t e = hipMemcpyAsync(c->node_bc, flags, c->G * 4, kind_to_dev(memory), st);
    if (e == hipSuccess && memory == PG_HOST) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_set_boundary");
    c->have_bc = true;
    c->tw.node_bc = c->bc_flags();
    c->version++;
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

// true for device (or managed) memory; pageable or pinned host memory is host
bool is_device_ptr(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // unregistered host memory: clear the error it leaves
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
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

// the context's second stream and its events (created on first use)
hipError_t side_stream(pg_mpm_ctx* c) {
    hipError_t e = hipSuccess;
    if (!c->bstream) e = hipStreamCreateWithFlags(&c->bstream, hipStreamNonBlocking);
    for (hipEvent_t& ev : c->bev)
        if (e == hipSuccess && !ev) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    return e;
}

int ensure_tape(pg_mpm_ctx* c, int K) {
    const bool plain = c->p.impl == PG_IMPL_PLAIN;   // plain tapes hold no E / mass copies
    if (c->tape_len >= K + 1 && c->abar && (plain || !c->have_E || c->tape[0].E) &&
        (plain || !c->have_mass || c->tape[0].mass))
        return PG_OK;
    hipDeviceSynchronize();
    free_all(c->tape_allocs);
    c->tape.assign(K + 1, pg::TapeStep{});
    c->tape2.assign(plain ? 0 : K + 1, pg::TapeStep{});
    c->tape_len = 0;
    c->tape_filled = 0;   // the new tape's grids start cleared
    c->tape_seg = -1;
    const size_t G = c->G;
    const pg::TiledWork& w = c->tw;
    const size_t nblk = (size_t)w.total_blocks;
    auto& A = c->tape_allocs;
    int rc;
    for (int ts = 0; ts < 2 * (K + 1); ++ts) {
        if (ts > K && plain) break;
        pg::TapeStep& t = ts <= K ? c->tape[ts] : c->tape2[ts - K - 1];
        if (plain) {   // user order throughout: states and dense grids only
            if ((rc = valloc(c, A, &t.S, 6 * c->cap, false)) || (rc = valloc(c, A, &t.mp, G, true)) ||
                (rc = valloc(c, A, &t.gv, G, true)))
                goto fail;
            continue;
        }
        if ((rc = valloc(c, A, &t.S, 6 * c->cap, false)) || (rc = valloc(c, A, &t.pid, c->cap, false)) ||
            (c->have_E && (rc = valloc(c, A, &t.E, c->cap, false))) ||
            (c->have_mass && (rc = valloc(c, A, &t.mass, c->cap, false))) ||
            (rc = valloc(c, A, &t.perm, c->cap, false)) || (rc = valloc(c, A, &t.src, c->cap, false)) ||
            (rc = valloc(c, A, &t.alist, 4 * (size_t)w.max_entries, false)) ||
            (rc = valloc(c, A, &t.tslot, 2 * (size_t)w.tile_slots, false)) ||
            (rc = valloc(c, A, &t.nactive, 1, true)) || (rc = valloc(c, A, &t.mp, G, true)) ||
            (rc = valloc(c, A, &t.gv, G, true)) || (rc = valloc(c, A, &t.flags, nblk, true)))
            goto fail;
    }
    if ((rc = valloc(c, A, &c->abar, 6 * c->cap, false)) || (rc = valloc(c, A, &c->part, 3 * c->cap, false)) ||
        (rc = valloc(c, A, &c->agv, G, true)) || (rc = valloc(c, A, &c->apam, G, true)) ||
        (rc = valloc(c, A, &c->aE, c->cap, false)) || (rc = valloc(c, A, &c->aEs, c->cap, false)) ||
        (rc = valloc(c, A, &c->aEp, c->cap, false)) || (rc = valloc(c, A, &c->adst, c->cap, false)) ||
        (!plain && ((rc = valloc(c, A, &c->part2, 3 * c->cap, false)) ||
                    (rc = valloc(c, A, &c->aEp2, c->cap, false)) || (rc = valloc(c, A, &c->adst2, c->cap, false)) ||
                    (rc = valloc(c, A, &c->ainv, c->cap, false)) ||
                    (rc = valloc(c, A, &c->rw.count, (size_t)w.tile_slots, true)) ||
                    (rc = valloc(c, A, &c->rw.start, (size_t)w.tile_slots, false)) ||
                    (rc = valloc(c, A, &c->rw.key, c->cap, false)) || (rc = valloc(c, A, &c->rw.rank, c->cap, false)) ||
                    (rc = valloc(c, A, &c->rw.bsum, 5 * (((size_t)w.tile_slots + 1023) / 1024), false)) ||
                    (rc = valloc(c, A, &c->rw.bflag, (size_t)w.tile_slots, true)) ||
                    (rc = valloc(c, A, &c->rw.blist, (size_t)w.tile_slots, false)) ||
                    (rc = valloc(c, A, &c->rw.nblist, 1, true)) ||
                    (rc = valloc(c, A, &c->rw.wbuf, (size_t)w.max_entries * 512, false)))) ||
        (rc = valloc(c, A, &c->ux, 3 * c->cap, false)) ||
        (rc = valloc(c, A, &c->uv, 3 * c->cap, false)) || (rc = valloc(c, A, &c->uC, 9 * c->cap, false)) ||
        (rc = valloc(c, A, &c->uF, 9 * c->cap, false)))
        goto fail;
    c->tape_len = K + 1;
    return PG_OK;
fail:
    c->tape.clear();
    c->tape2.clear();
    c->rw = pg::TiledWork{};
    free_all(c->tape_allocs);
    c->abar = c->part = c->agv = c->apam = nullptr;
    c->aE = c->aEs = c->aEp = c->ux = c->uv = c->uC = c->uF = nullptr;
    c->adst = c->adst2 = nullptr;
    c->part2 = nullptr;
    c->aEp2 = nullptr;
    c->ainv = nullptr;
    return rc;
}

}  // namespace

extern "C" {

int pg_mpm_forward_checkpointed(pg_mpm_ctx* c, int nsteps, int checkpoint_every, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM) return fail(c, PG_ERR_INVALID, "checkpointed rollouts need mode MLS_MPM");
    if (nsteps <= 0 || checkpoint_every <= 0) return fail(c, PG_ERR_INVALID, "nsteps and checkpoint_every must be > 0");
    const int K = checkpoint_every;
    const int ncp = (nsteps + K - 1) / K;
    int rc;
    if ((rc = ensure_checkpoints(c, ncp)) || (rc = ensure_tape(c, K))) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = settle(c, st);
    const bool plain = c->p.impl == PG_IMPL_PLAIN;
    const pg::MlsGeom g = mls_geom(c);
    // a tape filled by an earlier forward and never reversed: clear its grids
    if (!plain && c->tape_filled > 0) {
        for (int s = 0; s < c->tape_filled && e == hipSuccess; ++s) e = pg::tape_clear(c->tw, c->tape[s], st);
        c->tape_filled = 0;
    }
    c->tape_seg = -1;
    for (int k = 0; k < ncp && e == hipSuccess; ++k) {
        pg::TapeStep& ck = c->ckpt[k];
        const int L = std::min(K, nsteps - k * K);
        if (plain) {   // E and mass stay in user order: only the state moves
            e = copy_planes(ck.S, c->S, c->cap, c->n, 6, st);
            for (int s = 0; s < L && e == hipSuccess; ++s)
                e = pg::mls_plain_step(g, c->n, c->cap, c->S, c->have_mass ? c->mass : nullptr,
                                       c->have_E ? c->E : nullptr, c->mat_table(), c->bc_flags(), c->mp, c->gv, st,
                                       &c->prof);
            continue;
        }
        // tiled: a checkpoint before a segment of two or more steps that is not the
        // last is copied on the context's second stream, beside the segment's
        // first step (which does not write the state it reads)
        const bool side = k + 1 < ncp && L >= 2;
        hipStream_t cs = st;
        if (side) {
            if (e == hipSuccess) e = side_stream(c);
            if (e == hipSuccess) e = hipEventRecord(c->bev[5], st);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->bstream, c->bev[5], 0);
            cs = c->bstream;
        }
        if (e == hipSuccess) e = copy_planes(ck.S, c->S, c->cap, c->n, 6, cs);
        if (e == hipSuccess) e = copy_u32(ck.pid, c->pid, c->n, cs);
        if (e == hipSuccess && c->have_E) e = hipMemcpyAsync(ck.E, c->E, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, cs);
        if (e == hipSuccess && c->have_mass)
            e = hipMemcpyAsync(ck.mass, c->mass, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, cs);
        if (e == hipSuccess && side) e = hipEventRecord(c->bev[6], cs);
        float* E = c->have_E ? c->E : nullptr;
        float* E2 = c->have_E ? c->E2 : nullptr;
        float* m = c->have_mass ? c->mass : nullptr;
        float* m2 = c->have_mass ? c->mass2 : nullptr;
        if (k + 1 == ncp) {
            // last segment: run it into the tape, so that the backward reverses it
            // without recomputing (the tape holds K steps anyway)
            pg::TapeStep* T = c->tape.data();
            if (e == hipSuccess) e = copy_planes(T[0].S, c->S, c->cap, c->n, 6, st);
            if (e == hipSuccess) e = copy_u32(T[0].pid, c->pid, c->n, st);
            if (e == hipSuccess && c->have_E)
                e = hipMemcpyAsync(T[0].E, c->E, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess && c->have_mass)
                e = hipMemcpyAsync(T[0].mass, c->mass, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess) e = pg::tiled_tape_forward(c->tw, c->n, c->cap, T, L, st);
            c->tape_filled = L;
            // the end state and the last grid become the context's
            if (e == hipSuccess) e = copy_planes(c->S, T[L].S, c->cap, c->n, 6, st);
            if (e == hipSuccess) e = copy_u32(c->pid, T[L].pid, c->n, st);
            if (e == hipSuccess && c->have_E)
                e = hipMemcpyAsync(c->E, T[L].E, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess && c->have_mass)
                e = hipMemcpyAsync(c->mass, T[L].mass, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(c->gv, T[L - 1].gv, c->G * sizeof(float4), hipMemcpyDeviceToDevice, st);
            c->tape_seg = k;
            continue;
        }
        if (e == hipSuccess && L >= 2) {
            // the first step reads the state and writes the other buffer; the
            // second overwrites it, so it waits for the checkpoint copy, which runs
            // on the context's second stream beside the first step
            e = pg::tiled_steps(c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv, 1,
                                st, false, true);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, c->bev[6], 0);
            if (e == hipSuccess)
                e = pg::tiled_steps(c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv,
                                    L - 1, st, true, false);
        } else if (e == hipSuccess) {
            e = pg::tiled_steps(c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv, L, st);
        }
        if (c->have_E) { c->E = E; c->E2 = E2; }
        if (c->have_mass) { c->mass = m; c->mass2 = m2; }
    }
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_forward_checkpointed");
    c->time += (double)c->p.dt * nsteps;
    c->steps += nsteps;
    c->version++;
    c->ck_version = c->version;
    c->tape_version = c->version;
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
    hipError_t e = hipSuccess;
    // seeds and outputs may be host or device arrays (copies use hipMemcpyDefault);
    // with any host array the call returns after the outputs are written
    bool any_host = false;
    for (const void* q : {(const void*)ax, (const void*)av, (const void*)aC, (const void*)aF, (const void*)gx,
                          (const void*)gv, (const void*)gC, (const void*)gF, (const void*)gE})
        if (q && !is_device_ptr(q)) any_host = true;
    auto seed = [&](float* dst, const float* src, size_t count) {
        if (e != hipSuccess) return;
        if (src) e = hipMemcpyAsync(dst, src, count * 4, hipMemcpyDefault, st);
        else e = pg::fill_u32(reinterpret_cast<uint32_t*>(dst), count, 0u, st);
    };
    seed(c->ux, ax, 3 * (size_t)n);
    seed(c->uv, av, 3 * (size_t)n);
    seed(c->uC, aC, 9 * (size_t)n);
    seed(c->uF, aF, 9 * (size_t)n);
    seed(c->aE, nullptr, (size_t)n);
    const bool plain = c->p.impl == PG_IMPL_PLAIN;
    const pg::MlsGeom g = mls_geom(c);
    float* aE = gE ? c->aE : nullptr;      // dL/dE only when asked for
    float* aEs = gE ? c->aEs : nullptr;
    const float* pE = c->have_E ? c->E : nullptr;
    const float* pm = c->have_mass ? c->mass : nullptr;
    // tiled: between segments the adjoint stays in storage order (that of
    // checkpoint k, where segment k's reverse ends); segment k-1's first reverse
    // kernel reads it through the inverse of that order
    if (plain) {
        for (int k = ncp - 1; k >= 0 && e == hipSuccess; --k) {
            const int L = std::min(K, c->ck_steps - k * K);
            pg::TapeStep* T = c->tape.data();
            const pg::TapeStep& ck = c->ckpt[k];
            e = copy_planes(T[0].S, ck.S, c->cap, n, 6, st);
            // recompute the segment: tape[s].S is the state before step s, tape[s].mp/gv its grid
            for (int s = 0; s < L && e == hipSuccess; ++s) {
                e = copy_planes(T[s + 1].S, T[s].S, c->cap, n, 6, st);
                if (e == hipSuccess)
                    e = pg::mls_plain_step(g, n, c->cap, T[s + 1].S, pm, pE, c->mat_table(), c->bc_flags(), T[s].mp,
                                           T[s].gv, st, &c->prof);
            }
            if (e == hipSuccess)
                e = pg::adjoint_from_user(n, c->cap, nullptr, c->ux, c->uv, c->uC, c->uF, c->abar, nullptr, nullptr,
                                          st);
            if (e == hipSuccess)
                e = pg::adjoint_segment_plain(g, n, c->cap, T, L, c->abar, c->part, c->agv, c->apam, aE, pE, pm,
                                              c->mat_table(), c->bc_flags(), st, &c->prof);
            if (e == hipSuccess)
                e = pg::adjoint_to_user(n, c->cap, nullptr, c->abar, c->ux, c->uv, c->uC, c->uF, nullptr, nullptr, st);
        }
    } else {
        // Tiled: segment k is reversed on the caller's stream from one tape while
        // segment k-1 is recomputed into the other on the context's second
        // stream, with its own binning scratch and window buffers; events hand
        // each tape over (recomputed -> reversed -> recomputed again).  Between
        // segments the adjoint stays in storage order (that of checkpoint k, where
        // segment k's reverse ends); segment k-1's first reverse kernel reads it
        // through the inverse of that order.
        pg::TapeStep* Ts[2] = {c->tape.data(), c->tape2.data()};
        const pg::TapeStep own0[2] = {c->tape[0], c->tape2[0]};
        pg::TiledWork w2 = c->tw;
        w2.count = c->rw.count;
        w2.start = c->rw.start;
        w2.key = c->rw.key;
        w2.rank = c->rw.rank;
        w2.bsum = c->rw.bsum;
        w2.bflag = c->rw.bflag;
        w2.blist = c->rw.blist;
        w2.nblist = c->rw.nblist;
        w2.wbuf = c->rw.wbuf;
        if (e == hipSuccess) e = side_stream(c);
        hipStream_t st2 = c->bstream;
        hipEvent_t ev_start = c->bev[0], *ev_recomputed = c->bev + 1, *ev_reversed = c->bev + 3;
        // the recompute stream starts after everything the caller queued (the forward, the seeds)
        if (e == hipSuccess) e = hipEventRecord(ev_start, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(st2, ev_start, 0);
        auto seg_len = [&](int k) { return std::min(K, c->ck_steps - k * K); };
        auto recompute = [&](int k, int set) {
            // from the checkpoint in place: neither the recompute nor the reverse
            // writes step 0's state
            pg::TapeStep* T = Ts[set];
            const pg::TapeStep& ck = c->ckpt[k];
            T[0].S = ck.S;
            T[0].pid = ck.pid;
            T[0].E = ck.E;
            T[0].mass = ck.mass;
            if (e == hipSuccess) e = pg::tiled_tape_forward(w2, n, c->cap, T, seg_len(k), st2);
            if (e == hipSuccess) e = hipEventRecord(ev_recomputed[set], st2);
        };
        const int k0 = ncp - 1;
        const bool taped = k0 == c->tape_seg && c->tape_version == c->version && c->tape_filled == seg_len(k0);
        c->tape_seg = -1;
        c->tape_filled = 0;   // the reverse pass clears the grids it reads
        if (!taped) recompute(k0, 0);
        bool used[2] = {false, false};
        int set = 0;
        for (int k = k0; k >= 0 && e == hipSuccess; --k) {
            const int L = seg_len(k);
            pg::TapeStep* T = Ts[set];
            const bool carried = k < k0;
            if (carried || !taped) e = hipStreamWaitEvent(st, ev_recomputed[set], 0);
            if (e == hipSuccess && k >= 1) {
                // the other tape is free once the reverse that last read it is done
                if (used[set ^ 1]) e = hipStreamWaitEvent(st2, ev_reversed[set ^ 1], 0);
                recompute(k - 1, set ^ 1);
            }
            if (e == hipSuccess && !carried)
                e = pg::adjoint_from_user(n, c->cap, T[L].pid, c->ux, c->uv, c->uC, c->uF, c->abar, aE, aEs, st);
            if (e == hipSuccess && carried) e = pg::slot_of_caller(n, c->ckpt[k + 1].pid, c->ainv, st);
            if (e == hipSuccess) {
                float4* const part[2] = {c->part, c->part2};
                float* const aEp[2] = {c->aEp, c->aEp2};
                uint32_t* const dst[2] = {c->adst, c->adst2};
                e = pg::adjoint_segment(c->tw, n, c->cap, T, L, c->abar, carried ? c->ainv : nullptr, part, c->agv,
                                        c->apam, aEs, aEp, dst, st);
            }
            if (e == hipSuccess && k == 0)
                e = pg::adjoint_to_user(n, c->cap, T[0].pid, c->abar, c->ux, c->uv, c->uC, c->uF, aEs, aE, st);
            if (e == hipSuccess) e = hipEventRecord(ev_reversed[set], st);
            used[set] = true;
            set ^= 1;
        }
        // every recompute was waited for by the reverse of its segment; the tapes'
        // own step-0 buffers again (the forward writes into them)
        Ts[0][0] = own0[0];
        Ts[1][0] = own0[1];
    }
    auto out = [&](float* dst, const float* src, size_t count) {
        if (e == hipSuccess && dst) e = hipMemcpyAsync(dst, src, count * 4, hipMemcpyDefault, st);
    };
    out(gx, c->ux, 3 * (size_t)n);
    out(gv, c->uv, 3 * (size_t)n);
    out(gC, c->uC, 9 * (size_t)n);
    out(gF, c->uF, 9 * (size_t)n);
    out(gE, c->aE, (size_t)n);
    if (e == hipSuccess && any_host) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(c, e, "pg_mpm_backward");
    return PG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches, graphs, autotuning
namespace {

std::string tune_cache_path() {
    if (const char* p = std::getenv("PG_AUTOTUNE_CACHE")) return p;
    const char* home = std::getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/physgrad_amd_autotune.txt";
}

std::string tune_key(const pg_mpm_ctx* c) {
    size_t nb = 1;
    while (nb < (size_t)std::max(c->n, 1)) nb <<= 1;
    return "slots=" + std::to_string(c->tw.tile_slots) + ";blocks=" + std::to_string(c->tw.total_blocks) +
           ";n=" + std::to_string(nb) + ";scenes=" + std::to_string(c->nscenes);
}

// cached (launch size, work-item size, workgroup threads) for this workload
// shape; false if none (missing trailing values keep their defaults)
bool tune_lookup(const pg_mpm_ctx* c, int* wg, int* chunk, int* block) {
    FILE* f = std::fopen(tune_cache_path().c_str(), "r");
    if (!f) return false;
    const std::string key = tune_key(c);
    char line[512];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        char k[400];
        int v = 0, ch = pg::CHUNK, bt = 256;
        const int got = std::sscanf(line, "%399s %d %d %d", k, &v, &ch, &bt);
        if (got >= 2 && key == k && v >= 8 && v <= (1 << 20) && ch >= pg::MIN_CHUNK && ch <= pg::MAX_CHUNK &&
            (bt == 256 || bt == 512)) {
            *wg = v;
            *chunk = got >= 3 ? ch : pg::CHUNK;
            *block = got >= 4 ? bt : 256;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void tune_store(const pg_mpm_ctx* c, int wg, int chunk, int block) {
    FILE* f = std::fopen(tune_cache_path().c_str(), "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d\n", tune_key(c).c_str(), wg, chunk, block);
    std::fclose(f);
}

}  // namespace

extern "C" {

int pg_mpm_batch_create(const pg_scene_desc* scenes, int nscenes, pg_mpm_ctx** out) {
    if (!scenes || !out || nscenes < 1 || nscenes > 65536)
        return fail(nullptr, PG_ERR_INVALID, "pg_mpm_batch_create: bad arguments");
    *out = nullptr;
    for (int s = 0; s < nscenes; ++s)
        if (const char* why = check_scene(scenes[s]))
            return fail(nullptr, PG_ERR_INVALID, "scene " + std::to_string(s) + ": " + why);
    pg_mpm_ctx* c = new (std::nothrow) pg_mpm_ctx();
    if (!c) return fail(nullptr, PG_ERR_NOMEM, "host allocation failed");
    pg_mpm_default_params(&c->p);
    c->p.mode = PG_MODE_MLS_MPM;
    c->p.impl = PG_IMPL_TILED;
    c->p.dt = scenes[0].dt;
    c->p.density = scenes[0].density;
    c->p.particle_volume = scenes[0].particle_volume;
    hipGetDevice(&c->device);
    int rc = init_mls(c, scenes, nscenes);
    if (rc == PG_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(c, e, "pg_mpm_batch_create");
    }
    if (rc != PG_OK) {
        std::string m = c->err;
        pg_mpm_destroy(c);
        return fail(nullptr, rc, m);
    }
    *out = c;
    return PG_OK;
}

int pg_mpm_batch_set_particles(pg_mpm_ctx* c, const int* counts, const float* x, const float* v, const float* C,
                               const float* F, const float* mass, const float* E, int memory, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM || c->p.impl != PG_IMPL_TILED)
        return fail(c, PG_ERR_INVALID, "batched particles need a tiled MLS-MPM context");
    if (!counts) return fail(c, PG_ERR_INVALID, "counts is required");
    long long total = 0;
    for (int s = 0; s < c->nscenes; ++s) {
        if (counts[s] < 0) return fail(c, PG_ERR_INVALID, "negative particle count");
        total += counts[s];
    }
    if ((size_t)total > c->cap) return fail(c, PG_ERR_INVALID, "particle total exceeds the batch capacity");
    return set_particles_impl(c, (int)total, counts, x, v, C, F, mass, E, memory, stream);
}

int pg_mpm_num_scenes(const pg_mpm_ctx* c) { return c ? c->nscenes : -1; }

int pg_mpm_set_graph(pg_mpm_ctx* c, int enable) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (enable && (c->p.mode != PG_MODE_MLS_MPM || c->p.impl != PG_IMPL_TILED))
        return fail(c, PG_ERR_INVALID, "graph replay is implemented for the tiled MLS-MPM path");
    if (!enable) drop_graphs(c);
    c->use_graph = enable != 0;
    return PG_OK;
}

int pg_mpm_get_launch(const pg_mpm_ctx* c, int* tile_workgroups) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (tile_workgroups) *tile_workgroups = c->tw.nblocks_launch;
    return PG_OK;
}

int pg_mpm_set_launch(pg_mpm_ctx* c, int tile_workgroups) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (tile_workgroups < 8 || tile_workgroups > (1 << 20))
        return fail(c, PG_ERR_INVALID, "tile_workgroups out of range");
    drop_graphs(c);
    c->tw.nblocks_launch = tile_workgroups;
    return PG_OK;
}

int pg_mpm_set_work_item(pg_mpm_ctx* c, int particles) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM || c->p.impl != PG_IMPL_TILED)
        return fail(c, PG_ERR_INVALID, "the work-item size applies to the tiled MLS-MPM path");
    if (particles < pg::MIN_CHUNK || particles > pg::MAX_CHUNK)
        return fail(c, PG_ERR_INVALID, "particles per work item must be in [" + std::to_string(pg::MIN_CHUNK) + ", " +
                                           std::to_string(pg::MAX_CHUNK) + "]");
    if (particles == c->tw.chunk) return PG_OK;
    // both scan passes of a binning run in the same call, so the new size takes
    // effect at the next binning; captured graphs hold the old one
    drop_graphs(c);
    c->tw.chunk = particles;
    c->version++;
    return PG_OK;
}

int pg_mpm_set_block(pg_mpm_ctx* c, int threads) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM || c->p.impl != PG_IMPL_TILED)
        return fail(c, PG_ERR_INVALID, "the tile workgroup size applies to the tiled MLS-MPM path");
    if (threads != 256 && threads != 512) return fail(c, PG_ERR_INVALID, "tile workgroup threads must be 256 or 512");
    if (threads == c->tw.block) return PG_OK;
    drop_graphs(c);
    c->tw.block = threads;
    c->version++;
    return PG_OK;
}

int pg_mpm_get_tuning(const pg_mpm_ctx* c, int* tile_workgroups, int* work_item, int* tile_edge,
                      int* block_threads) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (tile_workgroups) *tile_workgroups = c->tw.nblocks_launch;
    if (work_item) *work_item = c->tw.chunk;
    if (tile_edge) *tile_edge = pg::TT;
    if (block_threads) *block_threads = c->tw.block;
    return PG_OK;
}

int pg_mpm_autotune(pg_mpm_ctx* c, int trial_steps, int use_cache, void* stream) {
    if (!c) return fail(nullptr, PG_ERR_INVALID, "NULL context");
    if (c->p.mode != PG_MODE_MLS_MPM || c->p.impl != PG_IMPL_TILED)
        return fail(c, PG_ERR_INVALID, "the autotuner tunes the tiled MLS-MPM path");
    if (trial_steps < 1) return fail(c, PG_ERR_INVALID, "trial_steps must be >= 1");
    if (c->n == 0) return PG_OK;
    if (use_cache) {
        int wg = 0, ch = 0, bt = 0;
        if (tune_lookup(c, &wg, &ch, &bt)) {
            int rc = pg_mpm_set_launch(c, wg);
            if (rc == PG_OK) rc = pg_mpm_set_work_item(c, ch);
            return rc != PG_OK ? rc : pg_mpm_set_block(c, bt);
        }
    }
    hipStream_t st = (hipStream_t)stream;
    drop_graphs(c);
    // save the state (and the grid) so that the trials leave no trace
    const size_t n = (size_t)c->n;
    float4 *S0 = c->S, *S20 = c->S2;
    uint32_t *pid0 = c->pid, *pid20 = c->pid2;
    float *E0 = c->E, *E20 = c->E2, *m0 = c->mass, *m20 = c->mass2;
    float4 *sS = nullptr, *sG = nullptr;
    uint32_t* sP = nullptr;
    float *sE = nullptr, *sM = nullptr;
    auto release = [&]() {
        hipStreamSynchronize(st);
        hipFree(sS); hipFree(sG); hipFree(sP); hipFree(sE); hipFree(sM);
    };
    if (hipMalloc(&sS, 6 * n * sizeof(float4)) != hipSuccess || hipMalloc(&sG, c->G * sizeof(float4)) != hipSuccess ||
        hipMalloc(&sP, n * 4) != hipSuccess || (c->have_E && hipMalloc(&sE, n * 4) != hipSuccess) ||
        (c->have_mass && hipMalloc(&sM, n * 4) != hipSuccess)) {
        release();
        return fail(c, PG_ERR_NOMEM, "autotune scratch allocation failed");
    }
    auto save = [&](bool back) {
        for (int p = 0; p < 6; ++p)
            hipMemcpyAsync(back ? S0 + p * c->cap : sS + p * n, back ? sS + p * n : S0 + p * c->cap,
                           n * sizeof(float4), hipMemcpyDeviceToDevice, st);
        hipMemcpyAsync(back ? (void*)c->gv : (void*)sG, back ? (void*)sG : (void*)c->gv, c->G * sizeof(float4),
                       hipMemcpyDeviceToDevice, st);
        hipMemcpyAsync(back ? (void*)pid0 : (void*)sP, back ? (void*)sP : (void*)pid0, n * 4, hipMemcpyDeviceToDevice,
                       st);
        if (c->have_E)
            hipMemcpyAsync(back ? (void*)E0 : (void*)sE, back ? (void*)sE : (void*)E0, n * 4, hipMemcpyDeviceToDevice,
                           st);
        if (c->have_mass)
            hipMemcpyAsync(back ? (void*)m0 : (void*)sM, back ? (void*)sM : (void*)m0, n * 4, hipMemcpyDeviceToDevice,
                           st);
        c->S = S0; c->S2 = S20; c->pid = pid0; c->pid2 = pid20;
        c->E = E0; c->E2 = E20; c->mass = m0; c->mass2 = m20;
    };
    save(false);
    // after the grid is saved, so that the restored grid is the last step's
    if (hipError_t es = settle(c, st); es != hipSuccess) {
        release();
        return hip_fail(c, es, "settle");
    }
    hipEvent_t a = nullptr, b = nullptr;
    hipEventCreate(&a);
    hipEventCreate(&b);
    const int cand_wg[] = {256, 512, 1024, 2048, 4096, 8192};
    const int cand_chunk[] = {640, 1280, 2560};
    const int cand_block[] = {256, 512};
    int best = c->tw.nblocks_launch, best_chunk = c->tw.chunk, best_block = c->tw.block;
    float best_ms = 1e30f;
    hipError_t e = hipSuccess;
    for (int bt : cand_block)
    for (int ch : cand_chunk) {
        for (int v : cand_wg) {
            c->tw.nblocks_launch = v;
            c->tw.chunk = ch;
            c->tw.block = bt;
            // steady-state steps only: an untimed call of trial_steps steps that hands
            // its pending work over, then a timed call that takes it and hands over
            // in turn (the per-call start, binning from scratch, is not timed), then
            // the pending work completed before the state is restored
            float ms_sum = 0.f;
            if (e == hipSuccess) {
                save(true);
                float* E = c->have_E ? c->E : nullptr;
                float* E2 = c->have_E ? c->E2 : nullptr;
                float* m = c->have_mass ? c->mass : nullptr;
                float* m2 = c->have_mass ? c->mass2 : nullptr;
                e = pg::tiled_steps(c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp, c->gv,
                                    trial_steps, st, false, true);
                hipEventRecord(a, st);
                if (e == hipSuccess)
                    e = pg::tiled_steps(c->tw, c->n, c->cap, c->S, c->S2, c->pid, c->pid2, E, E2, m, m2, c->mp,
                                        c->gv, trial_steps, st, true, true);
                hipEventRecord(b, st);
                if (e == hipSuccess) e = pg::tiled_settle(c->tw, c->n, c->mp, c->gv, st);
                hipEventSynchronize(b);
                hipEventElapsedTime(&ms_sum, a, b);
            }
            if (e != hipSuccess) break;
            if (ms_sum < best_ms) {
                best_ms = ms_sum;
                best = v;
                best_chunk = ch;
                best_block = bt;
            }
        }
        if (e != hipSuccess) break;
    }
    c->tw.nblocks_launch = best;
    c->tw.chunk = best_chunk;
    c->tw.block = best_block;
    save(true);
    hipEventDestroy(a);
    hipEventDestroy(b);
    release();
    if (e != hipSuccess) return hip_fail(c, e, "autotune");
    tune_store(c, best, best_chunk, best_block);
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
