// Host-side launch functions of the MPM kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pg {

// ref-compat geometry (reference MPMGrid: dims, cell size, origin; order 1..3)
struct RefGeom {
    int nx, ny, nz;
    float o[3];
    float h[3];
    int order;
};

// MLS-MPM scene description used by every MLS kernel.
struct MlsGeom {
    int nx, ny, nz;
    float dx, inv_dx;
    float o[3];
    float dt;
    float g[3];
    int bc;          // sticky wall thickness (nodes)
    int material;    // MAT_FIXED_COROTATED / MAT_NEO_HOOKEAN
    float nu;
    float E;         // uniform Young's modulus (used when no per-particle E)
    float p_vol;     // V0
    float p_mass;    // uniform mass (used when no per-particle mass)
};

// message of the last failure on this host thread (pg_mpm_last_error(NULL))
void set_thread_error(const char* msg);

// Cell coordinate of the spatial hash: the truncating float -> unsigned
// conversion of (x - lo) / cell (SpatialHashGrid::getGridCoords,
// src/sparse_data_structures.h:134-140), saturated where C++ leaves it
// undefined: values below 0 (points below domain_min, NaN) give 0, values of
// 2^32 and above give 2^32 - 1.
__host__ __device__ __forceinline__ uint32_t hash_coord(float x, float lo, float cell) {
    const float t = (x - lo) / cell;
    if (!(t > 0.f)) return 0u;
    return t < 4294967296.f ? (uint32_t)t : 0xFFFFFFFFu;
}

// Spatial hash build (cell ids, stable radix sort, cell ranges) over a device
// workspace of pg_spatial_hash_workspace_size(n) bytes; domain_min and res are
// host arrays.  linear: cell id (k res1 + j) res0 + i instead of the
// reference's Morton code mod the cell count.  cell_start NULL: no cell ranges
// (only the sorted order).  Used by the coupling build.
int spatial_hash_build_ws(const float* pos, unsigned n, float cell_size, const float* domain_min, const unsigned* res,
                          unsigned total, unsigned* cell_ids, unsigned* particle_indices, unsigned* cell_start,
                          unsigned* cell_count, void* workspace, hipStream_t st, bool linear);

// Per-particle material table of a context (device memory): recs[ids[u]] is
// the material of the particle with caller index u.  recs == NULL: every
// particle has its scene's material, Poisson ratio and volume.
struct MatRec;
struct MatTable {
    const MatRec* recs;
    const uint8_t* ids;
};
constexpr int MAX_MATERIALS = 256;

struct KernelProf;
hipError_t refcompat_step(const RefGeom& g, int n, float* x, float* v, const float* m, float* gm, float* gp, float* gv,
                          float dt, const float* grav, hipStream_t st, KernelProf* prof);

// ---- MLS-MPM plain port: one thread per particle, global float atomics, dense grid
// mt: material table (recs NULL: the scene's); node_bc: boundary flags per node (NULL: none)
hipError_t mls_plain_step(const MlsGeom& g, int n, size_t cap, float4* S, const float* mass, const float* E,
                          const MatTable& mt, const uint32_t* node_bc, float4* mp, float4* gv, hipStream_t st,
                          KernelProf* prof);

// ---- layout conversion between user AoS arrays and the six float4 planes
// pid (optional) maps storage slot -> user index; NULL means identity.
hipError_t pack_state(int n, size_t cap, float4* S, const float* x, const float* v, const float* C, const float* F,
                      hipStream_t st);
hipError_t unpack_state(int n, size_t cap, const float4* S, const uint32_t* pid, float* x, float* v, float* C, float* F,
                        hipStream_t st);
hipError_t unpermute_f32(int n, const float* src, const uint32_t* pid, float* dst, hipStream_t st);
hipError_t fill_f4(float4* p, size_t n, float4 v, hipStream_t st);
hipError_t fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st);
hipError_t iota_u32(uint32_t* p, size_t n, hipStream_t st);
hipError_t grid_split(size_t G, const float4* mp, const float4* gv, float* mass, float* mom, float* vel, hipStream_t st);

// ---- per-kernel timing with HIP events on the launch stream (opt-in)
enum ProfKernel {
    PK_TILE_P2G = 0,   // tiled P2G only (first step of a call)
    PK_TILE_G2P2G = 1, // tiled fused G2P(n) + P2G(n+1)
    PK_TILE_G2P = 2,   // tiled G2P only (last step of a call)
    PK_GRID = 3,       // grid update (tiled: flagged blocks; plain: dense)
    PK_BIN = 4,        // tile keys scan + scatter
    PK_PLAIN_P2G = 5,
    PK_PLAIN_G2P = 6,
    PK_CLEAR = 7,
    PK_REF_STEP = 8,
    PK_ADJ_G2P = 9,    // reverse G2P (tiled)
    PK_ADJ_GRID = 10,  // reverse grid update
    PK_ADJ_P2G = 11,   // reverse P2G (tiled)
    PK_ADJ_FUSED = 12, // reverse P2G of step n+1 fused with the reverse G2P of step n (tiled)
    PK_COUNT = 16
};

struct KernelProf {
    bool on = false;
    double total_ms[PK_COUNT] = {};
    long long count[PK_COUNT] = {};
    struct Rec { int k; hipEvent_t a, b; };
    Rec recs[4096];
    int nrec = 0;
    hipEvent_t pool[8192];
    int npool = 0;
    hipEvent_t get();
    void begin(int k, hipStream_t st, hipEvent_t* out);
    void end(int k, hipEvent_t a, hipStream_t st);
    hipError_t flush();  // synchronizes recorded events and accumulates
};

// RAII helper: times the launches issued in its scope when prof is on.
struct ProfScope {
    KernelProf* p;
    int k;
    hipStream_t st;
    hipEvent_t a = nullptr;
    ProfScope(KernelProf* p_, int k_, hipStream_t s) : p(p_), k(k_), st(s) {
        if (p && p->on) p->begin(k, st, &a);
    }
    ~ProfScope() {
        if (p && p->on && a) p->end(k, a, st);
    }
};

// ---- MLS-MPM tiled path (cell-binned particles, LDS-staged transfers)
//
// One context advances one or many independent scenes.  Scene s owns the user
// particle range [poff, poff+np), the grid nodes [goff, goff + nx*ny*nz) of
// the concatenated grid, the tile ids [toff, toff + tslots) and the 8^3
// blocks [boff, boff + gbx*gby*gbz); every kernel looks its scene up per tile
// or per block.
struct SceneInfo {
    MlsGeom g;
    unsigned long long goff;
    uint32_t toff, tslots;
    uint32_t boff, nblk;
    uint32_t poff, np;
    int ntx, nty, ntz;
    int gbx, gby, gbz;
    int pad;
};

// Work items of the tile kernels: the particles of a tile in chunks of at
// most `chunk` particles (TiledWork::chunk, default CHUNK, a runtime knob of
// the autotuner in [MIN_CHUNK, MAX_CHUNK]), so that a dense tile is spread over
// several workgroups instead of serialising the step on one.
constexpr int CHUNK = 1280;
constexpr int MIN_CHUNK = 640;
constexpr int MAX_CHUNK = 5120;

// Largest grid side of the tiled path (64 tiles of 4 cells per axis).  One
// step on a 264^3 grid (66 tiles per axis, 2^21 tile slots) raised an illegal
// memory access on MI355X with an earlier version of the window buffers whose
// cause was not found; until such grids have a passing test the tiled path
// refuses them at context creation.
constexpr int MAX_TILED_NODES = 256;

struct TiledWork {
    KernelProf* prof;      // optional per-kernel timing
    const SceneInfo* scenes;  // device scene table
    int nscenes;
    int tile_slots;        // total tile ids over all scenes
    int total_blocks;      // total 8^3 grid blocks over all scenes
    uint32_t* count;       // [tile_slots]
    uint32_t* start;       // [tile_slots]
    uint32_t* active;      // [4 max_entries] active list: (tile, start, count, scene) per chunk of a tile
    uint32_t* nactive;     // [1]
    uint32_t* key;         // [cap] tile id per storage slot
    uint32_t* rank;        // [cap] rank within tile
    uint32_t* perm;        // [cap] sorted position -> storage slot
    uint32_t* flags;       // [total_blocks] 8^3 block flags (unused by the forward step)
    uint32_t* bsum;        // [5 * ceil(tile_slots/1024)] scan block totals (particles, entries per size class)
    uint32_t* tslot;       // [tile_slots] uint2 (first active-list entry of the tile or ~0, entry count)
    uint32_t* tscene;      // [tile_slots] scene of each tile slot
    uint32_t* bscene;      // [total_blocks] scene of each 8^3 grid block
    uint32_t* bflag;       // [tile_slots] 4^3 node blocks to update: bit 0 covered by a buffered P2G
                           // window, bit 1 hit by out-of-window P2G (global atomics)
    uint32_t* blist;       // [tile_slots] the flagged blocks, in the order they were first flagged
    uint32_t* nblist;      // [1] their count; emptied after every grid update
    float4* wbuf;          // [max_entries][512] P2G window sums by active-list entry
    int max_entries;       // capacity of the active list (sized for MIN_CHUNK)
    int nblocks_launch;    // persistent grid size of the tile kernels
    int chunk;             // particles per work item (CHUNK by default)
    int block;             // threads per tile workgroup: 256 (default) or 512
    uint32_t* err;         // [1] binning error word: ERR_* bits set by guarded writes, never cleared by a step
    MatTable mt;           // per-particle materials (recs NULL: the scenes')
    const uint32_t* node_bc;  // [total grid nodes] boundary flags per node, or NULL
};

// Bits of the binning error word.  The binning kernels check every index they
// form against its buffer and, instead of writing out of bounds, set a bit;
// pg_mpm_check_state reports and clears them.
enum BinningError : uint32_t {
    ERR_ALIST = 1u,   // more active-list entries than max_entries
    ERR_PERM = 2u,    // a sorted position outside [0, n) (tile counts out of step with the keys)
    ERR_BLOCK = 4u,   // a node block id outside the tile slots, or a full block list
};

// Validation of the particle state: out[0] = lowest user index of a particle
// with a non-finite x, v, C or F (0xFFFFFFFF: none), out[1] = their number,
// out[2] = lowest user index of a particle outside its scene's grid box,
// out[3] = their number.  S: the six planes (pid: storage slot -> user index,
// NULL = identity); out must hold 4 words and is written on the stream.
hipError_t check_state_mls(int n, size_t cap, const float4* S, const uint32_t* pid, const SceneInfo* scenes,
                           int nscenes, uint32_t* out, hipStream_t st);
// the same for the ref-compat layout (x, v interleaved [3n]), one grid box
hipError_t check_state_ref(int n, const float* x, const float* v, const float* lo, const float* hi, uint32_t* out,
                           hipStream_t st);

// One step of a segment tape: the step's input state in its storage order, its
// binning, and its grid (valid on the flagged 8^3 blocks).
struct TapeStep {
    float4* S;          // 6*cap state planes
    uint32_t* pid;      // storage slot -> user index
    float* E;           // per-particle E in storage order (or NULL)
    float* mass;        // per-particle mass in storage order (or NULL)
    uint32_t* perm;     // sorted position -> storage slot
    uint32_t* src;      // output position of this step -> storage slot of this step
    uint32_t* alist;    // uint4 (tile, start, count, scene) per chunk of an active tile
    uint32_t* nactive;
    uint32_t* tslot;    // tile slot -> uint2 (first entry in alist or ~0, entry count)
    float4* mp;         // P2G result (m, p) of the step; cleared by the reverse pass
    float4* gv;         // (v, m) after the grid update
    uint32_t* flags;    // 8^3 blocks holding the step's grid; cleared by the reverse pass
};

// Recompute L steps from tape[0].S (tape[0].pid/E/mass set), filling tape[0..L-1]
// and tape[L].S/pid/E/mass (the end state in its storage order).  The steps'
// grids (mp) and 8^3 flags must be clear; the reverse pass leaves them clear,
// tape_clear clears a tape that was filled but not reversed.
hipError_t tape_clear(const TiledWork& w, const TapeStep& t, hipStream_t st);
hipError_t tiled_tape_forward(TiledWork& w, int n, size_t cap, TapeStep* tape, int L,
                              hipStream_t st);

// Reverse pass over a recomputed segment tape[0..L-1] (see mls_adjoint.hip).
// abar: the adjoint state by storage slot, in (after step L-1) and out
// (before step 0).  aEs: dL/dE by storage slot, in and out like abar (NULL:
// not wanted).  in_inv (NULL: abar and aEs come in tape[L]'s order): abar and
// aEs come in another storage order of the same particles, whose slot of
// caller index u is in_inv[u].  Two sets of per-step buffers, used
// alternately: part (3*cap) and aEp (cap) by output position, dst (cap) from a
// step's storage slots to its output positions.
hipError_t adjoint_segment(const TiledWork& w, int n, size_t cap, TapeStep* tape, int L,
                           float4* abar, const uint32_t* in_inv, float4* const part[2], float4* agv,
                           float4* apam, float* aEs, float* const aEp[2], uint32_t* const dst[2],
                           hipStream_t st);
// The plain port of the reverse pass over a segment recomputed with
// mls_plain_step: tape[s].S is the state before step s (user order),
// tape[s].mp / .gv the dense grid of step s.
hipError_t adjoint_segment_plain(const MlsGeom& g, int n, size_t cap, TapeStep* tape, int L, float4* abar,
                                 float4* part, float4* agv, float4* apam, float* aE_user, const float* E,
                                 const float* mass, const MatTable& mt, const uint32_t* node_bc, hipStream_t st,
                                 KernelProf* prof);
// user-order adjoint arrays (x,v [3n], C,F [9n]; NULL = 0) -> adjoint planes in storage order (pid;
// NULL = identity), and dL/dE aE (user order) -> aEs (storage order) when aEs is not NULL
hipError_t adjoint_from_user(int n, size_t cap, const uint32_t* pid, const float* ax, const float* av,
                             const float* aC, const float* aF, float4* abar, const float* aE, float* aEs,
                             hipStream_t st);
hipError_t adjoint_to_user(int n, size_t cap, const uint32_t* pid, const float4* abar, float* ax, float* av,
                           float* aC, float* aF, const float* aEs, float* aE, hipStream_t st);
// inv[pid[s]] = s: the slot of each caller index in a storage order
hipError_t slot_of_caller(int n, const uint32_t* pid, uint32_t* inv, hipStream_t st);

hipError_t tiled_init_keys(TiledWork& w, int n, const float4* S, const uint32_t* pid, hipStream_t st);
// Runs nsteps full steps.  S/S2, pid/pid2, E/E2, mass/mass2 ping-pong; on
// return the current buffers are S, pid, E, mass (the references are swapped).
// primed: the previous call left this state's P2G in the window buffers (and
// the global accumulator) and its tile counts counted, so the call starts with
// that grid update and binning instead of a binning from scratch and a P2G.
// leave_primed: the last step runs the fused kernel, leaving the new state's
// P2G and counts pending in the same way (else a G2P-only kernel, nothing
// pending).  Pending work must be consumed by the next primed call or by
// tiled_settle before the particles, the binning buffers or the grid are used
// otherwise.
hipError_t tiled_steps(TiledWork& w, int n, size_t cap, float4*& S, float4*& S2, uint32_t*& pid,
                       uint32_t*& pid2, float*& E, float*& E2, float*& mass, float*& mass2, float4* mp, float4* gv,
                       int nsteps, hipStream_t st, bool primed = false, bool leave_primed = false);
// consumes the work a primed call left pending: the grid update of its windows
// (into gv) and the binning of its counts
hipError_t tiled_settle(TiledWork& w, int n, float4* mp, float4* gv, hipStream_t st);

}  // namespace pg
