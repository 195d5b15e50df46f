// Spatial hash: cell ids and a stable LSD radix sort of particle indices by
// cell id, bit-identical to SpatialHashGrid::buildHashTable
// (src/sparse_data_structures.h:121-190), whose host counting sort scatters
// in index order, i.e. is stable; any stable sort on the same keys yields the
// same permutation.
//
// Radix sort: 8-bit digits, 4096 keys per workgroup (256 threads x 16 rounds).
// Per pass: per-block digit histogram, one-workgroup exclusive scan of the
// digit-major histogram table, then a stable scatter in which each 256-key
// round ranks keys inside a wave64 with eight ballots (lanes with equal digit
// form a peer mask) and across the four waves with an LDS table.  Cell start
// and count come from a binary search of the sorted ids for every cell, which
// also defines start for empty cells as the reference's prefix sum does.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <string>
#include "physgrad_mpm.h"
#include "pg_kernels.h"
#include "pg_scan.h"

namespace pg {
namespace {

constexpr int RS_T = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_T * RS_ROUNDS;
constexpr int RADIX = 256;

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// getGridCoords + getCellID (src/sparse_data_structures.h:134-152); the
// float -> unsigned conversion saturates (below domain_min: 0), see hash_coord.
// linear: cell id (k r1 + j) r0 + i instead of the Morton code mod the cell
// count (no two cells share an id; used by the coupling's binning).
__global__ void k_cell_ids(const float* __restrict__ pos, unsigned n, float cell, float m0, float m1, float m2,
                           unsigned r0, unsigned r1, unsigned r2, int linear, uint32_t* __restrict__ ids,
                           uint32_t* __restrict__ idx) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned c0 = hash_coord(pos[3 * (size_t)i], m0, cell);
    unsigned c1 = hash_coord(pos[3 * (size_t)i + 1], m1, cell);
    unsigned c2 = hash_coord(pos[3 * (size_t)i + 2], m2, cell);
    c0 = c0 < r0 - 1 ? c0 : r0 - 1;
    c1 = c1 < r1 - 1 ? c1 : r1 - 1;
    c2 = c2 < r2 - 1 ? c2 : r2 - 1;
    if (linear) {
        ids[i] = (c2 * r1 + c1) * r0 + c0;
    } else {
        uint32_t code = expand_bits(c0) | (expand_bits(c1) << 1) | (expand_bits(c2) << 2);
        ids[i] = code % (r0 * r1 * r2);
    }
    idx[i] = i;
}

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint32_t* __restrict__ keys, unsigned n, int shift,
                                                  uint32_t* __restrict__ hist, unsigned nb) {
    __shared__ uint32_t h[RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        size_t i = base + (size_t)r * RS_T + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                     unsigned n, int shift, const uint32_t* __restrict__ hist,
                                                     unsigned nb) {
    __shared__ uint32_t base[RADIX];
    __shared__ uint32_t wc[RS_T / 64][RADIX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = hist[(size_t)tid * nb + blockIdx.x];
    for (int w = 0; w < RS_T / 64; ++w) wc[w][tid] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t b0 = (size_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        size_t i = b0 + (size_t)r * RS_T + tid;
        bool act = i < n;
        uint32_t k = act ? kin[i] : 0u;
        uint32_t v = act ? vin[i] : 0u;
        uint32_t d = (k >> shift) & (RADIX - 1);
        uint64_t peers = __ballot(act);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            uint64_t bm = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bm : ~bm;
        }
        uint32_t pre = __popcll(peers & lt);
        uint32_t cnt = __popcll(peers);
        if (act && pre == 0) wc[wave][d] = cnt;
        __syncthreads();
        if (act) {
            uint32_t off = base[d] + pre;
            for (int w = 0; w < wave; ++w) off += wc[w][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        {
            uint32_t tot = 0;
            for (int w = 0; w < RS_T / 64; ++w) { tot += wc[w][tid]; wc[w][tid] = 0; }
            base[tid] += tot;
        }
        __syncthreads();
    }
}

// cell_start[c] = lower_bound(ids, c); cell_count[c] = lower_bound(ids, c+1) - cell_start[c]
__global__ void k_cell_ranges(const uint32_t* __restrict__ ids, unsigned n, unsigned total,
                              uint32_t* __restrict__ start, uint32_t* __restrict__ count) {
    unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    auto lb = [&](uint32_t key) {
        unsigned lo = 0, hi = n;
        while (lo < hi) {
            unsigned mid = (lo + hi) >> 1;
            if (ids[mid] < key) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    unsigned a = lb(c), b = lb(c + 1);
    start[c] = a;
    count[c] = b - a;
}

inline size_t rs_blocks(unsigned n) { return (n + RS_TILE - 1) / RS_TILE; }

}  // namespace
}  // namespace pg

extern "C" size_t pg_spatial_hash_workspace_size(unsigned n) {
    size_t nb = pg::rs_blocks(n ? n : 1);
    size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    return 3 * a + (((size_t)pg::RADIX * nb * 4 + 255) & ~(size_t)255) +
           pg::scan::ws_bytes((unsigned)(pg::RADIX * nb));
}

namespace pg {
namespace {

// Arguments of a spatial-hash build; returns the cell count or 0 (and sets the
// thread's error message) when they are invalid.
unsigned check_hash_args(const float* pos, unsigned n, float cell_size, const float* domain_min,
                         const unsigned* res, const unsigned* cell_ids, const unsigned* particle_indices,
                         const unsigned* cell_start, const unsigned* cell_count) {
    if (!domain_min || !res || !(cell_size > 0.f)) {
        set_thread_error("spatial hash: domain_min, res and a cell size > 0 are required");
        return 0;
    }
    const uint64_t total64 = (uint64_t)res[0] * res[1] * res[2];
    if (total64 == 0 || total64 > 0xFFFFFFFFull) {
        set_thread_error("spatial hash: the cell count res0*res1*res2 must be in [1, 2^32)");
        return 0;
    }
    if ((n > 0 && (!pos || !cell_ids || !particle_indices)) || !cell_start || !cell_count) {
        set_thread_error("spatial hash: NULL output or position array");
        return 0;
    }
    return (unsigned)total64;
}

// The build over a device workspace of pg_spatial_hash_workspace_size(n)
// bytes at ws.  Every workspace word a launch reads was written by an earlier
// launch of the same build, so the workspace needs no clearing.
int hash_build(const float* pos, unsigned n, float cell_size, const float* domain_min, const unsigned* res,
               unsigned total, unsigned* cell_ids, unsigned* particle_indices, unsigned* cell_start,
               unsigned* cell_count, void* workspace, hipStream_t st, bool linear = false) {
    const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    uint32_t* kb[2] = {(uint32_t*)workspace, (uint32_t*)((char*)workspace + a)};
    uint32_t* v2 = (uint32_t*)((char*)workspace + 2 * a);
    uint32_t* hist = (uint32_t*)((char*)workspace + 3 * a);
    const size_t nbh = rs_blocks(n ? n : 1);
    uint64_t* sws = (uint64_t*)((char*)hist + ((((size_t)RADIX * nbh * 4) + 255) & ~(size_t)255));
    const uint32_t* sorted = cell_ids;
    if (n > 0) {
        k_cell_ids<<<(n + 255) / 256, 256, 0, st>>>(pos, n, cell_size, domain_min[0], domain_min[1], domain_min[2],
                                                   res[0], res[1], res[2], linear ? 1 : 0, cell_ids,
                                                   particle_indices);
        int bits = 0;
        while (bits < 32 && ((total - 1) >> bits)) ++bits;
        int passes = (bits + 7) / 8;
        if (passes < 2) passes = 2;
        if (passes & 1) ++passes;  // an even pass count ends in particle_indices
        const unsigned nb = (unsigned)rs_blocks(n);
        for (int p = 0; p < passes; ++p) {
            const uint32_t* ki = p == 0 ? cell_ids : kb[(p - 1) & 1];
            const uint32_t* vi = (p & 1) ? v2 : particle_indices;
            uint32_t* ko = kb[p & 1];
            uint32_t* vo = (p & 1) ? particle_indices : v2;
            const int shift = 8 * p;  // passes <= 4, digits above the key width are zero
            k_rs_hist<<<nb, RS_T, 0, st>>>(ki, n, shift, hist, nb);
            scan::exclusive(hist, RADIX * nb, hist, false, sws, st);
            k_rs_scatter<<<nb, RS_T, 0, st>>>(ki, vi, ko, vo, n, shift, hist, nb);
            sorted = ko;
        }
    }
    if (cell_start) k_cell_ranges<<<(total + 255) / 256, 256, 0, st>>>(sorted, n, total, cell_start, cell_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_thread_error((std::string("spatial hash: ") + hipGetErrorString(e)).c_str());
        return PG_ERR_HIP;
    }
    return PG_OK;
}

}  // namespace

int spatial_hash_build_ws(const float* pos, unsigned n, float cell_size, const float* domain_min, const unsigned* res,
                          unsigned total, unsigned* cell_ids, unsigned* particle_indices, unsigned* cell_start,
                          unsigned* cell_count, void* workspace, hipStream_t st, bool linear) {
    return hash_build(pos, n, cell_size, domain_min, res, total, cell_ids, particle_indices, cell_start, cell_count,
                      workspace, st, linear);
}

}  // namespace pg

extern "C" int pg_spatial_hash_build(const float* pos, unsigned n, float cell_size, const float* domain_min,
                                     const unsigned* res, unsigned* cell_ids, unsigned* particle_indices,
                                     unsigned* cell_start, unsigned* cell_count, void* workspace,
                                     size_t workspace_bytes, void* stream) {
    using namespace pg;
    const unsigned total =
        check_hash_args(pos, n, cell_size, domain_min, res, cell_ids, particle_indices, cell_start, cell_count);
    if (!total) return PG_ERR_INVALID;
    if (workspace_bytes < pg_spatial_hash_workspace_size(n) || (n > 0 && !workspace)) {
        set_thread_error("spatial hash: workspace smaller than pg_spatial_hash_workspace_size(n)");
        return PG_ERR_INVALID;
    }
    return hash_build(pos, n, cell_size, domain_min, res, total, cell_ids, particle_indices, cell_start, cell_count,
                      workspace, (hipStream_t)stream);
}

// ---------------------------------------------------------------- allocation-free build
//
// The workspace of pg_spatial_hash_build_noalloc is this file-scope device
// array, sized for PG_SPATIAL_HASH_NOALLOC_MAX particles; every device the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` only names the symbol: its device address on each device comes
// from hipGetSymbolAddress, looked up once per device on the first call with
// that device current, and only that address is handed to the kernels.
namespace pg {
namespace {

constexpr size_t NOALLOC_BLOCKS = (PG_SPATIAL_HASH_NOALLOC_MAX + RS_TILE - 1) / RS_TILE;
constexpr size_t NOALLOC_BYTES = 3 * (((size_t)PG_SPATIAL_HASH_NOALLOC_MAX * 4 + 255) & ~(size_t)255) +
                                 (((size_t)RADIX * NOALLOC_BLOCKS * 4 + 255) & ~(size_t)255) +
                                 (((((size_t)RADIX * NOALLOC_BLOCKS + scan::TILE - 1) / scan::TILE + 1) * 8 + 255) &
                                  ~(size_t)255);

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// device address of this device's scratch array, or NULL (message set)
void* scratch_on_current_device() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        set_thread_error("spatial hash (no-alloc): no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_thread_error((std::string("spatial hash (no-alloc): hipGetSymbolAddress failed: ") +
                              hipGetErrorString(e)).c_str());
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace pg

extern "C" int pg_spatial_hash_build_noalloc(const float* pos, unsigned n, float cell_size, const float* domain_min,
                                             const unsigned* res, unsigned* cell_ids, unsigned* particle_indices,
                                             unsigned* cell_start, unsigned* cell_count, void* stream) {
    using namespace pg;
    static_assert(NOALLOC_BYTES >= 256, "scratch holds at least the histogram of one block");
    if (n > PG_SPATIAL_HASH_NOALLOC_MAX) {
        set_thread_error(("spatial hash (no-alloc): " + std::to_string(n) + " particles exceed the maximum of " +
                          std::to_string(PG_SPATIAL_HASH_NOALLOC_MAX) + "; use pg_spatial_hash_build with a workspace")
                             .c_str());
        return PG_ERR_INVALID;
    }
    const unsigned total =
        check_hash_args(pos, n, cell_size, domain_min, res, cell_ids, particle_indices, cell_start, cell_count);
    if (!total) return PG_ERR_INVALID;
    void* ws = scratch_on_current_device();
    if (!ws) return PG_ERR_HIP;
    if (pg_spatial_hash_workspace_size(n) > NOALLOC_BYTES) {
        set_thread_error("spatial hash (no-alloc): scratch smaller than the workspace");
        return PG_ERR_INVALID;
    }
    return hash_build(pos, n, cell_size, domain_min, res, total, cell_ids, particle_indices, cell_start, cell_count,
                      ws, (hipStream_t)stream);
}
