// Neighbour lists of the spatial hash (SpatialHashGrid::getNeighbors,
// src/sparse_data_structures.h:193-228) and the SPH-MPM coupling pairs and
// forces (SPHMPMCouplingSystem::findCouplingPairs / computeCouplingForces,
// src/sph_mpm_coupling.h:454-545), on the GPU.
//
// Both results are CSR: a count pass, an exclusive scan of the counts into
// row_ptr, the host reads the total (it sizes the output), a fill pass.
//
//   getNeighbors   count: one thread per query sums the counts of its
//                  (2R+1)^3 cells; fill: one wave per query copies each
//                  visited slot's particle range (coalesced).
//   coupling       both particle sets are binned with the spatial hash's
//                  radix sort on linear cell ids (cells of radius (1 + 2^-8),
//                  so distinct cells never share an id and the three cells of
//                  an x-row are one contiguous range); the MPM particles are
//                  gathered into cell order as float4 (x, y, z, index bits).
//                  One thread per SPH particle, SPH particles taken in cell
//                  order so that the threads of a wave read the same few
//                  candidate ranges: 9 ranges cover the 27 neighbour cells.
//                  The fill keeps rows of up to SMALL_ROW pairs sorted in
//                  registers (insertion by rank) with their penetrations;
//                  longer rows are written unsorted and sorted afterwards, by
//                  one wave each in LDS (bitonic, up to ROW_CAP pairs) or one
//                  workgroup each (LDS-sorted blocks, then merge passes in
//                  global memory through the penetration array as the second
//                  buffer), penetrations computed from the sorted indices.
//   forces         one thread per pair (coalesced writes); the two-way
//                  reaction is summed per SPH particle by one thread per row
//                  in pair order, so the float additions happen in the
//                  reference's order.
//
// The distance, force and friction arithmetic is written in the reference's
// operation order with contraction off, so results are bit-identical to it.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include "physgrad_mpm.h"
#include "pg_kernels.h"
#include "pg_tile.h"
#include "pg_scan.h"

namespace pg {
namespace {

constexpr int CPL_T = 256;                         // coupling workgroup (4 waves)
constexpr unsigned SMALL_ROW = 16;                 // pairs per row sorted in registers
constexpr unsigned ROW_CAP = 2048;                 // pairs per row sorted in a wave's LDS
constexpr unsigned MID_WG = 1024;                  // workgroups of the in-LDS row sort
constexpr unsigned LONG_BLOCK = 8192;              // block of a long row sorted in LDS
constexpr unsigned LONG_WG = 256;                  // workgroups of the long-row sort
constexpr uint64_t MAX_CPL_CELLS = 1ull << 28;

__device__ __forceinline__ uint32_t morton_slot(uint32_t x, uint32_t y, uint32_t z, uint32_t total) {
    return (expand3(x) | (expand3(y) << 1) | (expand3(z) << 2)) % total;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform_f(float v) {
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}

hipError_t read_total(const uint64_t* d, unsigned long long* h, hipStream_t st) {
    uint64_t v = 0;
    hipError_t e = hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    *h = v;
    return e;
}

// ------------------------------------------------------------------ getNeighbors

struct HashGeom {
    float lo[3];
    float cell;
    uint32_t res[3];
    uint32_t total;
    uint32_t R;  // search radius in cells
};

__global__ void k_nbr_count(const float* __restrict__ q, unsigned nq, HashGeom g, const uint32_t* __restrict__ ccount,
                            uint32_t* __restrict__ counts) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] = hash_coord(q[3 * (size_t)i + a], g.lo[a], g.cell);
    uint32_t cnt = 0;
    for (uint32_t dx = 0; dx <= 2 * g.R; ++dx) {
        const uint32_t gx = c[0] + dx - g.R;
        if (gx >= g.res[0]) continue;
        for (uint32_t dy = 0; dy <= 2 * g.R; ++dy) {
            const uint32_t gy = c[1] + dy - g.R;
            if (gy >= g.res[1]) continue;
            for (uint32_t dz = 0; dz <= 2 * g.R; ++dz) {
                const uint32_t gz = c[2] + dz - g.R;
                if (gz >= g.res[2]) continue;
                cnt += ccount[morton_slot(gx, gy, gz, g.total)];
            }
        }
    }
    counts[i] = cnt;
}

// one wave per query
__global__ void __launch_bounds__(256) k_nbr_fill(const float* __restrict__ q, unsigned nq, HashGeom g,
                                                  const uint32_t* __restrict__ pidx,
                                                  const uint32_t* __restrict__ cstart,
                                                  const uint32_t* __restrict__ ccount,
                                                  const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ out) {
    const unsigned w = uniform((blockIdx.x * 256u + threadIdx.x) >> 6);
    const unsigned lane = threadIdx.x & 63;
    if (w >= nq) return;
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] = uniform(hash_coord(q[3 * (size_t)w + a], g.lo[a], g.cell));
    uint32_t off = row_ptr[w];
    for (uint32_t dx = 0; dx <= 2 * g.R; ++dx) {
        const uint32_t gx = c[0] + dx - g.R;
        if (gx >= g.res[0]) continue;
        for (uint32_t dy = 0; dy <= 2 * g.R; ++dy) {
            const uint32_t gy = c[1] + dy - g.R;
            if (gy >= g.res[1]) continue;
            for (uint32_t dz = 0; dz <= 2 * g.R; ++dz) {
                const uint32_t gz = c[2] + dz - g.R;
                if (gz >= g.res[2]) continue;
                const uint32_t slot = morton_slot(gx, gy, gz, g.total);
                const uint32_t st = uniform(cstart[slot]), n = uniform(ccount[slot]);
                for (uint32_t j = lane; j < n; j += 64) out[off + j] = pidx[st + j];
                off += n;
            }
        }
    }
}

// ------------------------------------------------------------------ coupling pairs

struct CplGeom {
    float lo[3];
    float cell;
    uint32_t res[3];
    uint32_t total;
    float radius;
};

// |x_m - x_s| as the reference computes it: r = x_m - x_s, sqrt((rx rx + ry ry) + rz rz)
__device__ __forceinline__ float pair_distance(float mx, float my, float mz, float sx, float sy, float sz) {
#pragma clang fp contract(off)
    const float rx = mx - sx, ry = my - sy, rz = mz - sz;
    return sqrtf(rx * rx + ry * ry + rz * rz);
}

// MPM particles in cell order: (x, y, z, index bits)
__global__ void k_cpl_gather(const float* __restrict__ x, const uint32_t* __restrict__ pidx, unsigned nm,
                             float4* __restrict__ spos) {
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const uint32_t i = pidx[k];
    spos[k] = make_float4(x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], __uint_as_float(i));
}

// Visits the MPM particles within the radius of the point (sx, sy, sz):
// on_hit(mpm index, distance) for each, candidates taken from the 3 x 3 rows
// of 3 cells around the point's (clamped) cell.
template <class OnHit>
__device__ __forceinline__ void cpl_visit(const CplGeom& g, float sx, float sy, float sz,
                                          const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ ccount,
                                          const float4* __restrict__ spos, OnHit on_hit) {
    const float sp[3] = {sx, sy, sz};
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t t = hash_coord(sp[a], g.lo[a], g.cell);
        c[a] = (int)(t < g.res[a] - 1 ? t : g.res[a] - 1);
    }
    const int i0 = c[0] > 0 ? c[0] - 1 : 0;
    const int i1 = c[0] + 1 < (int)g.res[0] ? c[0] + 1 : c[0];
    for (int z = c[2] - 1; z <= c[2] + 1; ++z) {
        if (z < 0 || z >= (int)g.res[2]) continue;
        for (int y = c[1] - 1; y <= c[1] + 1; ++y) {
            if (y < 0 || y >= (int)g.res[1]) continue;
            const uint32_t row = ((uint32_t)z * g.res[1] + (uint32_t)y) * g.res[0];
            const uint32_t q0 = cstart[row + i0], q1 = cstart[row + i1] + ccount[row + i1];
            for (uint32_t q = q0; q < q1; ++q) {
                const float4 p = spos[q];
                const float d = pair_distance(p.x, p.y, p.z, sx, sy, sz);
                if (d < g.radius) on_hit(__float_as_uint(p.w), d);
            }
        }
    }
}

// one thread per SPH particle, in cell order (sidx)
__global__ void __launch_bounds__(CPL_T) k_cpl_count(const float* __restrict__ sx, const uint32_t* __restrict__ sidx,
                                                     unsigned ns, CplGeom g, const uint32_t* __restrict__ cstart,
                                                     const uint32_t* __restrict__ ccount,
                                                     const float4* __restrict__ spos, uint32_t* __restrict__ counts) {
    const unsigned t = blockIdx.x * CPL_T + threadIdx.x;
    if (t >= ns) return;
    const uint32_t s = sidx[t];
    uint32_t n = 0;
    cpl_visit(g, sx[3 * (size_t)s], sx[3 * (size_t)s + 1], sx[3 * (size_t)s + 2], cstart, ccount, spos,
              [&](uint32_t, float) { ++n; });
    counts[s] = n;
}

__global__ void __launch_bounds__(CPL_T) k_cpl_fill(const float* __restrict__ sx, const uint32_t* __restrict__ sidx,
                                                    unsigned ns, CplGeom g, const uint32_t* __restrict__ cstart,
                                                    const uint32_t* __restrict__ ccount,
                                                    const float4* __restrict__ spos,
                                                    const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ col,
                                                    uint32_t* __restrict__ penbits, uint32_t* __restrict__ prow,
                                                    uint32_t* __restrict__ long_rows, uint32_t* __restrict__ nlong) {
    const unsigned t = blockIdx.x * CPL_T + threadIdx.x;
    if (t >= ns) return;
    const uint32_t s = sidx[t];
    const uint32_t o = row_ptr[s], k = row_ptr[s + 1] - o;
    if (k == 0) return;
    const float x = sx[3 * (size_t)s], y = sx[3 * (size_t)s + 1], z = sx[3 * (size_t)s + 2];
    if (k <= SMALL_ROW) {
        // insertion by rank into a sorted register array (constant indices only)
        uint32_t ids[SMALL_ROW];
        float dist[SMALL_ROW];
#pragma unroll
        for (int i = 0; i < (int)SMALL_ROW; ++i) {
            ids[i] = 0xFFFFFFFFu;
            dist[i] = 0.f;
        }
        uint32_t n = 0;
        cpl_visit(g, x, y, z, cstart, ccount, spos, [&](uint32_t id, float d) {
            uint32_t pos = 0;
#pragma unroll
            for (int i = 0; i < (int)SMALL_ROW; ++i) pos += ids[i] < id ? 1u : 0u;
#pragma unroll
            for (int j = (int)SMALL_ROW - 1; j > 0; --j) {
                if ((uint32_t)j > pos) {
                    ids[j] = ids[j - 1];
                    dist[j] = dist[j - 1];
                } else if ((uint32_t)j == pos) {
                    ids[j] = id;
                    dist[j] = d;
                }
            }
            if (pos == 0) {
                ids[0] = id;
                dist[0] = d;
            }
            ++n;
        });
        // n == k: the count pass found the same hits
#pragma unroll
        for (int i = 0; i < (int)SMALL_ROW; ++i)
            if ((uint32_t)i < k) {
                col[o + i] = ids[i];
                penbits[o + i] = __float_as_uint(g.radius - dist[i]);
                prow[o + i] = s;
            }
        return;
    }
    uint32_t n = 0;
    cpl_visit(g, x, y, z, cstart, ccount, spos, [&](uint32_t id, float) {
        if (n < k) {
            col[o + n] = id;
            prow[o + n] = s;
        }
        ++n;
    });
    long_rows[atomicAdd(nlong, 1u)] = s;
}

// Rows of SMALL_ROW < k <= ROW_CAP pairs: one wave each sorts the row in LDS
// (bitonic) and writes the penetrations.
__global__ void __launch_bounds__(CPL_T) k_cpl_sort_mid(const float* __restrict__ sx, const float* __restrict__ mx,
                                                        float radius, const uint32_t* __restrict__ long_rows,
                                                        const uint32_t* __restrict__ nlong,
                                                        const uint32_t* __restrict__ row_ptr, uint32_t* col,
                                                        uint32_t* penbits) {
    __shared__ uint32_t buf[CPL_T / 64][ROW_CAP];
    uint32_t* lb = buf[threadIdx.x >> 6];
    const unsigned lane = threadIdx.x & 63;
    const uint32_t nl = *nlong;
    const unsigned nw = gridDim.x * (CPL_T / 64);
    for (uint32_t e = uniform(blockIdx.x * (CPL_T / 64) + (threadIdx.x >> 6)); e < nl; e += nw) {
        const uint32_t s = uniform(long_rows[e]);
        const uint32_t o = uniform(row_ptr[s]);
        const uint32_t k = uniform(row_ptr[s + 1]) - o;
        if (k > ROW_CAP) continue;
        uint32_t P = 1;
        while (P < k) P <<= 1;
        for (uint32_t i = lane; i < P; i += 64) lb[i] = i < k ? col[o + i] : 0xFFFFFFFFu;
        lds_order();
        for (uint32_t size = 2; size <= P; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t t = lane; t < (P >> 1); t += 64) {
                    const uint32_t i = 2 * t - (t & (stride - 1)), j = i + stride;
                    const uint32_t a = lb[i], b = lb[j];
                    if ((a > b) == ((i & size) == 0)) {
                        lb[i] = b;
                        lb[j] = a;
                    }
                }
                lds_order();
            }
        const float x = sx[3 * (size_t)s], y = sx[3 * (size_t)s + 1], z = sx[3 * (size_t)s + 2];
        for (uint32_t i = lane; i < k; i += 64) {
            const uint32_t id = lb[i];
            col[o + i] = id;
            const float d = pair_distance(mx[3 * (size_t)id], mx[3 * (size_t)id + 1], mx[3 * (size_t)id + 2], x, y, z);
            penbits[o + i] = __float_as_uint(radius - d);
        }
        lds_order();
    }
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Rows longer than ROW_CAP: one workgroup per row sorts col[row] ascending
// (LDS blocks, then merge passes between col and penbits), then writes the
// penetrations.
__global__ void __launch_bounds__(256) k_cpl_sort_long(const float* __restrict__ sx, const float* __restrict__ mx,
                                                       float radius, const uint32_t* __restrict__ long_rows,
                                                       const uint32_t* __restrict__ nlong,
                                                       const uint32_t* __restrict__ row_ptr, uint32_t* col,
                                                       uint32_t* penbits) {
    __shared__ uint32_t sb[LONG_BLOCK];
    const uint32_t nl = *nlong;
    const unsigned tid = threadIdx.x;
    for (uint32_t li = blockIdx.x; li < nl; li += gridDim.x) {
        const uint32_t s = long_rows[li];
        const uint32_t o = row_ptr[s], k = row_ptr[s + 1] - o;
        if (k <= ROW_CAP) continue;   // sorted by k_cpl_sort_mid
        uint32_t* a = col + o;
        uint32_t* b = penbits + o;
        for (uint32_t blk = 0; blk < k; blk += LONG_BLOCK) {
            const uint32_t m = k - blk < LONG_BLOCK ? k - blk : LONG_BLOCK;
            uint32_t P = 1;
            while (P < m) P <<= 1;
            for (uint32_t i = tid; i < P; i += 256) sb[i] = i < m ? a[blk + i] : 0xFFFFFFFFu;
            __syncthreads();
            for (uint32_t size = 2; size <= P; size <<= 1)
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t t = tid; t < (P >> 1); t += 256) {
                        const uint32_t i = 2 * t - (t & (stride - 1)), j = i + stride;
                        const uint32_t u = sb[i], v = sb[j];
                        if ((u > v) == ((i & size) == 0)) {
                            sb[i] = v;
                            sb[j] = u;
                        }
                    }
                    __syncthreads();
                }
            for (uint32_t i = tid; i < m; i += 256) a[blk + i] = sb[i];
            __syncthreads();
        }
        // merge sorted runs of width w pairwise; indices within a row are distinct
        uint32_t* src = a;
        uint32_t* dst = b;
        for (uint32_t w = LONG_BLOCK; w < k; w <<= 1) {
            for (uint32_t i = tid; i < k; i += 256) {
                const uint32_t run = i / w, p0 = (run & ~1u) * w;
                const uint32_t v = src[i];
                uint32_t pos;
                if (run & 1u) {
                    pos = i - w + lower_bound_u32(src + p0, w, v);
                } else {
                    const uint32_t b0 = p0 + w;
                    const uint32_t nb = b0 < k ? (k - b0 < w ? k - b0 : w) : 0u;
                    pos = i + (nb ? lower_bound_u32(src + b0, nb, v) : 0u);
                }
                dst[pos] = v;
            }
            __syncthreads();
            uint32_t* t = src;
            src = dst;
            dst = t;
        }
        const float x = sx[3 * (size_t)s], y = sx[3 * (size_t)s + 1], z = sx[3 * (size_t)s + 2];
        for (uint32_t i = tid; i < k; i += 256) {
            const uint32_t id = src[i];
            if (src != a) a[i] = id;
            const float d = pair_distance(mx[3 * (size_t)id], mx[3 * (size_t)id + 1], mx[3 * (size_t)id + 2], x, y, z);
            b[i] = __float_as_uint(radius - d);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ coupling forces

struct ForceCfg {
    float stiffness, damping, friction;
};

// One thread per pair (prow: its SPH particle): the reference's contact force.
// skip[k] = 1 for a pair closer than 1e-6 (zero forces, no reaction), as
// the reference skips it.
__global__ void k_cpl_forces(const float* __restrict__ sx, const float* __restrict__ sv,
                             const float* __restrict__ mx, const float* __restrict__ mv,
                             const uint32_t* __restrict__ prow, const uint32_t* __restrict__ col,
                             const float* __restrict__ pen, unsigned npairs, ForceCfg fc, float* __restrict__ solid,
                             float* __restrict__ fluid, uint8_t* __restrict__ skip) {
#pragma clang fp contract(off)
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const uint32_t s = prow[k], m = col[k];
    float fs[3] = {0.f, 0.f, 0.f}, ff[3] = {0.f, 0.f, 0.f};
    const float r[3] = {mx[3 * (size_t)m] - sx[3 * (size_t)s], mx[3 * (size_t)m + 1] - sx[3 * (size_t)s + 1],
                        mx[3 * (size_t)m + 2] - sx[3 * (size_t)s + 2]};
    const float d = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const bool sk = d < 1e-6f;
    if (!sk) {
        const float inv = 1.0f / d;
        const float nrm[3] = {r[0] * inv, r[1] * inv, r[2] * inv};
        float vd[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) vd[a] = mv[3 * (size_t)m + a] - sv[3 * (size_t)s + a];
        const float fn = fc.stiffness * pen[k];
        const float vn = vd[0] * nrm[0] + vd[1] * nrm[1] + vd[2] * nrm[2];
        const float fd = fc.damping * vn;
        const float sn = fn + fd;
        const float tn[3] = {nrm[0] * sn, nrm[1] * sn, nrm[2] * sn};
        float t[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) t[a] = vd[a] - nrm[a] * vn;
        const float ts = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        float fr[3] = {0.f, 0.f, 0.f};
        if (ts > 1e-6f) {
            const float it = 1.0f / ts;
            const float fm = fc.friction * fn;
#pragma unroll
            for (int a = 0; a < 3; ++a) fr[a] = (t[a] * it) * -fm;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            fs[a] = tn[a] + fr[a];
            ff[a] = (tn[a] + fr[a]) * -1.0f;
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        solid[3 * (size_t)k + a] = fs[a];
        fluid[3 * (size_t)k + a] = ff[a];
    }
    skip[k] = sk ? 1 : 0;
}

// Two-way reaction, one thread per SPH particle: its force plus the fluid
// forces of its pairs, added in pair order (the reference's order).
__global__ void k_cpl_accum(unsigned ns, const uint32_t* __restrict__ row_ptr, const float* __restrict__ fluid,
                            const uint8_t* __restrict__ skip, float* __restrict__ sph_force) {
#pragma clang fp contract(off)
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const uint32_t b = row_ptr[s], e = row_ptr[s + 1];
    if (b == e) return;
    float f0 = sph_force[3 * (size_t)s], f1 = sph_force[3 * (size_t)s + 1], f2 = sph_force[3 * (size_t)s + 2];
    for (uint32_t k = b; k < e; ++k) {
        if (skip[k]) continue;
        f0 = f0 + fluid[3 * (size_t)k];
        f1 = f1 + fluid[3 * (size_t)k + 1];
        f2 = f2 + fluid[3 * (size_t)k + 2];
    }
    sph_force[3 * (size_t)s] = f0;
    sph_force[3 * (size_t)s + 1] = f1;
    sph_force[3 * (size_t)s + 2] = f2;
}

// CSR sparse matrix-vector product, one thread per row, the row's products
// summed in column-index order (SparseMatrix::multiply,
// src/sparse_data_structures.h:392-410).  A column index >= ncols makes its
// row NaN instead of reading outside x.
__global__ void k_csr_spmv(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                           const float* __restrict__ val, const float* __restrict__ x, uint32_t ncols,
                           float* __restrict__ y, uint32_t nrows) {
#pragma clang fp contract(off)
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    float sum = 0.f;
    for (uint32_t k = row_ptr[r], e = row_ptr[r + 1]; k < e; ++k) {
        const uint32_t c = col[k];
        if (c >= ncols) {
            sum = __builtin_nanf("");
            break;
        }
        sum += val[k] * x[c];
    }
    y[r] = sum;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

unsigned nblocks(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace
}  // namespace pg

// ====================================================================== C-ABI

extern "C" size_t pg_hash_neighbors_workspace_size(unsigned nq) {
    return (((size_t)nq * 4 + 255) & ~(size_t)255) + pg::scan::ws_bytes(nq);
}

namespace {

bool hash_geom(float radius, float cell_size, const float* domain_min, const unsigned* res, pg::HashGeom* g) {
    if (!domain_min || !res || !(cell_size > 0.f) || !(radius >= 0.f) || !std::isfinite(radius)) {
        pg::set_thread_error("hash neighbours: domain_min, res, a cell size > 0 and a finite radius >= 0 are required");
        return false;
    }
    const uint64_t total = (uint64_t)res[0] * res[1] * res[2];
    if (total == 0 || total > 0xFFFFFFFFull) {
        pg::set_thread_error("hash neighbours: the cell count res0*res1*res2 must be in [1, 2^32)");
        return false;
    }
    const float rc = std::ceil(radius / cell_size);
    if (!(rc < 1024.f)) {
        pg::set_thread_error("hash neighbours: radius / cell_size must be below 1024");
        return false;
    }
    for (int a = 0; a < 3; ++a) {
        g->lo[a] = domain_min[a];
        g->res[a] = res[a];
    }
    g->cell = cell_size;
    g->total = (uint32_t)total;
    g->R = (uint32_t)rc;
    return true;
}

}  // namespace

extern "C" int pg_hash_neighbors_count(const float* queries, unsigned nq, float radius, float cell_size,
                                       const float* domain_min, const unsigned* res, const unsigned* cell_count,
                                       unsigned* row_ptr, unsigned long long* total, void* workspace,
                                       size_t workspace_bytes, void* stream) {
    pg::HashGeom g;
    if (!hash_geom(radius, cell_size, domain_min, res, &g)) return PG_ERR_INVALID;
    if (!row_ptr || !total || !cell_count || (nq > 0 && !queries)) {
        pg::set_thread_error("hash neighbours: NULL argument");
        return PG_ERR_INVALID;
    }
    if (workspace_bytes < pg_hash_neighbors_workspace_size(nq) || !workspace) {
        pg::set_thread_error("hash neighbours: workspace smaller than pg_hash_neighbors_workspace_size(nq)");
        return PG_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    uint32_t* counts = (uint32_t*)workspace;
    uint64_t* sws = (uint64_t*)((char*)workspace + (((size_t)nq * 4 + 255) & ~(size_t)255));
    if (nq) pg::k_nbr_count<<<pg::nblocks(nq, 256), 256, 0, st>>>(queries, nq, g, cell_count, counts);
    uint64_t* dtot = nullptr;
    hipError_t e = pg::scan::exclusive(counts, nq, row_ptr, true, sws, st, &dtot);
    if (e == hipSuccess) e = pg::read_total(dtot, total, st);
    if (e != hipSuccess) {
        pg::set_thread_error((std::string("hash neighbours: ") + hipGetErrorString(e)).c_str());
        return PG_ERR_HIP;
    }
    if (*total > 0xFFFFFFFFull) {
        pg::set_thread_error("hash neighbours: 2^32 or more neighbours in total");
        return PG_ERR_INVALID;
    }
    return PG_OK;
}

extern "C" int pg_hash_neighbors_fill(const float* queries, unsigned nq, float radius, float cell_size,
                                      const float* domain_min, const unsigned* res, const unsigned* particle_indices,
                                      const unsigned* cell_start, const unsigned* cell_count, const unsigned* row_ptr,
                                      unsigned* neighbors, void* stream) {
    pg::HashGeom g;
    if (!hash_geom(radius, cell_size, domain_min, res, &g)) return PG_ERR_INVALID;
    if (!row_ptr || !cell_count || !cell_start || (nq > 0 && (!queries || !neighbors || !particle_indices))) {
        pg::set_thread_error("hash neighbours: NULL argument");
        return PG_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    if (nq)
        pg::k_nbr_fill<<<pg::nblocks((size_t)nq * 64, 256), 256, 0, st>>>(queries, nq, g, particle_indices, cell_start,
                                                                          cell_count, row_ptr, neighbors);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        pg::set_thread_error((std::string("hash neighbours: ") + hipGetErrorString(e)).c_str());
        return PG_ERR_HIP;
    }
    return PG_OK;
}

extern "C" int pg_csr_spmv(const unsigned* row_ptr, const unsigned* col, const float* val, const float* x,
                           unsigned ncols, float* y, unsigned nrows, void* stream) {
    if (nrows > 0 && (!row_ptr || !y || (ncols > 0 && !x))) {
        pg::set_thread_error("csr spmv: NULL argument");
        return PG_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    if (nrows) pg::k_csr_spmv<<<pg::nblocks(nrows, 256), 256, 0, st>>>(row_ptr, col, val, x, ncols, y, nrows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        pg::set_thread_error((std::string("csr spmv: ") + hipGetErrorString(e)).c_str());
        return PG_ERR_HIP;
    }
    return PG_OK;
}

// ---------------------------------------------------------------------- coupling context

struct pg_coupling {
    pg_coupling_config cfg;
    pg::CplGeom g;
    std::string err;
    // binning: MPM cell ranges (linear cell ids), sorted MPM and SPH orders
    uint32_t *cstart = nullptr, *ccount = nullptr;
    uint32_t *cell_ids = nullptr, *pidx = nullptr, *sidx = nullptr;
    void* hws = nullptr;
    float4* spos = nullptr;
    unsigned cap_m = 0, cap_bin = 0;
    // rows
    uint32_t *counts = nullptr, *row_ptr = nullptr, *long_rows = nullptr, *nlong = nullptr;
    uint64_t* sws = nullptr;
    unsigned cap_s = 0;
    // pairs and scratch forces
    uint32_t *col = nullptr, *pen = nullptr, *prow = nullptr;
    uint8_t* skip = nullptr;
    float *solid = nullptr, *fluid = nullptr;
    size_t cap_p = 0;
    unsigned ns = 0, nm = 0;
    uint64_t npairs = 0;
    bool have_pairs = false;
};

namespace {

int cfail(pg_coupling* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    pg::set_thread_error(msg.c_str());
    return code;
}

template <typename T>
int grow(pg_coupling* c, T** p, size_t count) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return cfail(c, PG_ERR_NOMEM, "coupling: hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    *p = (T*)q;
    return PG_OK;
}

}  // namespace

extern "C" void pg_coupling_default_config(pg_coupling_config* cfg) {
    if (!cfg) return;
    cfg->coupling_stiffness = 1e5f;
    cfg->coupling_damping = 1e3f;
    cfg->friction_coefficient = 0.3f;
    cfg->coupling_radius = 0.15f;
    cfg->enable_two_way_coupling = 1;
    for (int a = 0; a < 3; ++a) {
        cfg->domain_min[a] = 0.f;
        cfg->domain_max[a] = 1.f;
    }
}

extern "C" int pg_coupling_create(const pg_coupling_config* cfg, pg_coupling** out) {
    if (!cfg || !out) return cfail(nullptr, PG_ERR_INVALID, "coupling: NULL argument");
    *out = nullptr;
    const float r = cfg->coupling_radius;
    if (!(r > 0.f) || !std::isfinite(r)) return cfail(nullptr, PG_ERR_INVALID, "coupling: coupling_radius must be > 0");
    // 2^-8 of margin over the radius keeps every pair within the 27 cells
    // despite the rounding of the cell coordinates
    const float cell = r * 1.00390625f;
    uint64_t total = 1;
    unsigned res[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = cfg->domain_max[a] - cfg->domain_min[a];
        if (!(ext > 0.f) || !std::isfinite(ext))
            return cfail(nullptr, PG_ERR_INVALID, "coupling: domain_max must exceed domain_min on every axis");
        const float rc = std::ceil(ext / cell);
        if (!(rc <= 32768.f))
            return cfail(nullptr, PG_ERR_INVALID, "coupling: more than 32768 binning cells on an axis");
        res[a] = (unsigned)rc;
        total *= res[a];
    }
    if (total > pg::MAX_CPL_CELLS)
        return cfail(nullptr, PG_ERR_INVALID,
                     "coupling: the box holds more than 2^28 cells of the coupling radius; enlarge the radius or "
                     "shrink the box");
    pg_coupling* c = new pg_coupling();
    c->cfg = *cfg;
    for (int a = 0; a < 3; ++a) {
        c->g.lo[a] = cfg->domain_min[a];
        c->g.res[a] = res[a];
    }
    c->g.cell = cell;
    c->g.total = (uint32_t)total;
    c->g.radius = r;
    int rc;
    if ((rc = grow(c, &c->cstart, total)) || (rc = grow(c, &c->ccount, total)) || (rc = grow(c, &c->nlong, 1))) {
        pg_coupling_destroy(c);
        return rc;
    }
    *out = c;
    return PG_OK;
}

extern "C" void pg_coupling_destroy(pg_coupling* c) {
    if (!c) return;
    void* ps[] = {c->cstart, c->ccount, c->cell_ids, c->pidx, c->sidx, c->hws, c->spos, c->counts, c->row_ptr,
                  c->long_rows, c->nlong, c->sws, c->col, c->pen, c->prow, c->skip, c->solid, c->fluid};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete c;
}

extern "C" const char* pg_coupling_last_error(const pg_coupling* c) { return c ? c->err.c_str() : ""; }

extern "C" int pg_coupling_find_pairs(pg_coupling* c, const float* sph_x, unsigned ns, const float* mpm_x,
                                      unsigned nm, unsigned long long* npairs, void* stream) {
    if (!c) return cfail(nullptr, PG_ERR_INVALID, "NULL coupling context");
    if (!npairs || (ns > 0 && !sph_x) || (nm > 0 && !mpm_x)) return cfail(c, PG_ERR_INVALID, "coupling: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    c->have_pairs = false;
    int rc;
    const unsigned nbin = std::max(nm, ns);
    if (nm > c->cap_m) {
        if ((rc = grow(c, &c->pidx, nm)) || (rc = grow(c, &c->spos, nm))) return rc;
        c->cap_m = nm;
    }
    if (nbin > c->cap_bin || !c->hws) {
        if ((rc = grow(c, &c->cell_ids, nbin)) ||
            (rc = grow(c, (unsigned char**)&c->hws, pg_spatial_hash_workspace_size(nbin))))
            return rc;
        c->cap_bin = nbin;
    }
    if (ns + 1 > c->cap_s || !c->row_ptr) {
        if ((rc = grow(c, &c->counts, ns + 1)) || (rc = grow(c, &c->row_ptr, ns + 1)) ||
            (rc = grow(c, &c->long_rows, ns + 1)) || (rc = grow(c, &c->sidx, ns + 1)) ||
            (rc = grow(c, (unsigned char**)&c->sws, pg::scan::ws_bytes(ns))))
            return rc;
        c->cap_s = ns + 1;
    }
    const unsigned res[3] = {c->g.res[0], c->g.res[1], c->g.res[2]};
    // MPM particles: cell ranges and cell order; SPH particles: cell order
    rc = pg::spatial_hash_build_ws(mpm_x, nm, c->g.cell, c->g.lo, res, c->g.total, c->cell_ids, c->pidx, c->cstart,
                                   c->ccount, c->hws, st, true);
    if (rc == PG_OK)
        rc = pg::spatial_hash_build_ws(sph_x, ns, c->g.cell, c->g.lo, res, c->g.total, c->cell_ids, c->sidx, nullptr,
                                       nullptr, c->hws, st, true);
    if (rc != PG_OK) return cfail(c, rc, std::string("coupling: ") + pg_mpm_last_error(nullptr));
    if (nm) pg::k_cpl_gather<<<pg::nblocks(nm, 256), 256, 0, st>>>(mpm_x, c->pidx, nm, c->spos);
    if (ns)
        pg::k_cpl_count<<<pg::nblocks(ns, pg::CPL_T), pg::CPL_T, 0, st>>>(sph_x, c->sidx, ns, c->g, c->cstart,
                                                                        c->ccount, c->spos, c->counts);
    uint64_t* dtot = nullptr;
    unsigned long long tot = 0;
    hipError_t e = pg::scan::exclusive(c->counts, ns, c->row_ptr, true, c->sws, st, &dtot);
    if (e == hipSuccess) e = pg::read_total(dtot, &tot, st);
    if (e != hipSuccess) return cfail(c, PG_ERR_HIP, std::string("coupling: ") + hipGetErrorString(e));
    if (tot > 0xFFFFFFFFull) return cfail(c, PG_ERR_INVALID, "coupling: 2^32 or more pairs");
    if (tot > c->cap_p) {
        const size_t cap = std::max<size_t>(tot, c->cap_p + c->cap_p / 2);
        if ((rc = grow(c, &c->col, cap)) || (rc = grow(c, &c->pen, cap)) || (rc = grow(c, &c->prow, cap)) ||
            (rc = grow(c, &c->skip, cap)) || (rc = grow(c, &c->solid, 3 * cap)) || (rc = grow(c, &c->fluid, 3 * cap)))
            return rc;
        c->cap_p = cap;
    }
    if (tot) {
        pg::k_set_u32<<<1, 64, 0, st>>>(c->nlong, 0u);
        pg::k_cpl_fill<<<pg::nblocks(ns, pg::CPL_T), pg::CPL_T, 0, st>>>(sph_x, c->sidx, ns, c->g, c->cstart,
                                                                       c->ccount, c->spos, c->row_ptr, c->col, c->pen,
                                                                       c->prow, c->long_rows, c->nlong);
        pg::k_cpl_sort_mid<<<pg::MID_WG, pg::CPL_T, 0, st>>>(sph_x, mpm_x, c->g.radius, c->long_rows, c->nlong,
                                                             c->row_ptr, c->col, c->pen);
        pg::k_cpl_sort_long<<<pg::LONG_WG, 256, 0, st>>>(sph_x, mpm_x, c->g.radius, c->long_rows, c->nlong,
                                                         c->row_ptr, c->col, c->pen);
        e = hipGetLastError();
        if (e != hipSuccess) return cfail(c, PG_ERR_HIP, std::string("coupling: ") + hipGetErrorString(e));
    }
    c->ns = ns;
    c->nm = nm;
    c->npairs = tot;
    c->have_pairs = true;
    *npairs = tot;
    return PG_OK;
}

extern "C" int pg_coupling_get_pairs(pg_coupling* c, unsigned* row_ptr, unsigned* col, float* penetration,
                                     void* stream) {
    if (!c) return cfail(nullptr, PG_ERR_INVALID, "NULL coupling context");
    if (!c->have_pairs) return cfail(c, PG_ERR_STATE, "coupling: get_pairs before find_pairs");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (row_ptr) e = hipMemcpyAsync(row_ptr, c->row_ptr, ((size_t)c->ns + 1) * 4, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && col && c->npairs)
        e = hipMemcpyAsync(col, c->col, (size_t)c->npairs * 4, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && penetration && c->npairs)
        e = hipMemcpyAsync(penetration, c->pen, (size_t)c->npairs * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return cfail(c, PG_ERR_HIP, std::string("coupling: ") + hipGetErrorString(e));
    return PG_OK;
}

extern "C" int pg_coupling_compute_forces(pg_coupling* c, const float* sph_x, const float* sph_v, const float* mpm_x,
                                          const float* mpm_v, float* sph_force, float* solid_force, float* fluid_force,
                                          void* stream) {
    if (!c) return cfail(nullptr, PG_ERR_INVALID, "NULL coupling context");
    if (!c->have_pairs) return cfail(c, PG_ERR_STATE, "coupling: compute_forces before find_pairs");
    if ((c->ns && (!sph_x || !sph_v)) || (c->nm && (!mpm_x || !mpm_v)) ||
        (c->cfg.enable_two_way_coupling && c->ns && !sph_force))
        return cfail(c, PG_ERR_INVALID, "coupling: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    float* fs = solid_force ? solid_force : c->solid;
    float* ff = fluid_force ? fluid_force : c->fluid;
    if (c->npairs) {
        const pg::ForceCfg fc{c->cfg.coupling_stiffness, c->cfg.coupling_damping, c->cfg.friction_coefficient};
        pg::k_cpl_forces<<<pg::nblocks(c->npairs, 256), 256, 0, st>>>(sph_x, sph_v, mpm_x, mpm_v, c->prow, c->col,
                                                                       (const float*)c->pen, (unsigned)c->npairs, fc,
                                                                       fs, ff, c->skip);
        if (c->cfg.enable_two_way_coupling)
            pg::k_cpl_accum<<<pg::nblocks(c->ns, 256), 256, 0, st>>>(c->ns, c->row_ptr, ff, c->skip, sph_force);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cfail(c, PG_ERR_HIP, std::string("coupling: ") + hipGetErrorString(e));
    return PG_OK;
}
