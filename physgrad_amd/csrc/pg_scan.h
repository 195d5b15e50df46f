// Exclusive prefix sum of uint32 counts on the device, in three launches:
// per-tile sums (4096 values per workgroup), one workgroup scanning the tile
// sums in 64 bits, per-tile scan with the tile's offset.  out may alias in.
// Used by the radix sort of the spatial hash and by the CSR builds.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {
namespace scan {
namespace {

constexpr int T = 256;
constexpr int PER = 16;
constexpr unsigned TILE = T * PER;

__device__ __forceinline__ uint32_t wave_incl_u32(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

__global__ void __launch_bounds__(T) k_tiles(const uint32_t* in, unsigned n, uint64_t* __restrict__ tsum) {
    __shared__ uint64_t ws[T / 64];
    const size_t base = (size_t)blockIdx.x * TILE;
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const size_t i = base + (size_t)r * T + threadIdx.x;
        if (i < n) s += in[i];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < T / 64; ++w) t += ws[w];
        tsum[blockIdx.x] = t;
    }
}

// exclusive scan of nt tile sums in place (one workgroup); tsum[nt] = total
__global__ void __launch_bounds__(1024) k_top(uint64_t* __restrict__ tsum, unsigned nt) {
    __shared__ uint64_t ws[16];
    __shared__ uint64_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (unsigned b = 0; b < nt; b += 1024) {
        const unsigned i = b + threadIdx.x;
        const uint64_t v = i < nt ? tsum[i] : 0;
        uint64_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t u = __shfl_up(x, o, 64);
            if (lane >= o) x += u;
        }
        if (lane == 63) ws[w] = x;
        __syncthreads();
        uint64_t before = carry_s;
        for (int k = 0; k < w; ++k) before += ws[k];
        if (i < nt) tsum[i] = before + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) tsum[nt] = carry_s;
}

// each thread scans 16 consecutive values (all read before any is written,
// so out may alias in); write_total: out[n] = the total
__global__ void __launch_bounds__(T) k_apply(const uint32_t* in, unsigned n, const uint64_t* __restrict__ tsum,
                                             unsigned nt, uint32_t* out, int write_total) {
    __shared__ uint32_t ws[T / 64];
    const size_t base = (size_t)blockIdx.x * TILE + (size_t)threadIdx.x * PER;
    uint32_t v[PER];
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        v[r] = base + r < n ? in[base + r] : 0u;
        s += v[r];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_u32(s);
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t run = (uint32_t)tsum[blockIdx.x] + incl - s;
    for (int k = 0; k < w; ++k) run += ws[k];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        if (base + r < n) out[base + r] = run;
        run += v[r];
    }
    if (write_total && blockIdx.x == 0 && threadIdx.x == 0) out[n] = (uint32_t)tsum[nt];
}

inline unsigned tiles(unsigned n) { return n ? (n + TILE - 1) / TILE : 1u; }

// device bytes of the scan's scratch for n values
inline size_t ws_bytes(unsigned n) { return (((size_t)tiles(n) + 1) * sizeof(uint64_t) + 255) & ~(size_t)255; }

// out[i] = in[0] + ... + in[i-1] for i < n (and out[n] = the total when
// write_total); *total (may be NULL) receives the device address of the
// 64-bit total inside ws.
inline hipError_t exclusive(const uint32_t* in, unsigned n, uint32_t* out, bool write_total, uint64_t* ws,
                            hipStream_t st, uint64_t** total = nullptr) {
    const unsigned nt = tiles(n);
    k_tiles<<<nt, T, 0, st>>>(in, n, ws);
    k_top<<<1, 1024, 0, st>>>(ws, nt);
    k_apply<<<nt, T, 0, st>>>(in, n, ws, nt, out, write_total ? 1 : 0);
    if (total) *total = ws + nt;
    return hipGetLastError();
}

}  // namespace
}  // namespace scan
}  // namespace pg
