// Validation of the particle state: finds particles whose state is not
// finite and particles outside their grid box, by user index.
//
// The reference has no failure detection on this path (its solvers step on
// whatever the particles hold); the tiled path clamps the tile key of a
// particle outside the grid and skips its out-of-grid stencil nodes, so such
// a particle would otherwise be handled silently.
#include "pg_tile.h"

namespace pg {
namespace {

struct Box {
    float lo[3], hi[3];
};

// out[0] = lowest bad user index (non-finite), out[1] = count,
// out[2] = lowest user index outside the box, out[3] = count
__device__ __forceinline__ void report(uint32_t uid, bool bad, bool outside, uint32_t* out) {
    const uint64_t mb = __ballot(bad), mo = __ballot(outside);
    const int lane = threadIdx.x & 63;
    // one atomic per wave and kind: the wave's lowest index and its count
    if (mb) {
        uint32_t lo = bad ? uid : 0xFFFFFFFFu;
        for (int off = 32; off > 0; off >>= 1) lo = min(lo, (uint32_t)__shfl_xor(lo, off));
        if (lane == 0) {
            atomicMin(out + 0, lo);
            atomicAdd(out + 1, (uint32_t)__popcll(mb));
        }
    }
    if (mo) {
        uint32_t lo = outside ? uid : 0xFFFFFFFFu;
        for (int off = 32; off > 0; off >>= 1) lo = min(lo, (uint32_t)__shfl_xor(lo, off));
        if (lane == 0) {
            atomicMin(out + 2, lo);
            atomicAdd(out + 3, (uint32_t)__popcll(mo));
        }
    }
}

__global__ void __launch_bounds__(256) k_check_mls(int n, size_t cap, const float4* __restrict__ S,
                                                   const uint32_t* __restrict__ pid,
                                                   const SceneInfo* __restrict__ scenes, int nscenes,
                                                   uint32_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false, outside = false;
    uint32_t uid = 0xFFFFFFFFu;
    if (i < (size_t)n) {
        uid = pid ? pid[i] : (uint32_t)i;
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const float4 q = S[p * cap + i];
            bad = bad || !(__builtin_isfinite(q.x) && __builtin_isfinite(q.y) && __builtin_isfinite(q.z) &&
                           __builtin_isfinite(q.w));
        }
        const SceneInfo& sc = scenes[scene_of(scenes, nscenes, uid, [](const SceneInfo& t) { return t.poff; })];
        const float4 q0 = S[i];
        const float x[3] = {q0.x, q0.y, q0.z};
        const int nn[3] = {sc.g.nx, sc.g.ny, sc.g.nz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float lo = sc.g.o[a], hi = sc.g.o[a] + (float)(nn[a] - 1) * sc.g.dx;
            outside = outside || !(x[a] >= lo && x[a] <= hi);
        }
        outside = outside && !bad;
    }
    report(uid, bad, outside, out);
}

__global__ void __launch_bounds__(256) k_check_ref(int n, const float* __restrict__ x, const float* __restrict__ v,
                                                   Box box, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false, outside = false;
    if (i < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) bad = bad || !(__builtin_isfinite(x[3 * i + c]) && __builtin_isfinite(v[3 * i + c]));
#pragma unroll
        for (int c = 0; c < 3; ++c) outside = outside || !(x[3 * i + c] >= box.lo[c] && x[3 * i + c] <= box.hi[c]);
        outside = outside && !bad;
    }
    report((uint32_t)i, bad, outside, out);
}

__global__ void k_check_init(uint32_t* out) {
    if (threadIdx.x < 4) out[threadIdx.x] = (threadIdx.x & 1) ? 0u : 0xFFFFFFFFu;
}

}  // namespace

hipError_t check_state_mls(int n, size_t cap, const float4* S, const uint32_t* pid, const SceneInfo* scenes,
                           int nscenes, uint32_t* out, hipStream_t st) {
    k_check_init<<<1, 64, 0, st>>>(out);
    if (n > 0) k_check_mls<<<(n + 255) / 256, 256, 0, st>>>(n, cap, S, pid, scenes, nscenes, out);
    return hipGetLastError();
}

hipError_t check_state_ref(int n, const float* x, const float* v, const float* lo, const float* hi, uint32_t* out,
                           hipStream_t st) {
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = lo[a];
        b.hi[a] = hi[a];
    }
    k_check_init<<<1, 64, 0, st>>>(out);
    if (n > 0) k_check_ref<<<(n + 255) / 256, 256, 0, st>>>(n, x, v, b, out);
    return hipGetLastError();
}

}  // namespace pg
