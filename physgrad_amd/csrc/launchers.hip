// The reference's standalone transfer launchers with their own semantics
// (src/mpm_g2p2g_kernels.cu:72-271 kernels, :506-617 launchers):
//   P2G  PIC scatter of mass and momentum from float4 (x,y,z,mass) particles,
//        stencil base floor((x-o)/h), offsets {-1,0,1} (order 2) or
//        {-2,-1,0,1} (order 3), weight threshold 1e-10, one float atomic per
//        node and channel into caller-owned grid arrays (not cleared here);
//   G2P  normalised gather over nodes with mass > 1e-10, FLIP/PIC blend with
//        the old velocity, x += v dt;
//   fused clear -> P2G with gravity on the particle velocity -> v = p/m ->
//        G2P;
//   the launcher class G2P2GKernelLauncher<T>::launchG2P2G
//        (src/mpm_g2p2g_kernels.h:494-547 with the header kernels :211-453):
//        the fused sequence with an un-normalised G2P sum and the blend
//        applied only when use_affine_pic is set (pure PIC otherwise).  The reference needs a cooperative launch with grid.sync() for
//        this; here the phases are stream-ordered kernels on the caller's
//        stream, which gives the same ordering without co-residency limits.
// Launch shape: 256 threads, one particle per thread.
#include "pg_device.h"
#include "pg_kernels.h"
#include "physgrad_mpm.h"

namespace pg {
namespace {

struct LGeom {
    float o[3], h[3];
    int d[3];
    int order;
};

// evalQuadraticGPU / evalCubicGPU (src/mpm_g2p2g_kernels.cu:27-66), weights only
__device__ __forceinline__ float l_weight(int order, float x) {
    const float a = fabsf(x);
    if (order == 2) {
        if (a < 0.5f) return 0.75f - x * x;
        if (a < 1.5f) {
            const float t = 1.5f - a;
            return 0.5f * t * t;
        }
        return 0.f;
    }
    if (a < 1.f) {
        const float a2 = a * a, a3 = a2 * a;
        return 0.666666686534881591796875f - a2 + 0.5f * a3;
    }
    if (a < 2.f) {
        const float t = 2.f - a, t2 = t * t, t3 = t2 * t;
        return 0.16666667163372039794921875f * t3;
    }
    return 0.f;
}

__device__ __forceinline__ void l_base(const float* p, const LGeom& g, int* b, float* f) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = (int)floorf((p[a] - g.o[a]) / g.h[a]);
        f[a] = (p[a] - g.o[a]) / g.h[a] - (float)b[a];
    }
}

__global__ void k_l_clear(float* __restrict__ gm, float* __restrict__ gp, float* __restrict__ gv, size_t G) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x) {
        gm[i] = 0.f;
        gp[3 * i] = gp[3 * i + 1] = gp[3 * i + 2] = 0.f;
        gv[3 * i] = gv[3 * i + 1] = gv[3 * i + 2] = 0.f;
    }
}

// optimizedP2GKernel (:72-157); gravity is added to the particle velocity in
// the fused variant (:307-330)
__global__ void k_l_p2g(const float4* __restrict__ pos, const float4* __restrict__ vel, float* __restrict__ gm,
                        float* __restrict__ gp, LGeom g, int n, float gdt0, float gdt1, float gdt2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float4 pm = pos[q], vp = vel[q];
    const float p[3] = {pm.x, pm.y, pm.z};
    const float mass = pm.w;
    const float mom[3] = {(vp.x + gdt0) * mass, (vp.y + gdt1) * mass, (vp.z + gdt2) * mass};
    int b[3];
    float f[3];
    l_base(p, g, b, f);
    const int ks = g.order == 2 ? 3 : 4, off = ks / 2;
    for (int di = -off; di < ks - off; ++di)
        for (int dj = -off; dj < ks - off; ++dj)
            for (int dk = -off; dk < ks - off; ++dk) {
                const int i = b[0] + di, j = b[1] + dj, k = b[2] + dk;
                if (i < 0 || i >= g.d[0] || j < 0 || j >= g.d[1] || k < 0 || k >= g.d[2]) continue;
                const float w = l_weight(g.order, f[0] - di) * l_weight(g.order, f[1] - dj) * l_weight(g.order, f[2] - dk);
                if (w > 1e-10f) {
                    const size_t id = (size_t)i + (size_t)j * g.d[0] + (size_t)k * g.d[0] * g.d[1];
                    atomic_add_f32(gm + id, mass * w);
                    atomic_add_f32(gp + 3 * id + 0, mom[0] * w);
                    atomic_add_f32(gp + 3 * id + 1, mom[1] * w);
                    atomic_add_f32(gp + 3 * id + 2, mom[2] * w);
                }
            }
}

// grid velocity phase of the fused kernel (:387-398) / calculateGridVelocityKernel
__global__ void k_l_vel(const float* __restrict__ gm, const float* __restrict__ gp, float* __restrict__ gv, size_t G) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x) {
        const float m = gm[i];
        if (m > 1e-10f) {
            gv[3 * i] = gp[3 * i] / m;
            gv[3 * i + 1] = gp[3 * i + 1] / m;
            gv[3 * i + 2] = gp[3 * i + 2] / m;
        }
    }
}

// optimizedG2PKernel (:163-271, normalise = true) and the header launcher's
// fusedG2P2GKernel (src/mpm_g2p2g_kernels.h:211-335, normalise = false: the
// plain weighted sum over nodes with mass > 1e-10)
__global__ void k_l_g2p(float4* __restrict__ pos, float4* __restrict__ vel, const float* __restrict__ gm,
                        const float* __restrict__ gv, LGeom g, int n, float dt, float flip, bool normalise) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float4 pm = pos[q], vo = vel[q];
    const float p[3] = {pm.x, pm.y, pm.z};
    int b[3];
    float f[3];
    l_base(p, g, b, f);
    const int ks = g.order == 2 ? 3 : 4, off = ks / 2;
    float nv[3] = {0.f, 0.f, 0.f}, tw = 0.f;
    for (int di = -off; di < ks - off; ++di)
        for (int dj = -off; dj < ks - off; ++dj)
            for (int dk = -off; dk < ks - off; ++dk) {
                const int i = b[0] + di, j = b[1] + dj, k = b[2] + dk;
                if (i < 0 || i >= g.d[0] || j < 0 || j >= g.d[1] || k < 0 || k >= g.d[2]) continue;
                const float w = l_weight(g.order, f[0] - di) * l_weight(g.order, f[1] - dj) * l_weight(g.order, f[2] - dk);
                if (w > 1e-10f) {
                    const size_t id = (size_t)i + (size_t)j * g.d[0] + (size_t)k * g.d[0] * g.d[1];
                    if (gm[id] > 1e-10f) {
                        nv[0] += w * gv[3 * id];
                        nv[1] += w * gv[3 * id + 1];
                        nv[2] += w * gv[3 * id + 2];
                        tw += w;
                    }
                }
            }
    if (normalise && tw > 1e-10f) {
        nv[0] /= tw;
        nv[1] /= tw;
        nv[2] /= tw;
    }
    const float bv[3] = {vo.x * flip + nv[0] * (1.f - flip), vo.y * flip + nv[1] * (1.f - flip),
                         vo.z * flip + nv[2] * (1.f - flip)};
    pos[q] = make_float4(p[0] + bv[0] * dt, p[1] + bv[1] * dt, p[2] + bv[2] * dt, pm.w);
    vel[q] = make_float4(bv[0], bv[1], bv[2], 0.f);
}

int check_geom(const float* origin, const float* cell, const int* dims, int n, int order, LGeom& g) {
    if (!origin || !cell || !dims || n < 0 || (order != 2 && order != 3)) return PG_ERR_INVALID;
    for (int a = 0; a < 3; ++a) {
        if (dims[a] < 1 || !(cell[a] > 0.f)) return PG_ERR_INVALID;
        g.o[a] = origin[a];
        g.h[a] = cell[a];
        g.d[a] = dims[a];
    }
    g.order = order;
    return PG_OK;
}

int blocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1 << 20)); }

}  // namespace
}  // namespace pg

extern "C" int pg_launch_optimized_p2g(const float* positions, const float* velocities, float* grid_mass,
                                       float* grid_momentum, const float* origin, const float* cell, const int* dims,
                                       int n, int order, void* stream) {
    pg::LGeom g;
    if (pg::check_geom(origin, cell, dims, n, order, g) != PG_OK) return PG_ERR_INVALID;
    if (n > 0 && (!positions || !velocities || !grid_mass || !grid_momentum)) return PG_ERR_INVALID;
    if (n > 0)
        pg::k_l_p2g<<<pg::blocks(n), 256, 0, (hipStream_t)stream>>>(
            reinterpret_cast<const float4*>(positions), reinterpret_cast<const float4*>(velocities), grid_mass,
            grid_momentum, g, n, 0.f, 0.f, 0.f);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

extern "C" int pg_launch_optimized_g2p(float* positions, float* velocities, const float* grid_mass,
                                       const float* grid_velocity, const float* origin, const float* cell,
                                       const int* dims, int n, float dt, float flip_ratio, int order, void* stream) {
    pg::LGeom g;
    if (pg::check_geom(origin, cell, dims, n, order, g) != PG_OK) return PG_ERR_INVALID;
    if (n > 0 && (!positions || !velocities || !grid_mass || !grid_velocity)) return PG_ERR_INVALID;
    if (n > 0)
        pg::k_l_g2p<<<pg::blocks(n), 256, 0, (hipStream_t)stream>>>(
            reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities), grid_mass, grid_velocity, g, n,
            dt, flip_ratio, true);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

extern "C" int pg_launch_fully_fused_g2p2g(float* positions, float* velocities, float* grid_mass,
                                           float* grid_momentum, float* grid_velocity, const float* origin,
                                           const float* cell, const int* dims, int n, float dt,
                                           const float* gravity, float flip_ratio, int order, void* stream) {
    pg::LGeom g;
    if (pg::check_geom(origin, cell, dims, n, order, g) != PG_OK || !gravity) return PG_ERR_INVALID;
    if (!grid_mass || !grid_momentum || !grid_velocity || (n > 0 && (!positions || !velocities)))
        return PG_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    const int gb = std::min(pg::blocks(G), 16384);
    pg::k_l_clear<<<gb, 256, 0, st>>>(grid_mass, grid_momentum, grid_velocity, G);
    if (n > 0)
        pg::k_l_p2g<<<pg::blocks(n), 256, 0, st>>>(reinterpret_cast<const float4*>(positions),
                                                   reinterpret_cast<const float4*>(velocities), grid_mass,
                                                   grid_momentum, g, n, gravity[0] * dt, gravity[1] * dt,
                                                   gravity[2] * dt);
    pg::k_l_vel<<<gb, 256, 0, st>>>(grid_mass, grid_momentum, grid_velocity, G);
    if (n > 0)
        pg::k_l_g2p<<<pg::blocks(n), 256, 0, st>>>(reinterpret_cast<float4*>(positions),
                                                   reinterpret_cast<float4*>(velocities), grid_mass, grid_velocity, g,
                                                   n, dt, flip_ratio, true);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

namespace pg {
namespace {

// clear -> P2G (gravity on the particle velocity) -> v = p/m -> G2P (plain
// weighted sum) on device arrays
hipError_t header_g2p2g(float4* pos, float4* vel, float* gm, float* gp, float* gv, const LGeom& g, int n, float dt,
                        const float* gravity, float flip, hipStream_t st) {
    const size_t G = (size_t)g.d[0] * g.d[1] * g.d[2];
    const int gb = std::min(blocks(G), 16384);
    k_l_clear<<<gb, 256, 0, st>>>(gm, gp, gv, G);
    if (n > 0)
        k_l_p2g<<<blocks(n), 256, 0, st>>>(pos, vel, gm, gp, g, n, gravity[0] * dt, gravity[1] * dt, gravity[2] * dt);
    k_l_vel<<<gb, 256, 0, st>>>(gm, gp, gv, G);
    if (n > 0) k_l_g2p<<<blocks(n), 256, 0, st>>>(pos, vel, gm, gv, g, n, dt, flip, false);
    return hipGetLastError();
}

}  // namespace
}  // namespace pg

extern "C" int pg_launch_g2p2g(float* positions, float* velocities, float* grid_mass, float* grid_momentum,
                               float* grid_velocity, const float* origin, const float* cell, const int* dims, int n,
                               float dt, const float* gravity, int use_affine_pic, float flip_ratio, int order,
                               int memory, void* stream) {
    pg::LGeom g;
    if (pg::check_geom(origin, cell, dims, n, order, g) != PG_OK || !gravity) return PG_ERR_INVALID;
    if (!grid_mass || !grid_momentum || !grid_velocity || (n > 0 && (!positions || !velocities)))
        return PG_ERR_INVALID;
    if (memory != PG_HOST && memory != PG_DEVICE) return PG_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    const float flip = use_affine_pic ? flip_ratio : 0.f;
    if (memory == PG_DEVICE)
        return pg::header_g2p2g(reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities),
                                grid_mass, grid_momentum, grid_velocity, g, n, dt, gravity, flip, st) == hipSuccess
                   ? PG_OK
                   : PG_ERR_HIP;
    // host arrays: stage through device copies for the duration of the call
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    const size_t pb = 16 * (size_t)std::max(n, 1), gbytes = 4 * G;
    char* buf = nullptr;
    if (hipMalloc(&buf, 2 * pb + 7 * gbytes) != hipSuccess) return PG_ERR_NOMEM;
    float4* dp = reinterpret_cast<float4*>(buf);
    float4* dv = reinterpret_cast<float4*>(buf + pb);
    float* dm = reinterpret_cast<float*>(buf + 2 * pb);
    float* dmom = dm + G;
    float* dvel = dmom + 3 * G;
    hipError_t e = hipSuccess;
    if (n > 0) {
        e = hipMemcpyAsync(dp, positions, 16 * (size_t)n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dv, velocities, 16 * (size_t)n, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) e = pg::header_g2p2g(dp, dv, dm, dmom, dvel, g, n, dt, gravity, flip, st);
    if (e == hipSuccess && n > 0) {
        e = hipMemcpyAsync(positions, dp, 16 * (size_t)n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(velocities, dv, 16 * (size_t)n, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(grid_mass, dm, gbytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(grid_momentum, dmom, 3 * gbytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(grid_velocity, dvel, 3 * gbytes, hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    hipFree(buf);
    return e == hipSuccess && es == hipSuccess ? PG_OK : PG_ERR_HIP;
}
