/*
 * Neighbour lists of the spatial hash and SPH-MPM coupling pairs and forces.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates:
 *   SpatialHashGrid::getNeighbors   src/sparse_data_structures.h:193-228
 *       (2R+1)^3 cells around the query's truncated cell coordinates,
 *       R = ceil(radius / cell_size), x outermost, z innermost, cells whose
 *       unsigned coordinates reach res skipped, every particle of each visited
 *       hash slot appended in sorted order; no distance test, no de-duplication
 *   SPHMPMCouplingSystem::findCouplingPairs      src/sph_mpm_coupling.h:454-486
 *       all (sph, mpm) with |x_mpm - x_sph| < radius, SPH-major, MPM ascending,
 *       penetration = radius - distance
 *   SPHMPMCouplingSystem::computeCouplingForces  src/sph_mpm_coupling.h:488-545
 *   SparseMatrix::multiply                       src/sparse_data_structures.h:392-410
 *       per row, sum = sum + value * x[column] in stored order
 *       spring-damper normal force, Coulomb-style friction, two-way reaction
 *       accumulated onto the SPH particle in pair order
 * Float arithmetic in the order the reference writes it (built with
 * -ffp-contract=off).  The float->unsigned conversion of a cell coordinate,
 * undefined in C and C++ below 0 and from 2^32, saturates here (0, 2^32-1).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static unsigned coord(float x, float lo, float cell) {
    const float t = (x - lo) / cell;
    if (!(t > 0.0f)) return 0u;
    return t < 4294967296.0f ? (unsigned)t : 0xFFFFFFFFu;
}

/* Pass 1 (out == NULL): counts[q] = neighbours of query q.  Pass 2: the lists,
 * back to back in query order (offsets from the counts). */
unsigned long long orc_hash_neighbors(const float* queries, unsigned nq, float radius, float cell_size,
                                      const float* dmin, const unsigned* res, const unsigned* particle_indices,
                                      const unsigned* cell_start, const unsigned* cell_count, unsigned* counts,
                                      unsigned* out) {
    const unsigned total = res[0] * res[1] * res[2];
    const unsigned R = (unsigned)ceilf(radius / cell_size);
    unsigned long long k = 0;
    for (unsigned q = 0; q < nq; ++q) {
        unsigned c[3];
        for (int a = 0; a < 3; ++a) c[a] = coord(queries[3 * q + a], dmin[a], cell_size);
        unsigned cnt = 0;
        for (unsigned dx = 0; dx <= 2 * R; ++dx)
            for (unsigned dy = 0; dy <= 2 * R; ++dy)
                for (unsigned dz = 0; dz <= 2 * R; ++dz) {
                    const unsigned gx = c[0] + dx - R, gy = c[1] + dy - R, gz = c[2] + dz - R;
                    if (gx >= res[0] || gy >= res[1] || gz >= res[2]) continue;
                    const unsigned slot = orc_morton3(gx, gy, gz) % total;
                    for (unsigned i = 0; i < cell_count[slot]; ++i) {
                        if (out) out[k] = particle_indices[cell_start[slot] + i];
                        ++k;
                        ++cnt;
                    }
                }
        if (counts) counts[q] = cnt;
    }
    return k;
}

static float dist3(const float* a, const float* b) {
    const float rx = a[0] - b[0], ry = a[1] - b[1], rz = a[2] - b[2];
    return sqrtf(rx * rx + ry * ry + rz * rz);
}

unsigned long long orc_coupling_pairs(const float* sph_x, unsigned ns, const float* mpm_x, unsigned nm, float radius,
                                      unsigned* row_counts, unsigned* col, float* penetration) {
    unsigned long long k = 0;
    for (unsigned s = 0; s < ns; ++s) {
        unsigned cnt = 0;
        for (unsigned m = 0; m < nm; ++m) {
            const float d = dist3(mpm_x + 3 * (size_t)m, sph_x + 3 * (size_t)s);
            if (d < radius) {
                if (col) col[k] = m;
                if (penetration) penetration[k] = radius - d;
                ++k;
                ++cnt;
            }
        }
        if (row_counts) row_counts[s] = cnt;
    }
    return k;
}

void orc_coupling_forces(const float* sph_x, const float* sph_v, unsigned ns, const float* mpm_x,
                         const float* mpm_v, const unsigned* row_ptr, const unsigned* col, const float* penetration,
                         float stiffness, float damping, float friction, int two_way, float* solid_force,
                         float* fluid_force, float* sph_force) {
    for (unsigned s = 0; s < ns; ++s)
        for (unsigned k = row_ptr[s]; k < row_ptr[s + 1]; ++k) {
            const unsigned m = col[k];
            const float* xs = sph_x + 3 * (size_t)s;
            const float* xm = mpm_x + 3 * (size_t)m;
            float* fs = solid_force + 3 * (size_t)k;
            float* ff = fluid_force + 3 * (size_t)k;
            fs[0] = fs[1] = fs[2] = 0.0f;
            ff[0] = ff[1] = ff[2] = 0.0f;
            const float r[3] = {xm[0] - xs[0], xm[1] - xs[1], xm[2] - xs[2]};
            const float d = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (d < 1e-6f) continue;
            const float inv = 1.0f / d;
            const float nrm[3] = {r[0] * inv, r[1] * inv, r[2] * inv};
            float vd[3];
            for (int a = 0; a < 3; ++a) vd[a] = mpm_v[3 * (size_t)m + a] - sph_v[3 * (size_t)s + a];
            const float fn = stiffness * penetration[k];
            const float vn = vd[0] * nrm[0] + vd[1] * nrm[1] + vd[2] * nrm[2];
            const float fd = damping * vn;
            const float sn = fn + fd;
            const float tn[3] = {nrm[0] * sn, nrm[1] * sn, nrm[2] * sn};
            float t[3];
            for (int a = 0; a < 3; ++a) t[a] = vd[a] - nrm[a] * vn;
            const float ts = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            float fr[3] = {0.0f, 0.0f, 0.0f};
            if (ts > 1e-6f) {
                const float it = 1.0f / ts;
                const float fm = friction * fn;
                for (int a = 0; a < 3; ++a) fr[a] = (t[a] * it) * -fm;
            }
            for (int a = 0; a < 3; ++a) {
                fs[a] = tn[a] + fr[a];
                ff[a] = (tn[a] + fr[a]) * -1.0f;
            }
            if (two_way)
                for (int a = 0; a < 3; ++a) sph_force[3 * (size_t)s + a] = sph_force[3 * (size_t)s + a] + ff[a];
        }
}

void orc_csr_spmv(const unsigned* row_ptr, const unsigned* col, const float* val, const float* x, unsigned nrows,
                  float* y) {
    for (unsigned r = 0; r < nrows; ++r) {
        float sum = 0.0f;
        for (unsigned k = row_ptr[r]; k < row_ptr[r + 1]; ++k) sum += val[k] * x[col[k]];
        y[r] = sum;
    }
}
