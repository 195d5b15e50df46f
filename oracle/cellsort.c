/*
 * Cell keys and stable counting sort.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates SpatialHashGrid (src/sparse_data_structures.h):
 *   mortonEncode3D   :121-131  10-bit expansion by multiply-and-mask
 *   getGridCoords    :134-140  truncating float->uint32 conversion of (x-min)/cell,
 *                              saturated (below domain_min: 0)
 *   getCellID        :143-152  clamp to res-1, morton % total_cells
 *   buildHashTable   :155-190  count, exclusive prefix sum, scatter in index order
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static unsigned expand_bits(unsigned v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

unsigned orc_morton3(unsigned x, unsigned y, unsigned z) {
    return expand_bits(x) | (expand_bits(y) << 1) | (expand_bits(z) << 2);
}

void orc_cell_ids(const float* pos, unsigned n, float cell_size, const float* dmin,
                  const unsigned* res, unsigned* ids) {
    const unsigned total = res[0] * res[1] * res[2];
    for (unsigned i = 0; i < n; ++i) {
        unsigned c[3];
        for (int a = 0; a < 3; ++a) {
            const float t = (pos[3 * i + a] - dmin[a]) / cell_size;
            /* float -> unsigned saturated where C leaves it undefined (t < 0, NaN, t >= 2^32) */
            c[a] = !(t > 0.0f) ? 0u : (t < 4294967296.0f ? (unsigned)t : 0xFFFFFFFFu);
            if (c[a] > res[a] - 1) c[a] = res[a] - 1;
        }
        ids[i] = orc_morton3(c[0], c[1], c[2]) % total;
    }
}

void orc_counting_sort(const unsigned* ids, unsigned n, unsigned total_cells,
                       unsigned* particle_indices, unsigned* cell_start, unsigned* cell_count) {
    memset(cell_count, 0, (size_t)total_cells * sizeof(unsigned));
    for (unsigned i = 0; i < n; ++i) cell_count[ids[i]]++;
    cell_start[0] = 0;
    for (unsigned c = 1; c < total_cells; ++c) cell_start[c] = cell_start[c - 1] + cell_count[c - 1];
    unsigned* next = (unsigned*)malloc((size_t)total_cells * sizeof(unsigned));
    memcpy(next, cell_start, (size_t)total_cells * sizeof(unsigned));
    for (unsigned i = 0; i < n; ++i) particle_indices[next[ids[i]]++] = i;
    free(next);
}
