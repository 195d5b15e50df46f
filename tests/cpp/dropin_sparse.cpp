// Drives the sparse data-structure drop-in (include/physgrad/sparse_data_structures.h)
// the way a caller of the reference's SpatialHashGrid / SparseMatrix would.
//   hash:   in.bin: float32 cell, dmin[3], dmax[3], radius; uint32 n, nq; float32 pos[3n], q[3nq]
//           out.bin: uint32 cell_ids[n], particle_indices[n], cell_start[C], cell_count[C],
//                    row_ptr[nq+1], neighbors[...]; then the list of query 0 by getNeighbors
//   matrix: in.bin: uint32 nrows, ncols, nnz, rows[nnz], cols[nnz]; float32 vals[nnz], x[ncols]
//           out.bin: uint32 row_ptr[nrows+1]; float32 y[nrows]
#include <cstdio>
#include <cstring>
#include <vector>

#include "physgrad/sparse_data_structures.h"

using namespace physgrad::sparse;

template <typename T>
static bool rd(FILE* f, T* p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <typename T>
static void wr(FILE* f, const T* p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    FILE* f = std::fopen(argv[2], "rb");
    FILE* o = std::fopen(argv[3], "wb");
    if (!f || !o) return 2;
    if (std::strcmp(argv[1], "hash") == 0) {
        float hdr[8];
        unsigned cnt[2];
        if (!rd(f, hdr, 8) || !rd(f, cnt, 2)) return 2;
        const unsigned n = cnt[0], nq = cnt[1];
        std::vector<float> pos(3 * (size_t)n), q(3 * (size_t)nq);
        if (!rd(f, pos.data(), pos.size()) || !rd(f, q.data(), q.size())) return 2;
        SpatialHashGrid<float> h(hdr[0], {hdr[1], hdr[2], hdr[3]}, {hdr[4], hdr[5], hdr[6]}, n);
        bool threw = false;
        try {
            h.getNeighbors(0.f, 0.f, 0.f, hdr[7]);
        } catch (const std::runtime_error&) {
            threw = true;   // before buildHashTable
        }
        h.buildHashTable(pos, n);
        std::vector<unsigned> rp, nb;
        h.getNeighborsBatch(q, hdr[7], rp, nb);
        const auto one = h.getNeighbors(q[0], q[1], q[2], hdr[7]);
        wr(o, h.getParticleCellIDs().data(), n);
        wr(o, h.getParticleIndices().data(), n);
        wr(o, h.getCellStartIndices().data(), h.getTotalCells());
        wr(o, h.getCellParticleCounts().data(), h.getTotalCells());
        wr(o, rp.data(), rp.size());
        wr(o, nb.data(), nb.size());
        wr(o, one.data(), one.size());
        const auto st = h.getStats();
        std::printf("cells %u occupied %u max %u threw %d one %zu\n", h.getTotalCells(), st.occupied_cells,
                    st.max_particles_per_cell, threw ? 1 : 0, one.size());
    } else {
        unsigned hdr[3];
        if (!rd(f, hdr, 3)) return 2;
        const unsigned nr = hdr[0], nc = hdr[1], nnz = hdr[2];
        std::vector<unsigned> rows(nnz), cols(nnz);
        std::vector<float> vals(nnz), x(nc);
        if (!rd(f, rows.data(), nnz) || !rd(f, cols.data(), nnz) || !rd(f, vals.data(), nnz) || !rd(f, x.data(), nc))
            return 2;
        SparseMatrix<float> A(nr, nc, nnz);
        for (unsigned k = 0; k < nnz; ++k) A.addElement(rows[k], cols[k], vals[k]);
        A.finalize();
        const auto y = A.multiply(x);
        wr(o, A.getRowPointers().data(), nr + 1);
        wr(o, y.data(), y.size());
        const auto ms = A.getMemoryStats();
        std::printf("nnz %u efficiency %.6g\n", A.getNNZ(), ms.storage_efficiency);
    }
    std::fclose(f);
    std::fclose(o);
    return 0;
}
