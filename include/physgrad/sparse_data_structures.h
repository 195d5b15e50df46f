// Drop-in for the reference's sparse data-structure header
// (src/sparse_data_structures.h): SpatialHashGrid<T, IndexType> (:44-294) and
// SparseMatrix<T, IndexType> (:304-443), same namespace, names and members.
// The hierarchical octree of the same header (:454-610) is not on the MPM
// path and is not provided.
//
// The work runs on the GPU through the C-ABI (include/physgrad_mpm.h):
//   buildHashTable -> pg_spatial_hash_build (bit-identical cell ids, order,
//                     starts and counts); the host vectors are refreshed
//   getNeighbors   -> pg_hash_neighbors_count/fill (the reference's list);
//                     getNeighborsBatch (not in the reference) answers many
//                     points at once as a CSR list
//   multiply       -> pg_csr_spmv (each row summed in stored order, float)
// Differences from the reference, each because its version cannot work as
// written:
//   * SparseMatrix::addElement drops the row index and finalize() leaves every
//     row pointer at 0, so the reference's multiply returns zeros; here
//     addElement records the row (elements must come in row-major order, as
//     the reference's comment requires; an earlier row throws) and finalize()
//     builds the row pointers from the counts.
//   * Cell coordinates saturate where the reference's float -> unsigned
//     conversion is undefined (points below domain_min get coordinate 0).
// T must be float and IndexType uint32_t (the GPU kernels' types).  Needs the
// HIP runtime headers (-D__HIP_PLATFORM_AMD__ -I/opt/rocm/include) and links
// with libphysgrad_mpm.so and libamdhip64.so.
#ifndef PHYSGRAD_DROPIN_SPARSE_DATA_STRUCTURES_H
#define PHYSGRAD_DROPIN_SPARSE_DATA_STRUCTURES_H

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../physgrad_mpm.h"

namespace physgrad::sparse {

namespace detail {

inline void check_pg(int rc, const char* what) {
    if (rc != PG_OK) throw std::runtime_error(std::string(what) + ": " + pg_mpm_last_error(nullptr));
}
inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// owning device buffer of n elements of T
template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void resize(size_t n) {
        if (n <= cap_ && p_) {
            n_ = n;
            return;
        }
        release();
        check_hip(hipMalloc(reinterpret_cast<void**>(&p_), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        cap_ = std::max<size_t>(n, 1);
        n_ = n;
    }
    void upload(const T* h, size_t n) {
        resize(n);
        if (n) check_hip(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
    void download(T* h, size_t n) const {
        if (n) check_hip(hipMemcpy(h, p_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        cap_ = n_ = 0;
    }
    T* p_ = nullptr;
    size_t cap_ = 0, n_ = 0;
};

}  // namespace detail

template <typename T, typename IndexType = uint32_t>
class SpatialHashGrid {
    static_assert(std::is_same<T, float>::value && std::is_same<IndexType, uint32_t>::value,
                  "the GPU spatial hash works on float positions and uint32_t indices");

public:
    using index_type = IndexType;
    using float_type = T;

    SpatialHashGrid(T cell_size, const std::array<T, 3>& domain_min, const std::array<T, 3>& domain_max,
                    IndexType max_particles = 10000000)
        : cell_size_(cell_size), domain_min_(domain_min), domain_max_(domain_max), max_particles_(max_particles) {
        for (int i = 0; i < 3; ++i)
            grid_resolution_[i] = static_cast<IndexType>(std::ceil((domain_max_[i] - domain_min_[i]) / cell_size_));
        total_cells_ = grid_resolution_[0] * grid_resolution_[1] * grid_resolution_[2];
        cell_particle_counts_.assign(total_cells_, 0);
        cell_start_indices_.assign(total_cells_, 0);
        d_start_.resize(total_cells_);
        d_count_.resize(total_cells_);
    }

    IndexType mortonEncode3D(IndexType x, IndexType y, IndexType z) const {
        auto expand = [](IndexType v) -> IndexType {
            v = (v * 0x00010001u) & 0xFF0000FFu;
            v = (v * 0x00000101u) & 0x0F00F00Fu;
            v = (v * 0x00000011u) & 0xC30C30C3u;
            v = (v * 0x00000005u) & 0x49249249u;
            return v;
        };
        return expand(x) | (expand(y) << 1) | (expand(z) << 2);
    }

    std::array<IndexType, 3> getGridCoords(T x, T y, T z) const {
        const T p[3] = {x, y, z};
        std::array<IndexType, 3> c{};
        for (int a = 0; a < 3; ++a) {
            const T t = (p[a] - domain_min_[a]) / cell_size_;
            c[a] = !(t > T{0}) ? 0u : (t < T{4294967296.0f} ? static_cast<IndexType>(t) : 0xFFFFFFFFu);
        }
        return c;
    }

    IndexType getCellID(T x, T y, T z) const {
        auto c = getGridCoords(x, y, z);
        for (int a = 0; a < 3; ++a) c[a] = std::min(c[a], grid_resolution_[a] - 1);
        return mortonEncode3D(c[0], c[1], c[2]) % total_cells_;
    }

    void buildHashTable(const std::vector<T>& positions, IndexType num_particles) {
        if (num_particles > max_particles_) throw std::runtime_error("buildHashTable: more than max_particles");
        if (positions.size() < 3 * static_cast<size_t>(num_particles))
            throw std::runtime_error("buildHashTable: fewer than 3 num_particles positions");
        d_pos_.upload(positions.data(), 3 * static_cast<size_t>(num_particles));
        d_ids_.resize(num_particles);
        d_idx_.resize(num_particles);
        detail::DeviceArray<unsigned char> ws(pg_spatial_hash_workspace_size(num_particles));
        const float lo[3] = {domain_min_[0], domain_min_[1], domain_min_[2]};
        const unsigned res[3] = {grid_resolution_[0], grid_resolution_[1], grid_resolution_[2]};
        detail::check_pg(pg_spatial_hash_build(d_pos_.get(), num_particles, cell_size_, lo, res, d_ids_.get(),
                                               d_idx_.get(), d_start_.get(), d_count_.get(), ws.get(), ws.size(),
                                               nullptr),
                         "buildHashTable");
        detail::check_hip(hipDeviceSynchronize(), "buildHashTable");
        num_particles_ = num_particles;
        built_ = true;
        particle_indices_.resize(num_particles);
        particle_cell_ids_.resize(num_particles);
        d_idx_.download(particle_indices_.data(), num_particles);
        d_ids_.download(particle_cell_ids_.data(), num_particles);
        d_start_.download(cell_start_indices_.data(), total_cells_);
        d_count_.download(cell_particle_counts_.data(), total_cells_);
    }

    // The reference's list for one point (computed on the GPU).
    std::vector<IndexType> getNeighbors(T x, T y, T z, T radius) const {
        std::vector<IndexType> row_ptr, nb;
        const std::vector<T> q = {x, y, z};
        getNeighborsBatch(q, radius, row_ptr, nb);
        return nb;
    }

    // not in the reference: getNeighbors for every point of queries [3 nq] at
    // once; row i of the CSR list (row_ptr [nq+1]) is the list of point i
    void getNeighborsBatch(const std::vector<T>& queries, T radius, std::vector<IndexType>& row_ptr,
                           std::vector<IndexType>& neighbors) const {
        if (!built_) throw std::runtime_error("getNeighbors: buildHashTable first");
        const unsigned nq = static_cast<unsigned>(queries.size() / 3);
        detail::DeviceArray<float> dq;
        dq.upload(queries.data(), 3 * static_cast<size_t>(nq));
        detail::DeviceArray<unsigned> drp(nq + 1);
        detail::DeviceArray<unsigned char> ws(pg_hash_neighbors_workspace_size(nq));
        const float lo[3] = {domain_min_[0], domain_min_[1], domain_min_[2]};
        const unsigned res[3] = {grid_resolution_[0], grid_resolution_[1], grid_resolution_[2]};
        unsigned long long total = 0;
        detail::check_pg(pg_hash_neighbors_count(dq.get(), nq, radius, cell_size_, lo, res, d_count_.get(), drp.get(),
                                                 &total, ws.get(), ws.size(), nullptr),
                         "getNeighbors");
        detail::DeviceArray<unsigned> dnb(total);
        detail::check_pg(pg_hash_neighbors_fill(dq.get(), nq, radius, cell_size_, lo, res,
                                                d_idx_.get(), d_start_.get(),
                                                d_count_.get(), drp.get(), dnb.get(), nullptr),
                         "getNeighbors");
        detail::check_hip(hipDeviceSynchronize(), "getNeighbors");
        row_ptr.resize(nq + 1);
        drp.download(row_ptr.data(), nq + 1);
        neighbors.resize(total);
        dnb.download(neighbors.data(), total);
    }

    struct HashTableStats {
        T average_particles_per_cell;
        T load_factor;
        IndexType max_particles_per_cell;
        IndexType occupied_cells;
        T memory_usage_mb;
    };

    HashTableStats getStats() const {
        HashTableStats s{};
        IndexType occupied = 0, mx = 0, tot = 0;
        for (IndexType c : cell_particle_counts_) {
            if (c > 0) {
                ++occupied;
                mx = std::max(mx, c);
                tot += c;
            }
        }
        s.occupied_cells = occupied;
        s.max_particles_per_cell = mx;
        s.average_particles_per_cell = static_cast<T>(tot) / std::max(IndexType(1), occupied);
        s.load_factor = static_cast<T>(occupied) / total_cells_;
        const size_t bytes = (2 * static_cast<size_t>(total_cells_) + 2 * static_cast<size_t>(max_particles_)) *
                             sizeof(IndexType);
        s.memory_usage_mb = static_cast<T>(bytes) / (1024.0f * 1024.0f);
        return s;
    }

    T getCellSize() const { return cell_size_; }
    IndexType getTotalCells() const { return total_cells_; }
    const std::vector<IndexType>& getParticleIndices() const { return particle_indices_; }
    const std::vector<IndexType>& getCellStartIndices() const { return cell_start_indices_; }
    const std::vector<IndexType>& getCellParticleCounts() const { return cell_particle_counts_; }
    const std::vector<IndexType>& getParticleCellIDs() const { return particle_cell_ids_; }
    IndexType* getDeviceCellStartIndices() const { return d_start_.get(); }
    IndexType* getDeviceParticleIndices() const { return d_idx_.get(); }
    IndexType* getDeviceCellParticleCounts() const { return d_count_.get(); }

private:
    T cell_size_;
    std::array<T, 3> domain_min_, domain_max_;
    std::array<IndexType, 3> grid_resolution_{};
    IndexType total_cells_ = 0;
    IndexType max_particles_;
    IndexType num_particles_ = 0;
    bool built_ = false;
    std::vector<IndexType> cell_particle_counts_, cell_start_indices_, particle_indices_, particle_cell_ids_;
    detail::DeviceArray<float> d_pos_;
    detail::DeviceArray<unsigned> d_ids_, d_idx_, d_start_, d_count_;
};

template <typename T, typename IndexType = uint32_t>
class SparseMatrix {
    static_assert(std::is_same<T, float>::value && std::is_same<IndexType, uint32_t>::value,
                  "the GPU product works on float values and uint32_t indices");

public:
    SparseMatrix(IndexType num_rows, IndexType num_cols, IndexType estimated_nnz = 0)
        : num_rows_(num_rows), num_cols_(num_cols), nnz_(0) {
        if (estimated_nnz > 0) {
            values_.reserve(estimated_nnz);
            column_indices_.reserve(estimated_nnz);
        }
        row_pointers_.assign(static_cast<size_t>(num_rows) + 1, 0);
    }

    // Elements in row-major order (rows non-decreasing), as the reference requires.
    void addElement(IndexType row, IndexType col, T value) {
        if (row >= num_rows_ || col >= num_cols_) throw std::out_of_range("addElement: index out of range");
        if (!rows_.empty() && row < rows_.back()) throw std::runtime_error("addElement: rows must not decrease");
        values_.push_back(value);
        column_indices_.push_back(col);
        rows_.push_back(row);
        nnz_++;
        finalized_ = false;
    }

    void finalize() {
        std::fill(row_pointers_.begin(), row_pointers_.end(), IndexType{0});
        for (IndexType r : rows_) row_pointers_[static_cast<size_t>(r) + 1]++;
        for (IndexType i = 1; i <= num_rows_; ++i) row_pointers_[i] += row_pointers_[i - 1];
        d_rp_.upload(row_pointers_.data(), row_pointers_.size());
        d_col_.upload(column_indices_.data(), nnz_);
        d_val_.upload(values_.data(), nnz_);
        finalized_ = true;
    }

    // y = A x on the GPU (each row summed in stored order, in float)
    std::vector<T> multiply(const std::vector<T>& x) const {
        if (x.size() != num_cols_) throw std::runtime_error("multiply: x has the wrong size");
        if (!finalized_) throw std::runtime_error("multiply: finalize() first");
        detail::DeviceArray<float> dx, dy(num_rows_);
        dx.upload(x.data(), x.size());
        detail::check_pg(pg_csr_spmv(d_rp_.get(), nnz_ ? d_col_.get() : nullptr, nnz_ ? d_val_.get() : nullptr,
                                     dx.get(), num_cols_, dy.get(), num_rows_, nullptr),
                         "multiply");
        detail::check_hip(hipDeviceSynchronize(), "multiply");
        std::vector<T> y(num_rows_);
        dy.download(y.data(), num_rows_);
        return y;
    }

    struct MemoryStats {
        T storage_efficiency;
        T memory_usage_mb;
        T compression_ratio;
    };

    MemoryStats getMemoryStats() const {
        MemoryStats s;
        const size_t sparse_bytes = nnz_ * (sizeof(T) + sizeof(IndexType)) + (num_rows_ + 1) * sizeof(IndexType);
        const size_t dense_bytes = static_cast<size_t>(num_rows_) * num_cols_ * sizeof(T);
        s.memory_usage_mb = static_cast<T>(sparse_bytes) / (1024.0f * 1024.0f);
        s.compression_ratio = static_cast<T>(dense_bytes) / sparse_bytes;
        s.storage_efficiency = static_cast<T>(nnz_) / (static_cast<T>(num_rows_) * num_cols_);
        return s;
    }

    IndexType getNumRows() const { return num_rows_; }
    IndexType getNumCols() const { return num_cols_; }
    IndexType getNNZ() const { return nnz_; }
    const std::vector<T>& getValues() const { return values_; }
    const std::vector<IndexType>& getColumnIndices() const { return column_indices_; }
    const std::vector<IndexType>& getRowPointers() const { return row_pointers_; }
    // not in the reference: the device copies (valid after finalize)
    const IndexType* getDeviceRowPointers() const { return d_rp_.get(); }
    const IndexType* getDeviceColumnIndices() const { return d_col_.get(); }
    const T* getDeviceValues() const { return d_val_.get(); }

private:
    std::vector<T> values_;
    std::vector<IndexType> column_indices_, row_pointers_, rows_;
    IndexType num_rows_, num_cols_, nnz_;
    bool finalized_ = false;
    detail::DeviceArray<unsigned> d_rp_, d_col_;
    detail::DeviceArray<float> d_val_;
};

}  // namespace physgrad::sparse

#endif
