// lattice.hpp -- host-side lattice classes with the reference's interface
// (include/fk_mc/lattice.hpp:12-46, include/fk_mc/lattice/hypercubic.hpp:16-69).
// The hopping matrix is kept in CSR form; the lattice owns the device context
// (fkgpu_ctx) that holds its hopping template, created on first use.
#pragma once
#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <vector>

#include "fkgpu.h"

namespace fk {

// Compressed-row sparse matrix (the role of Eigen::SparseMatrix<double> in the reference).
struct sparse_m {
    int n = 0;
    std::vector<int> row_ptr{0}, col;
    std::vector<double> val;
    size_t rows() const { return (size_t)n; }
    size_t cols() const { return (size_t)n; }
    size_t nonZeros() const { return val.size(); }
    double coeff(int i, int j) const;
};

// Device context bound to one lattice.
struct device_lattice {
    fkgpu_ctx* ctx = nullptr;
    ~device_lattice() { if (ctx) fkgpu_destroy(ctx); }
};

struct abstract_lattice {
    explicit abstract_lattice(size_t norbitals = 1) : norbs_(norbitals) {}
    virtual ~abstract_lattice() = default;
    size_t msize() const { finalize(); return (size_t)hopping_m_.n; }
    size_t norbs() const { return norbs_; }
    size_t volume() const { return msize() / norbs(); }
    const sparse_m& hopping_m() const { finalize(); return hopping_m_; }
    virtual std::vector<size_t> neighbor_index(size_t index) const = 0;
    virtual size_t ndim() const = 0;
    // Sites reachable by one nearest-neighbour f-hop: neighbor_index(i), plus the two
    // diagonal (x-1,y-1), (x+1,y+1) sites on the triangular stencil.
    std::vector<size_t> hop_neighbors(size_t index) const;
    // add_hopping (src/lattice.cpp:9-16)
    abstract_lattice& add_hopping(size_t l, size_t r, double v, bool symmetrize = false);
    abstract_lattice& set_hopping_m(const sparse_m& in);
    // device context with this lattice's hopping template (throws std::logic_error on failure)
    fkgpu_ctx* device() const;

    // Stencil description set by fill_nearest_neighbors / fill_triangular so that the
    // device builds the same hopping matrix from (kind, D, L, t, tp); -1: use the CSR.
    int stencil_kind = -1;
    int stencil_L = 0;
    double stencil_t = 0.0, stencil_tp = 0.0;

protected:
    void finalize() const;  // pending triplets -> CSR
    mutable sparse_m hopping_m_;
    mutable std::vector<std::array<double, 3>> pending_;
    mutable int pending_n_ = 0;
    size_t norbs_;
    mutable std::shared_ptr<device_lattice> dev_;
};

template <size_t D>
struct hypercubic_lattice : public abstract_lattice {
    static constexpr size_t Ndim = D;
    typedef std::array<int, D> pos_t;
    explicit hypercubic_lattice(size_t lattice_size);
    pos_t index_to_pos(size_t index) const;        // src/lattice/hypercubic.cpp:31-39
    size_t pos_to_index(pos_t pos) const;          // :41-51
    std::vector<size_t> neighbor_index(size_t index) const override;  // :67-75
    size_t ndim() const override { return D; }
    const std::array<int, D>& dims() const { return dims_; }

protected:
    std::array<int, D> dims_;
};

// fill_nearest_neighbors (src/lattice/hypercubic.cpp:115-131): H_ij = -t for the +-1 neighbours.
template <size_t D>
hypercubic_lattice<D>& fill_nearest_neighbors(hypercubic_lattice<D>& l, double t);
// fill_triangular (src/lattice/hypercubic.cpp:137-155): square lattice plus -tp on (x-1,y-1),(x+1,y+1).
hypercubic_lattice<2>& fill_triangular(hypercubic_lattice<2>& l, double t, double tp);

}  // namespace fk
