// lattice.cpp -- see fkmc/lattice.hpp.
#include "fkmc/lattice.hpp"

#include <algorithm>
#include <string>

namespace fk {

double sparse_m::coeff(int i, int j) const {
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
        if (col[p] == j) return val[p];
    return 0.0;
}

abstract_lattice& abstract_lattice::add_hopping(size_t l, size_t r, double v, bool symmetrize) {
    pending_.push_back({(double)l, (double)r, v});
    if (symmetrize) pending_.push_back({(double)r, (double)l, v});
    stencil_kind = -1;  // a hand-added hopping no longer matches a stencil
    dev_.reset();
    return *this;
}

abstract_lattice& abstract_lattice::set_hopping_m(const sparse_m& in) {
    hopping_m_ = in;
    pending_.clear();
    stencil_kind = -1;
    dev_.reset();
    return *this;
}

// Triplets are summed per (row, col) as Eigen's setFromTriplets would do.
void abstract_lattice::finalize() const {
    if (pending_.empty()) return;
    const int n = pending_n_;
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    for (int i = 0; i < hopping_m_.n; ++i)
        for (int p = hopping_m_.row_ptr[i]; p < hopping_m_.row_ptr[i + 1]; ++p)
            rows[i].push_back({hopping_m_.col[p], hopping_m_.val[p]});
    for (auto& t : pending_) rows[(int)t[0]].push_back({(int)t[1], t[2]});
    pending_.clear();
    sparse_m m;
    m.n = n;
    m.row_ptr.assign(1, 0);
    for (int i = 0; i < n; ++i) {
        auto& r = rows[i];
        std::sort(r.begin(), r.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < r.size(); ++k) {
            if (k > 0 && r[k].first == r[k - 1].first) {
                m.val.back() += r[k].second;  // duplicate entry of this row: summed
            } else {
                m.col.push_back(r[k].first);
                m.val.push_back(r[k].second);
            }
        }
        m.row_ptr.push_back((int)m.col.size());
    }
    hopping_m_ = std::move(m);
}

fkgpu_ctx* abstract_lattice::device() const {
    finalize();
    if (dev_ && dev_->ctx) return dev_->ctx;
    auto d = std::make_shared<device_lattice>();
    if (fkgpu_create(0, &d->ctx) != FKGPU_OK) throw std::logic_error("fkgpu_create failed: no HIP device");
    int r;
    if (stencil_kind >= 0)
        r = fkgpu_set_lattice(d->ctx, stencil_kind, (int)ndim(), stencil_L, stencil_t, stencil_tp);
    else
        r = fkgpu_set_lattice_csr(d->ctx, hopping_m_.n, hopping_m_.row_ptr.data(), hopping_m_.col.data(),
                                  hopping_m_.val.data());
    if (r != FKGPU_OK) throw std::logic_error(std::string("fkgpu lattice setup failed: ") + fkgpu_last_error(d->ctx));
    dev_ = d;
    return dev_->ctx;
}

std::vector<size_t> abstract_lattice::hop_neighbors(size_t index) const {
    std::vector<size_t> out = neighbor_index(index);
    if (stencil_kind == FKGPU_LATTICE_TRIANGULAR) {
        const int L = stencil_L, x = (int)index / L, y = (int)index % L;
        out.push_back((size_t)(((x + L - 1) % L) * L + (y + L - 1) % L));
        out.push_back((size_t)(((x + 1) % L) * L + (y + 1) % L));
    }
    return out;
}

template <size_t D>
hypercubic_lattice<D>::hypercubic_lattice(size_t lattice_size) : abstract_lattice(1) {
    dims_.fill((int)lattice_size);
    size_t n = 1;
    for (size_t i = 0; i < D; ++i) n *= lattice_size;
    hopping_m_.n = (int)n;
    hopping_m_.row_ptr.assign(n + 1, 0);
    pending_n_ = (int)n;
}

template <size_t D>
typename hypercubic_lattice<D>::pos_t hypercubic_lattice<D>::index_to_pos(size_t index) const {
    pos_t out;
    for (int i = (int)D - 1; i >= 0; --i) {
        out[i] = (int)(index % dims_[i]);
        index /= dims_[i];
    }
    return out;
}

template <size_t D>
size_t hypercubic_lattice<D>::pos_to_index(pos_t pos) const {
    size_t out = 0, mult = 1;
    for (int i = (int)D - 1; i >= 0; --i) {
        out += pos[i] * mult;
        mult *= dims_[i];
    }
    return out;
}

template <size_t D>
std::vector<size_t> hypercubic_lattice<D>::neighbor_index(size_t index) const {
    std::vector<size_t> out;
    const pos_t p = index_to_pos(index);
    for (size_t d = 0; d < D; ++d) {
        pos_t l = p, r = p;
        l[d] = p[d] == 0 ? dims_[d] - 1 : p[d] - 1;
        r[d] = p[d] == dims_[d] - 1 ? 0 : p[d] + 1;
        out.push_back(pos_to_index(l));
        out.push_back(pos_to_index(r));
    }
    return out;
}

template <size_t D>
hypercubic_lattice<D>& fill_nearest_neighbors(hypercubic_lattice<D>& l, double t) {
    for (size_t i = 0; i < l.volume(); ++i)
        for (size_t j : l.neighbor_index(i)) l.add_hopping(i, j, -t);
    l.stencil_kind = FKGPU_LATTICE_HYPERCUBIC;
    l.stencil_L = l.dims()[0];
    l.stencil_t = t;
    l.stencil_tp = 0.0;
    return l;
}

hypercubic_lattice<2>& fill_triangular(hypercubic_lattice<2>& l, double t, double tp) {
    fill_nearest_neighbors(l, t);
    const int L0 = l.dims()[0], L1 = l.dims()[1];
    for (size_t i = 0; i < l.volume(); ++i) {
        auto p = l.index_to_pos(i);
        std::array<int, 2> pl{p[0] > 0 ? p[0] - 1 : L0 - 1, p[1] > 0 ? p[1] - 1 : L1 - 1};
        std::array<int, 2> pr{p[0] < L0 - 1 ? p[0] + 1 : 0, p[1] < L1 - 1 ? p[1] + 1 : 0};
        l.add_hopping(i, l.pos_to_index(pl), -tp);
        l.add_hopping(i, l.pos_to_index(pr), -tp);
    }
    l.stencil_kind = FKGPU_LATTICE_TRIANGULAR;
    l.stencil_tp = tp;
    return l;
}

template struct hypercubic_lattice<1>;
template struct hypercubic_lattice<2>;
template struct hypercubic_lattice<3>;
template hypercubic_lattice<1>& fill_nearest_neighbors(hypercubic_lattice<1>&, double);
template hypercubic_lattice<2>& fill_nearest_neighbors(hypercubic_lattice<2>&, double);
template hypercubic_lattice<3>& fill_nearest_neighbors(hypercubic_lattice<3>&, double);

}  // namespace fk
