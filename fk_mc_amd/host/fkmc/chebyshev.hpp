// chebyshev.hpp -- chebyshev_eval with the reference's interface
// (include/fk_mc/chebyshev.hpp:15-63): angle grid theta_i uniform in [0,1],
// Lobatto grid x_i = -cos(pi theta_i), T_k(x_i) tabulated, moment_f(f, k) =
// trapezoid rule in theta of f(x_i) T_k(x_i) (approximating
// (1/pi) int f T_k / sqrt(1-x^2) dx).  cheb_size is rounded up to even.
#pragma once
#include <cmath>
#include <vector>

namespace fk {
namespace chebyshev {

inline double chebyshev_t(double x, size_t n) { return std::cos((double)n * std::acos(x)); }

struct chebyshev_eval {
    chebyshev_eval(int max_moment, int grid_size) : M_(max_moment + max_moment % 2), G_(grid_size) {
        angle_.resize(G_);
        x_.resize(G_);
        T_.resize((size_t)M_ * G_);
        for (int i = 0; i < G_; ++i) {
            angle_[i] = G_ > 1 ? (double)i / (double)(G_ - 1) : 0.0;
            x_[i] = -std::cos(M_PI * angle_[i]);
            for (int k = 0; k < M_; ++k) T_[(size_t)k * G_ + i] = chebyshev_t(x_[i], k);
        }
    }
    int cheb_size() const { return M_; }
    int grid_size() const { return G_; }

    template <class F>
    double moment_f(const F& op, int order) const {
        std::vector<double> vals(G_);
        for (int i = 0; i < G_; ++i) vals[i] = op(x_[i]);
        return moment(vals, order);
    }
    template <class V>
    double moment(const V& in, int order) const {
        double s = 0.0;
        for (int i = 0; i + 1 < G_; ++i)
            s += (in[i + 1] * T_[(size_t)order * G_ + i + 1] + in[i] * T_[(size_t)order * G_ + i]) *
                 (angle_[i + 1] - angle_[i]);
        return s * 0.5;
    }

private:
    int M_, G_;
    std::vector<double> angle_, x_, T_;
};

}  // namespace chebyshev
}  // namespace fk
