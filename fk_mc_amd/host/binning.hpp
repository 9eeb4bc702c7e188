// binning.hpp -- the error analysis of the reference's output stage:
//   * binning (include/fk_mc/binning.hpp:21-160): level D holds floor(n / 2^D) values,
//     value j of level D being (value 2j + value 2j+1 of level D-1) / 2, and reports
//     {size, mean, variance with an n-1 denominator, standard error sqrt(variance/size)};
//   * estimate_bin (prog/data_save.hpp:107-122): walking down from the deepest level, the
//     first level whose error differs from the one above by >= 5 % (or more than the last
//     accepted change) stops the walk;
//   * the integrated correlation length 0.5 (2^D var_D / var_0 - 1) of save_binning
//     (binning.hpp:163-171, whose loop writes past the end of an empty vector: the value
//     it means is computed here);
//   * jackknife (include/fk_mc/jackknife.hpp:40-81): on each binning level, the function of
//     the leave-one-out means, bias-corrected U_0 - (n-1)(mean U_j - U_0), error
//     (n-1) sqrt(var U_j / n).
// The series is passed in the order it is to be binned (the reference bins its histories
// from the last measurement backwards: callers reverse them).
#pragma once
#include <cmath>
#include <functional>
#include <vector>

namespace fkmc {
namespace binning {

struct stats {
    double size, mean, disp, sqerror;
};

inline stats calc_stats(const std::vector<double>& v) {
    const double n = (double)v.size();
    double s = 0.0;
    for (double x : v) s += x;
    const double mean = s / n;
    double q = 0.0;
    for (double y : v) q += (y - mean) * (y - mean);
    const double var = q / (n - 1.0);
    return {n, mean, var, std::sqrt(var / n)};
}

// next binning level: pairwise means, floor(n / 2) values
inline std::vector<double> halve(const std::vector<double>& v) {
    std::vector<double> out(v.size() / 2);
    for (size_t j = 0; j < out.size(); ++j) out[j] = (v[2 * j] + v[2 * j + 1]) / 2.0;
    return out;
}

inline std::vector<double> level(const std::vector<double>& x, int depth) {
    std::vector<double> v = x;
    for (int d = 0; d < depth; ++d) v = halve(v);
    return v;
}

// stats of levels 0..depth
inline std::vector<stats> accumulate(const std::vector<double>& x, int depth) {
    std::vector<stats> out;
    std::vector<double> v = x;
    for (int d = 0; d <= depth; ++d) {
        out.push_back(calc_stats(v));
        v = halve(v);
    }
    return out;
}

inline stats bin(const std::vector<double>& x, int depth) { return calc_stats(level(x, depth)); }

inline int estimate_bin(const std::vector<stats>& data) {
    double rel = 1.0;
    int ind = (int)data.size() - 1;
    bool f = true;
    while (f && ind > 0) {
        const double cur = std::abs(data[ind - 1].sqerror / data[ind].sqerror - 1.0);
        f = (cur < 0.05 && cur < rel);
        if (f) {
            rel = cur;
            --ind;
        }
    }
    return ind;
}

inline std::vector<double> cor_length(const std::vector<stats>& data) {
    std::vector<double> out(data.size());
    for (size_t i = 0; i < data.size(); ++i) out[i] = 0.5 * (std::pow(2.0, (double)i) * data[i].disp / data[0].disp - 1.0);
    return out;
}

// max_bin_ of data_saver (prog/data_save.hpp:90)
inline int max_bin(size_t nmeasures) {
    const int b = (int)(std::log((double)nmeasures / 16) / std::log(2.) - 1);
    return b < 1 ? 1 : (b > 15 ? 15 : b);
}

// jackknife of F over the variables xs (equal lengths), on one binning level
inline stats jack(const std::function<double(const std::vector<double>&)>& F,
                  const std::vector<std::vector<double>>& xs) {
    const size_t L = xs.size(), n = xs[0].size();
    std::vector<double> m(L), loo(L);
    for (size_t i = 0; i < L; ++i) m[i] = calc_stats(xs[i]).mean;
    const double U0 = F(m);
    std::vector<double> U(n);
    for (size_t j = 0; j < n; ++j) {
        for (size_t i = 0; i < L; ++i) loo[i] = ((double)n * m[i] - xs[i][j]) / (double)(n - 1);
        U[j] = F(loo);
    }
    const stats us = calc_stats(U);
    const double avg = U0 - (double)(n - 1) * (us.mean - U0);
    const double dU = (double)(n - 1) * us.sqerror;
    return {(double)n, avg, dU * dU * (double)n, dU};
}

inline std::vector<stats> accumulate_jackknife(const std::function<double(const std::vector<double>&)>& F,
                                               const std::vector<std::vector<double>>& xs, int depth) {
    std::vector<stats> out;
    std::vector<std::vector<double>> v = xs;
    for (int d = 0; d <= depth; ++d) {
        out.push_back(jack(F, v));
        for (auto& s : v) s = halve(s);
    }
    return out;
}

}  // namespace binning
}  // namespace fkmc
