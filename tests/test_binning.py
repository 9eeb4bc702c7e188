"""The error analysis of the reference's output stage (include/fk_mc/binning.hpp,
include/fk_mc/jackknife.hpp, estimate_bin of prog/data_save.hpp:107-122) in the host
library, pinned by closed forms: hand-computed levels of a short series, the error of the
mean of an AR(1) process, and the jackknife identities (exact standard error for a linear
function, x-bar^2 - s^2/n for the square of a mean)."""
import numpy as np
import pytest

from fk_mc_amd import host_api as host


def test_levels_by_hand():
    t, b = host.binning([1.0, 2.0, 3.0, 4.0, 5.0], 2)
    # level 0: 5 values; level 1: [1.5, 3.5] (the odd value dropped); level 2: [2.5]
    np.testing.assert_allclose(t[0, :4], [5, 3.0, 2.5, np.sqrt(2.5 / 5)], rtol=1e-15)
    np.testing.assert_allclose(t[1, :4], [2, 2.5, 2.0, 1.0], rtol=1e-15)
    assert t[2, 0] == 1 and t[2, 1] == 2.5 and np.isnan(t[2, 2])  # one value: n - 1 = 0
    # correlation length 0.5 (2^D var_D / var_0 - 1)
    np.testing.assert_allclose(t[1, 4], 0.5 * (2 * 2.0 / 2.5 - 1), rtol=1e-15)
    assert t[0, 4] == 0.0


def test_constant_series():
    t, b = host.binning(np.full(64, 1.25), 3)
    assert (t[:, 1] == 1.25).all() and (t[:, 2] == 0.0).all()
    assert list(t[:, 0]) == [64, 32, 16, 8]


def test_ar1_error_converges():
    """x_t = rho x_{t-1} + sqrt(1 - rho^2) eps_t: unit variance, error of the mean
    sqrt((1 + rho) / (1 - rho) / n); level 0 sees only sqrt(1 / n)."""
    rng = np.random.default_rng(4)
    n, rho = 1 << 17, 0.8
    eps = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = eps[0]
    for i in range(1, n):
        x[i] = rho * x[i - 1] + np.sqrt(1 - rho * rho) * eps[i]
    depth = host.max_bin(n)
    assert depth == 12  # int(log2(n / 16) - 1)
    t, b = host.binning(x, depth)
    assert abs(t[0, 3] / np.sqrt(1.0 / n) - 1) < 0.05
    want = np.sqrt((1 + rho) / (1 - rho) / n)
    assert b >= 4, t[:, 3]
    assert abs(t[b, 3] / want - 1) < 0.25, (b, t[b, 3], want)
    # correlation length at the plateau ~ (1 + rho) / (1 - rho) / 2 - 1/2 = rho / (1 - rho)
    assert abs(t[b, 4] / (rho / (1 - rho)) - 1) < 0.5


def test_max_bin():
    assert host.max_bin(1024) == 5
    assert host.max_bin(16) == 1 and host.max_bin(1) == 1
    assert host.max_bin(1 << 30) == 15


def test_jackknife_linear_is_standard_error():
    """With E constant, cv = -beta^2 mean(d2E) / V is linear: the jackknife returns the
    plain mean and standard error on every level."""
    rng = np.random.default_rng(5)
    n, beta, V = 256, 3.0, 16
    d = rng.standard_normal(n)
    t, b = host.jackknife_cv(np.full(n, 0.7), d, beta, V, 3)
    for D in range(4):
        lv = d.reshape(-1, 1 << D).mean(axis=1) if D else d
        m = len(lv)
        assert t[D, 0] == m
        np.testing.assert_allclose(t[D, 1], -beta * beta * lv.mean() / V, rtol=1e-12)
        np.testing.assert_allclose(t[D, 3], beta * beta / V * lv.std(ddof=1) / np.sqrt(m), rtol=1e-10)


def test_jackknife_bias_correction():
    """For <E>^2 the jackknife estimate is x-bar^2 - s^2 / n exactly."""
    rng = np.random.default_rng(6)
    n, beta, V = 100, 2.0, 4
    e = 1.0 + 0.3 * rng.standard_normal(n)
    d = rng.standard_normal(n)
    t, _ = host.jackknife_cv(e, d, beta, V, 0)
    want = beta * beta / V * (np.mean(e * e) - d.mean() - (e.mean() ** 2 - e.var(ddof=1) / n))
    np.testing.assert_allclose(t[0, 1], want, rtol=1e-11)


def test_estimate_bin_walks_down_plateau():
    """i.i.d. data: every level has the same expected error, so the walk from the deepest
    level stops only where two neighbouring errors differ by >= 5 %."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal(1 << 12)
    t, b = host.binning(x, 6)
    err = t[:, 3]
    for i in range(b + 1, len(err)):
        assert abs(err[i - 1] / err[i] - 1) < 0.05
    if b > 0:
        r = abs(err[b - 1] / err[b] - 1)
        prev = abs(err[b] / err[b + 1] - 1) if b + 1 < len(err) else 1.0
        assert r >= 0.05 or r >= prev


def test_binning_errors():
    with pytest.raises(host.FkgpuError):
        host.binning([], 1)
    with pytest.raises(host.FkgpuError):
        host.jackknife_cv([1.0], [1.0], 1.0, 1, 0)
