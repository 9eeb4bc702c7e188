"""The library's generator port against libstdc++ itself (host side, no GPU).

The oracle draws with the real std::mt19937 / std::*_distribution objects
(the reference's calls: src/mc_metropolis.cpp:25,41,43, src/moves.cpp:7,55);
fkgpu_rng_draws runs the port that the sweep kernel executes on the device.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("seed", [0, 1, 5489, 4294967295, 32167])
def test_raw_words(fk, orc, seed):
    a = fk.fkgpu.rng_draws(seed, 0, 2000)
    b = orc.rng_draws(seed, 0, 2000)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("hi", [0, 1, 2, 63, 255, 511, 575, 4095, 1000003, 2**31 - 2])
def test_uniform_int(fk, orc, hi):
    a = fk.fkgpu.rng_draws(7, 1, 3000, hi=hi)
    b = orc.rng_draws(7, 1, 3000, hi=hi)
    assert np.array_equal(a, b)


def test_uniform_int_rejection_path(fk, orc):
    # range 2^30 + 1: low words below (2^32 - range) % range = 2^30 - 3 are
    # rejected, about a quarter of the draws
    a = fk.fkgpu.rng_draws(11, 1, 4000, hi=2**30)
    b = orc.rng_draws(11, 1, 4000, hi=2**30)
    assert np.array_equal(a, b)


def test_uniform_real(fk, orc):
    a = fk.fkgpu.rng_draws(3, 2, 5000)
    b = orc.rng_draws(3, 2, 5000)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() < 1.0


@pytest.mark.parametrize("probs", [[1.0], [0.5, 0.5], [0.2, 0.3, 0.5], [1.0, 0.0, 2.0, 1.0]])
def test_discrete(fk, orc, probs):
    a = fk.fkgpu.rng_draws(9, 3, 3000, probs=probs)
    b = orc.rng_draws(9, 3, 3000, probs=probs)
    assert np.array_equal(a, b)


def test_single_move_draws_nothing(fk, orc):
    """discrete_distribution with one weight returns 0 without consuming a word."""
    # draw 10 discrete values then 1 raw word; the raw word must be the first word
    a = fk.fkgpu.rng_draws(5, 3, 10, probs=[1.0])
    assert np.all(a == 0)
