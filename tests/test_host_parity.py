"""The product's host-side adaptation (chili_grid_adapt, chili_grid_split,
chili_alpha_adapt, chili_max_difference in libchili_gpu.so) against the
oracle's restatements of AdaptiveMap::Adapt / Split
(src/Integrator/AdaptiveMap.cc:67-160) and MultiChannel::Adapt /
MaxDifference (src/Channel/MultiChannel.cc:109-154), bit for bit.  These run
on the host: no GPU is needed."""
import numpy as np
import pytest

from chili_amd import lib
from oracle import oracle as O


def random_edges(rng, dims, B, zero_width=False):
    e = np.array([np.concatenate([[0.0], np.sort(rng.uniform(0, 1, B - 1)), [1.0]]) for _ in range(dims)])
    if zero_width and B > 3:
        e[0, 2] = e[0, 1]                  # a zero-width bin
        e[-1, B - 1] = e[-1, B]            # and a zero-width last bin
    return e


def histograms(rng, dims, B):
    """Random rows, a one-hot row, an all-zero row (the reference leaves that
    dimension's edges all zero, AdaptiveMap.cc:93), a row with isolated zeros
    and a row spanning 300 orders of magnitude."""
    h = rng.exponential(1.0, (dims, B))
    if dims > 1:
        h[1] = 0.0
        h[1, rng.integers(0, B)] = 3.5
    if dims > 2:
        h[2] = 0.0
    if dims > 3:
        h[3, ::3] = 0.0
    if dims > 4:
        h[4] = 10.0 ** rng.uniform(-150, 150, B)
    return h


@pytest.mark.parametrize("B", [2, 3, 4, 17, 50, 64, 200])
@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0, 1.5])
def test_grid_adapt_bitwise(B, alpha):
    rng = np.random.default_rng(B * 7 + int(alpha * 10))
    dims = 6
    for trial in range(4):
        e = random_edges(rng, dims, B, zero_width=trial % 2 == 1) if trial else O.uniform_grid(dims, B)
        h = histograms(rng, dims, B)
        got = lib.grid_adapt(e, h, alpha)
        want = O.grid_adapt(e, h, alpha)
        assert np.array_equal(got, want, equal_nan=True), f"B={B} alpha={alpha} trial {trial}: " \
            f"max |d| {np.nanmax(np.abs(got - want))}"
        if B > 2:
            assert np.all(got[2] == 0.0)   # all-zero histogram row: every edge 0 (Q9)


def test_grid_adapt_uniform_fixed_point():
    e = O.uniform_grid(3, 64)
    got = lib.grid_adapt(e, np.ones((3, 64)), 1.0)
    assert np.allclose(got, e, rtol=0, atol=1e-15)


@pytest.mark.parametrize("nsplit", [2, 3, 4])
@pytest.mark.parametrize("B", [2, 50, 64, 100])
def test_grid_split_bitwise(nsplit, B):
    rng = np.random.default_rng(nsplit * 100 + B)
    e = random_edges(rng, 5, B, zero_width=True)
    got = lib.grid_split(e, nsplit)
    assert got.shape == (5, nsplit * B + 1)
    assert np.array_equal(got, O.grid_split(e, nsplit))


def test_grid_split_rejects_bad_factor():
    with pytest.raises(lib.ChiliError):
        lib.grid_split(O.uniform_grid(2, 4), 5)


ALPHA_CASES = {
    "random": lambda r, n: (r.dirichlet(np.ones(n)), r.exponential(1.0, n)),
    "zero_alpha": lambda r, n: (np.where(np.arange(n) % 3 == 0, 0.0, r.uniform(0.1, 1, n)), r.exponential(1.0, n)),
    "zero_train": lambda r, n: (r.dirichlet(np.ones(n)), np.where(np.arange(n) % 2 == 0, 0.0, r.exponential(1.0, n))),
    "min_alpha_clamp": lambda r, n: (r.dirichlet(np.ones(n)), np.concatenate([[1e30], r.uniform(1e-3, 1e-2, n - 1)])),
    "huge_spread": lambda r, n: (r.dirichlet(np.ones(n)), 10.0 ** r.uniform(-200, 200, n)),
}


@pytest.mark.parametrize("case", list(ALPHA_CASES))
@pytest.mark.parametrize("n", [1, 2, 15, 52])
@pytest.mark.parametrize("beta,min_alpha", [(0.25, 1e-5), (0.5, 1e-2), (1.0, 0.0)])
def test_alpha_adapt_bitwise(case, n, beta, min_alpha):
    rng = np.random.default_rng(n + int(beta * 100))
    a, T = ALPHA_CASES[case](rng, n)
    got = lib.alpha_adapt(a, T, beta, min_alpha)
    want = O.alpha_adapt(a, T, beta, min_alpha)
    assert np.array_equal(got, want, equal_nan=True)
    if case == "min_alpha_clamp" and n > 1 and min_alpha > 0:
        # the dominated channels sit at the clamp before renormalisation:
        # min_alpha / (1 + (n - 1) min_alpha)
        assert np.all(got[1:] == got[1])
        assert abs(got[1] * (1 + (n - 1) * min_alpha) / min_alpha - 1) < 1e-5
    if case == "zero_alpha" and n > 1:
        assert np.all(got[np.arange(n) % 3 == 0] == 0.0)   # a zero alpha stays zero (Q21)


@pytest.mark.parametrize("n", [1, 2, 3, 15, 52])
def test_max_difference_bitwise(n):
    rng = np.random.default_rng(n)
    for T in (rng.exponential(1.0, n), np.zeros(n), 10.0 ** rng.uniform(-100, 100, n)):
        assert lib.max_difference(T) == O.max_difference(T)
        if n > 1:
            assert lib.max_difference(T) == np.max(T) - np.min(T)
