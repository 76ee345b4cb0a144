"""GPU parity of the group kernel (n lanes per pair, one geometry per call).

The group kernel serves one-geometry calls whose band does not fit one lane:
float32 DTW past 112 slots, float32 WDTW past 72, float64 DTW past 104 and
float64 WDTW past 56 slots.  Policy 2 forces it wherever it applies, so
narrow bands (and every strip size) are covered too.  Bars: float64
bit-identical to the oracle; float32 within 1e-5 relative.
"""

import numpy as np
import pytest

import oracle
from wildboar_amd import _lib
from wildboar_amd import distance as wd

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-5


def rel_err(got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert np.isfinite(got).all()
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-1)))


def _rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / max(float(x.std()), 1e-12)


def _check(X, Y, metric, r, g, precision, policy):
    params = {"r": r} if metric == "dtw" else {"r": r, "g": g}
    _lib.set_policy(policy)
    try:
        got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params,
                                   precision=precision)
    finally:
        _lib.set_policy(0)
    want = oracle.pairwise(X, Y, metric=metric, r=r, g=g)
    got = np.atleast_2d(got)
    if precision == "fp64":
        np.testing.assert_array_equal(got, want)
    else:
        assert rel_err(got, want) <= FP32_RTOL


# bands from 1 slot to ~400 (every strip size of the table appears), equal
# and unequal lengths, rows shorter than the group (lx < n)
GEOMS = [(40, 40, 0.0), (40, 40, 0.05), (60, 60, 0.1), (64, 64, 0.3), (100, 100, 0.5),
         (128, 128, 1.0), (200, 200, 0.6), (23, 180, 0.2), (180, 23, 0.2), (5, 200, 1.0),
         (200, 5, 1.0), (1, 50, 0.5), (50, 1, 0.5), (2, 2, 1.0), (97, 131, 0.37),
         (256, 256, 0.8)]


@pytest.mark.parametrize("lx,ly,r", GEOMS)
@pytest.mark.parametrize("metric", ["dtw", "wdtw"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_group_kernel_forced_matches_oracle(lx, ly, r, metric, precision):
    X = _rw(lx * 7 + ly, 9, lx)
    Y = _rw(ly * 5 + lx + 1, 13, ly)
    _check(X, Y, metric, r, 0.1, precision, 2)


@pytest.mark.parametrize("t,r", [(512, 1.0), (300, 0.25), (700, 0.1)])
@pytest.mark.parametrize("metric", ["dtw", "wdtw"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_group_kernel_default_routing_wide_bands(t, r, metric, precision):
    # bands past the lane classes take the group kernel under the default policy
    X = _rw(300 + t, 5, t)
    Y = _rw(301 + t, 7, t)
    _check(X, Y, metric, r, 0.05, precision, 0)


@pytest.mark.parametrize("metric", ["dtw", "wdtw"])
def test_group_kernel_layouts_and_ragged_waves(metric):
    # singleton (triangle), paired (diagonal) and argmin candidate lists,
    # with pair counts that leave the last wave partly empty
    X = _rw(400, 37, 150)
    Y = _rw(401, 37, 150)
    params = {"r": 1.0} if metric == "dtw" else {"r": 1.0, "g": 0.2}
    for prec in ("fp64", "fp32"):
        _lib.set_policy(2)
        try:
            s = wd.pairwise_distance(X, metric=metric, metric_params=params, precision=prec)
            pr = wd.paired_distance(X, Y, metric=metric, metric_params=params, precision=prec)
            idx, dist = wd.argmin_distance(X, Y, k=3, metric=metric, metric_params=params,
                                           precision=prec, return_distance=True)
        finally:
            _lib.set_policy(0)
        ws = oracle.pairwise(X, None, metric=metric, r=1.0, g=params.get("g", 0.05))
        wp = oracle.paired(X, Y, metric=metric, r=1.0, g=params.get("g", 0.05))
        oi, od = oracle.argmin(X, Y, k=3, metric=metric, r=1.0, g=params.get("g", 0.05))
        if prec == "fp64":
            np.testing.assert_array_equal(s, ws)
            np.testing.assert_array_equal(pr, wp)
        else:
            assert rel_err(s, ws) <= FP32_RTOL and rel_err(pr, wp) <= FP32_RTOL
        np.testing.assert_array_equal(idx, oi)
        np.testing.assert_array_equal(dist, od)


@pytest.mark.parametrize("lx,ly", [(1, 1), (1, 40), (40, 1), (2, 3), (7, 130), (130, 7),
                                   (33, 33), (64, 64), (100, 257), (257, 100), (512, 512),
                                   (777, 600)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_full_matrix_kernel_r1(lx, ly, precision):
    # r = 1 (the reference's default): the band is the whole matrix and the
    # column-owning kernel runs; also the singleton and paired layouts
    X = _rw(500 + lx, 6, lx)
    Y = _rw(501 + ly, 11, ly)
    _check(X, Y, "dtw", 1.0, 0.05, precision, 0)
    if lx == ly:
        for prec in (precision,):
            s = wd.pairwise_distance(X, metric="dtw", precision=prec)
            ws = oracle.pairwise(X, None, metric="dtw", r=1.0)
            pr = wd.paired_distance(X, Y[:6], metric="dtw", precision=prec)
            wp = oracle.paired(X, Y[:6], metric="dtw", r=1.0)
            if prec == "fp64":
                np.testing.assert_array_equal(s, ws)
                np.testing.assert_array_equal(pr, wp)
            else:
                assert rel_err(s, ws) <= FP32_RTOL and rel_err(pr, wp) <= FP32_RTOL
