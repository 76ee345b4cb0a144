"""GPU parity of the z-normalised subsequence scans (pairwise_subsequence_distance
with scale=True: ScaledEuclideanSubsequenceMetric, ScaledDtwSubsequenceMetric
and its LB_Kim / LB_Keogh cascade) against the float64 oracle.  Bar:
distances and indices bit-identical, for the register-band kernel (policy 0)
and the scratch-row kernel (policy 1)."""

import numpy as np
import pytest

import oracle
from wildboar_amd import _lib
from wildboar_amd import distance as wd

pytestmark = pytest.mark.gpu


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.standard_normal((n, t)), axis=1)


def patterns(X, seed):
    rng = np.random.default_rng(seed)
    t = X.shape[1]
    out = []
    for L in (1, 2, 3, 5, 6, 7, 13, 30, 64, t):
        s = int(rng.integers(0, t - L + 1))
        out.append(X[int(rng.integers(0, X.shape[0])), s:s + L] * rng.uniform(0.5, 2.0)
                   + rng.normal())
    out += [rng.standard_normal(9), np.full(8, 1.5), rng.standard_normal(40)]
    return out


def setup_module(module):
    _lib.set_policy(0)


def teardown_module(module):
    _lib.set_policy(0)


def run(pats, X, metric, params, policy):
    _lib.set_policy(policy)
    try:
        return wd.pairwise_subsequence_distance(pats, X, metric=metric, metric_params=params,
                                                scale=True, return_index=True)
    finally:
        _lib.set_policy(0)


@pytest.mark.parametrize("r", [0.0, 0.05, 0.1, 0.3, 1.0])
@pytest.mark.parametrize("policy", [0, 1])
def test_scaled_dtw_bitwise(r, policy):
    X = rw(900, 11, 150)
    pats = patterns(X, 901)
    d, i = run(pats, X, "dtw", {"r": r}, policy)
    od, oi = oracle.scaled_subsequence(pats, X, "dtw", r=r)
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)


def test_scaled_euclidean_bitwise_and_names():
    X = rw(902, 13, 200)
    pats = patterns(X, 903)
    d, i = run(pats, X, "euclidean", None, 0)
    od, oi = oracle.scaled_subsequence(pats, X, "euclidean")
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)
    d2 = wd.pairwise_subsequence_distance(pats, X, metric="scaled_euclidean")
    np.testing.assert_array_equal(d2, od)
    d3 = wd.pairwise_subsequence_distance(pats[6], X, metric="scaled_dtw",
                                          metric_params={"r": 0.2})
    np.testing.assert_array_equal(d3, oracle.scaled_subsequence([pats[6]], X, "dtw", r=0.2)[0][:, 0])


def test_scaled_dtw_where_lb_kim_prunes_the_best_window():
    # shapes where the reference's LB_Kim exceeds the DTW of the best window
    # (the oracle's cascade differs from the exhaustive minimum there)
    rng = np.random.default_rng(1)
    X = np.cumsum(rng.standard_normal((40, 300)), axis=1)
    pats = []
    for _ in range(30):
        L = int(rng.integers(6, 60))
        s = int(rng.integers(0, 300 - L))
        pats.append(X[int(rng.integers(0, 40)), s:s + L] * rng.uniform(0.5, 2) + rng.normal())
    od, oi = oracle.scaled_subsequence(pats, X, "dtw", r=0.3)
    ed, _ = oracle.scaled_subsequence(pats, X, "dtw", r=0.3, cascade=False)
    assert (od != ed).sum() > 0
    for policy in (0, 1):
        d, i = run(pats, X, "dtw", {"r": 0.3}, policy)
        np.testing.assert_array_equal(d, od)
        np.testing.assert_array_equal(i, oi)


def test_scaled_float32_inputs_and_validation():
    X = rw(904, 6, 90).astype(np.float32)
    pats = [X[1, 10:30].astype(np.float64), np.arange(12.0)]
    d, i = wd.pairwise_subsequence_distance(pats, X, metric="dtw", metric_params={"r": 0.1},
                                            scale=True, return_index=True)
    od, oi = oracle.scaled_subsequence(pats, X.astype(np.float64), "dtw", r=0.1)
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)
    with pytest.raises(NotImplementedError):
        wd.pairwise_subsequence_distance(pats, X, metric="scaled_lcss")
    with pytest.raises(ValueError):
        wd.pairwise_subsequence_distance(pats, X, metric="scaled_nope")
    with pytest.raises(ValueError):
        wd.pairwise_subsequence_distance(pats, X, metric="dtw", metric_params={"r": 1.5},
                                         scale=True)
    with pytest.raises(TypeError):
        wd.pairwise_subsequence_distance(pats, X, metric="euclidean", metric_params={"r": 0.1},
                                         scale=True)


@pytest.mark.parametrize("metric,params", [("euclidean", None), ("dtw", {"r": 0.1}),
                                           ("dtw", {"r": 0.5}), ("dtw", {"r": 1.0})])
@pytest.mark.parametrize("policy", [0, 1])
def test_scaled_subsequence_match_bitwise(metric, params, policy):
    X = rw(905, 7, 160)
    s = X[3, 40:72] * 2.5 - 1.0
    r = (params or {}).get("r", 1.0)
    _lib.set_policy(policy)
    try:
        for threshold in (0.5, 2.0, 4.0, np.inf):
            idx, dist = wd.subsequence_match(s, X, threshold, metric=metric, metric_params=params,
                                             scale=True, return_distance=True)
            for q in range(X.shape[0]):
                oi, od = oracle.scaled_subsequence_matches(s, X[q], threshold, metric, r)
                if oi.size == 0:
                    assert idx[q] is None and dist[q] is None
                else:
                    np.testing.assert_array_equal(idx[q], oi)
                    np.testing.assert_array_equal(dist[q], od)
        # default: the 10 closest windows of the full profile
        idx = wd.subsequence_match(s, X, metric=metric, metric_params=params, scale=True)
        oi, od = oracle.scaled_subsequence_matches(s, X[0], np.inf, metric, r)
        np.testing.assert_array_equal(idx[0], oi[np.argsort(od)[:10]])
    finally:
        _lib.set_policy(0)


@pytest.mark.parametrize("metric,r", [("euclidean", 1.0), ("dtw", 0.2)])
def test_scaled_paired_subsequence_match_bitwise(metric, r):
    X = rw(906, 6, 120)
    rng = np.random.default_rng(907)
    pats = [X[q, 10 + 5 * q:10 + 5 * q + L] * 1.5 + 0.2
            for q, L in enumerate((8, 20, 20, 33, 5, 64))]
    pats[1] = rng.standard_normal(20)
    params = None if metric == "euclidean" else {"r": r}
    idx, dist = wd.paired_subsequence_match(pats, X, 3.0, metric=metric, metric_params=params,
                                            scale=True, return_distance=True)
    for q in range(X.shape[0]):
        oi, od = oracle.scaled_subsequence_matches(pats[q], X[q], 3.0, metric, r)
        if oi.size == 0:
            assert idx[q] is None
        else:
            np.testing.assert_array_equal(idx[q], oi)
            np.testing.assert_array_equal(dist[q], od)


@pytest.mark.parametrize("metric,params", [("wdtw", {"r": 0.2, "g": 0.1}),
                                           ("wdtw", {"r": 1.0, "g": 0.05}),
                                           ("ddtw", {"r": 0.1}), ("ddtw", {"r": 1.0}),
                                           ("wddtw", {"r": 0.3, "g": -0.2}),
                                           ("adtw", {"r": 0.2, "p": 0.5}),
                                           ("adtw", {"r": 1.0, "p": 0.0})])
def test_scaled_wrapped_metrics_bitwise(metric, params):
    X = rw(908, 9, 130)
    pats = patterns(X, 909)
    d, i = wd.pairwise_subsequence_distance(pats, X, metric=metric, metric_params=params,
                                            scale=True, return_index=True)
    od, oi = oracle.scaled_subsequence(pats, X, metric, r=params.get("r", 1.0),
                                       g=params.get("g", 0.05), p=params.get("p", 1.0))
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)
    d2 = wd.pairwise_subsequence_distance(pats[:4], X, metric="scaled_" + metric,
                                          metric_params=params)
    np.testing.assert_array_equal(d2, od[:, :4])


def test_scaled_wrapped_refusals():
    X = rw(910, 3, 40)
    with pytest.raises(NotImplementedError):
        wd.pairwise_subsequence_distance([X[0, :5]], X, metric="adtw",
                                         metric_params={"p": -1.0}, scale=True)
    with pytest.raises(NotImplementedError):
        wd.subsequence_match(X[0, :5], X, 1.0, metric="adtw", metric_params={"p": -0.5},
                             scale=True)


@pytest.mark.parametrize("metric,params", [("euclidean", None), ("dtw", {"r": 0.2}),
                                           ("dtw", {"r": 1.0}), ("wdtw", {"r": 0.5, "g": 0.1}),
                                           ("ddtw", {"r": 0.3}), ("wddtw", {"r": 0.3, "g": 0.3}),
                                           ("adtw", {"r": 0.4, "p": 0.2})])
@pytest.mark.parametrize("k", [1, 4])
def test_scaled_argmin_subsequence_bitwise(metric, params, k):
    X = rw(911, 8, 100)
    rng = np.random.default_rng(912)
    pats = [X[q, 3 * q:3 * q + L] * 2.0 + 1.0 for q, L in enumerate((5, 12, 30, 2, 17, 40, 8, 25))]
    pats[3] = rng.standard_normal(2)
    pats[6] = np.full(8, 3.0)  # constant: fast_mean_std gives std 0 -> 1.0
    p = params or {}
    idx, dist = wd.argmin_subsequence_distance(pats, X, k=k, metric=metric, metric_params=params,
                                               scale=True, return_distance=True)
    oi, od = oracle.scaled_argmin_subsequence(pats, X, k=k, metric=metric, r=p.get("r", 1.0),
                                              g=p.get("g", 0.05), p=p.get("p", 1.0))
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, od)
    idx2 = wd.argmin_subsequence_distance(pats, X, k=k, metric="scaled_" + metric,
                                          metric_params=params)
    np.testing.assert_array_equal(idx2, oi)


@pytest.mark.parametrize("metric,params", [("wdtw", {"r": 0.3, "g": 0.1}), ("ddtw", {"r": 0.2}),
                                           ("wddtw", {"r": 1.0, "g": -0.1}),
                                           ("adtw", {"r": 0.5, "p": 0.3})])
def test_scaled_wrapped_matching_bitwise(metric, params):
    X = rw(913, 5, 110)
    s = X[2, 30:55] * 0.7 + 4.0
    p = params
    for threshold in (0.5, 1.5, np.inf):
        idx, dist = wd.subsequence_match(s, X, threshold, metric=metric, metric_params=params,
                                         scale=True, return_distance=True)
        for q in range(X.shape[0]):
            oi, od = oracle.scaled_subsequence_matches(s, X[q], threshold, metric,
                                                       p.get("r", 1.0), p.get("g", 0.05),
                                                       p.get("p", 1.0))
            if oi.size == 0:
                assert idx[q] is None
            else:
                np.testing.assert_array_equal(idx[q], oi)
                np.testing.assert_array_equal(dist[q], od)
    pats = [X[q, 5:5 + L] + q for q, L in enumerate((10, 2, 20, 33, 7))]
    idx, dist = wd.paired_subsequence_match(pats, X, 2.0, metric=metric, metric_params=params,
                                            scale=True, return_distance=True)
    for q in range(X.shape[0]):
        oi, od = oracle.scaled_subsequence_matches(pats[q], X[q], 2.0, metric, p.get("r", 1.0),
                                                   p.get("g", 0.05), p.get("p", 1.0))
        if oi.size == 0:
            assert idx[q] is None
        else:
            np.testing.assert_array_equal(idx[q], oi)
            np.testing.assert_array_equal(dist[q], od)
