"""GPU parity of the derivative and amercing metrics (ddtw, wddtw, adtw)
through pairwise_distance / paired_distance / argmin_distance against the
float64 oracle.  Bars: FP64 bit-identical; FP32 max relative error 1e-5 with
argmin indices and distances exact; FP32_ACC64 bit-identical on float32-
rounded inputs."""

import numpy as np
import pytest

import oracle
from wildboar_amd import _lib
from wildboar_amd import distance as wd

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-5
CASES = [("ddtw", {"r": 0.1}), ("ddtw", {"r": 1.0}), ("ddtw", {"r": 0.25}),
         ("wddtw", {"r": 0.1, "g": 0.05}), ("wddtw", {"r": 1.0, "g": 0.5}),
         ("adtw", {"r": 0.1, "p": 1.0}), ("adtw", {"r": 1.0, "p": 0.1}),
         ("adtw", {"r": 0.3, "p": 0.0}), ("adtw", {"r": 0.2, "p": float("inf")})]


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


def okw(params):
    return {"r": params.get("r", 1.0), "g": params.get("g", 0.05), "p": params.get("p", 1.0)}


def setup_module(module):
    _lib.set_policy(0)


@pytest.mark.parametrize("metric,params", CASES)
def test_pairwise_paired_fp64_bitwise(metric, params):
    X = rw(80, 30, 96)
    Y = rw(81, 25, 96)
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params)
    np.testing.assert_array_equal(got, oracle.pairwise(X, Y, metric=metric, **okw(params)))
    got = wd.pairwise_distance(X, metric=metric, metric_params=params)
    np.testing.assert_array_equal(got, oracle.pairwise(X, None, metric=metric, **okw(params)))
    got = wd.paired_distance(X[:25], Y, metric=metric, metric_params=params)
    np.testing.assert_array_equal(got, oracle.paired(X[:25], Y, metric=metric, **okw(params)))


@pytest.mark.parametrize("metric,params", CASES[:7])
def test_pairwise_fp64_plain_kernel_and_unequal_lengths(metric, params):
    X = rw(82, 9, 70)
    Y = rw(83, 8, 110)
    want = oracle.pairwise(X, Y, metric=metric, **okw(params))
    np.testing.assert_array_equal(wd.pairwise_distance(X, Y, metric=metric, metric_params=params),
                                  want)
    _lib.set_policy(1)
    try:
        got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params)
    finally:
        _lib.set_policy(0)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("metric,params", CASES)
@pytest.mark.parametrize("k", [1, 4])
def test_argmin_fp64_and_fp32_exact(metric, params, k):
    X = rw(84, 20, 64)
    Y = rw(85, 40, 64)
    oi, od = oracle.argmin(X, Y, k=k, metric=metric, **okw(params))
    for precision in ("fp64", "fp32"):
        gi, gd = wd.argmin_distance(X, Y, k=k, metric=metric, metric_params=params,
                                    return_distance=True, precision=precision)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)


@pytest.mark.parametrize("metric,params", CASES)
def test_fp32_error_and_acc64_bitwise(metric, params):
    X = rw(86, 24, 128)
    Y = rw(87, 20, 128)
    want = oracle.pairwise(X, Y, metric=metric, **okw(params))
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params, precision="fp32")
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-1)
    assert np.isfinite(got).all() and err.max() <= FP32_RTOL, float(err.max())
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params,
                               precision="fp32_acc64")
    np.testing.assert_array_equal(got, oracle.pairwise(f32(X), f32(Y), metric=metric,
                                                       **okw(params)))


def test_derivative_short_series_and_validation():
    X = rw(88, 4, 2)
    Y = rw(89, 3, 2)
    assert (wd.pairwise_distance(X, Y, metric="ddtw") == 0).all()
    assert (wd.paired_distance(X[:3], Y, metric="wddtw") == 0).all()
    idx, dist = wd.argmin_distance(X, Y, k=2, metric="ddtw", return_distance=True)
    assert (idx == 0).all() and (dist == 0).all()
    X = rw(90, 4, 30)
    with pytest.raises(NotImplementedError):
        wd.argmin_subsequence_distance([X[0, :5]] * 4, X, metric="adtw",
                                       metric_params={"p": -1.0})
    with pytest.raises(ValueError):
        wd.pairwise_distance(X, metric="adtw", metric_params={"p": float("-inf")})
    with pytest.raises(TypeError):
        wd.pairwise_distance(X, metric="adtw", metric_params={"g": 0.1})
    # WeightedDerivativeDtwMetric takes any float g
    got = wd.pairwise_distance(X, X[:2], metric="wddtw", metric_params={"g": -0.2})
    np.testing.assert_array_equal(got, oracle.pairwise(X, X[:2], metric="wddtw", g=-0.2))


@pytest.mark.parametrize("metric,params", [("ddtw", {"r": 0.1}), ("ddtw", {"r": 1.0}),
                                           ("wddtw", {"r": 0.2, "g": 0.1}),
                                           ("adtw", {"r": 0.1, "p": 0.5}),
                                           ("adtw", {"r": 1.0, "p": 2.0})])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_subsequence_derivative_and_amercing_bitwise(metric, params, precision):
    X = rw(84, 9, 120)
    rng = np.random.default_rng(85)
    pats = [X[2, 10:40].copy(), rng.standard_normal(7), X[5, 0:2].copy(), X[0, 3:120].copy(),
            X[8, 50:53].copy(), rng.standard_normal(64)]
    d, i = wd.pairwise_subsequence_distance(pats, X, metric=metric, metric_params=params,
                                            return_index=True, precision=precision)
    od, oi = oracle.subsequence(pats, X, metric=metric, **okw(params))
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)


@pytest.mark.parametrize("metric,params", [("ddtw", {"r": 0.1}), ("ddtw", {"r": 1.0}),
                                           ("wddtw", {"r": 0.2, "g": 0.1}),
                                           ("adtw", {"r": 0.1, "p": 0.5}),
                                           ("adtw", {"r": 1.0, "p": 0.0})])
def test_derivative_and_amercing_matching_and_argmin_bitwise(metric, params):
    X = rw(92, 6, 100)
    s = X[2, 20:50].copy()
    p = okw(params)
    for threshold in (1.0, 3.0, np.inf):
        idx, dist = wd.subsequence_match(s, X, threshold, metric=metric, metric_params=params,
                                         return_distance=True)
        for q in range(X.shape[0]):
            oi, od = oracle.subsequence_matches(s, X[q], threshold, metric, p["r"], p["g"],
                                                p["p"])
            if oi.size == 0:
                assert idx[q] is None
            else:
                np.testing.assert_array_equal(idx[q], oi)
                np.testing.assert_array_equal(dist[q], od)
    pats = [X[q, 3 + q:3 + q + L].copy() for q, L in enumerate((2, 9, 30, 17, 44, 6))]
    idx, dist = wd.paired_subsequence_match(pats, X, 2.0, metric=metric, metric_params=params,
                                            return_distance=True)
    for q in range(X.shape[0]):
        oi, od = oracle.subsequence_matches(pats[q], X[q], 2.0, metric, p["r"], p["g"], p["p"])
        if oi.size == 0:
            assert idx[q] is None
        else:
            np.testing.assert_array_equal(idx[q], oi)
            np.testing.assert_array_equal(dist[q], od)
    for k in (1, 3):
        gi, gd = wd.argmin_subsequence_distance(pats, X, k=k, metric=metric,
                                                metric_params=params, return_distance=True)
        oi, od = oracle.argmin_subsequence(pats, X, k=k, metric=metric, r=p["r"],
                                           g=p["p"] if metric == "adtw" else p["g"])
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)


@pytest.mark.parametrize("tx,ty,r,metric", [(300, 300, 0.1, "dtw"), (400, 400, 0.12, "dtw"),
                                            (512, 512, 0.1, "dtw"), (300, 340, 0.1, "dtw"),
                                            (512, 512, 0.1, "adtw"), (300, 300, 0.2, "ddtw")])
def test_fp64_wide_register_classes_bitwise(tx, ty, r, metric):
    # bands of 57..104 slots run the wide float64 register classes
    X = rw(93, 3, tx)
    Y = rw(94, 7, ty)
    params = {"r": r, "p": 0.5} if metric == "adtw" else {"r": r}
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params)
    np.testing.assert_array_equal(got, oracle.pairwise(X, Y, metric=metric, **okw(params)))
    if tx == ty:
        got = wd.paired_distance(X, Y[:3], metric=metric, metric_params=params)
        np.testing.assert_array_equal(got, oracle.paired(X, Y[:3], metric=metric, **okw(params)))


@pytest.mark.parametrize("metric,params", [("dtw", {"r": 1.0}), ("adtw", {"r": 1.0, "p": 0.5}),
                                           ("ddtw", {"r": 1.0})])
def test_fp64_wide_classes_in_subsequence_scans_and_profiles(metric, params):
    # the default r = 1 with patterns of length 29..52 gives 57..104 slots
    X = rw(95, 5, 160)
    pats = [X[q, 7:7 + L].copy() for q, L in enumerate((30, 41, 52, 36, 45))]
    d, i = wd.pairwise_subsequence_distance(pats, X, metric=metric, metric_params=params,
                                            return_index=True)
    od, oi = oracle.subsequence(pats, X, metric=metric, **okw(params))
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)
    p = okw(params)
    idx, dist = wd.subsequence_match(pats[1], X, 4.0, metric=metric, metric_params=params,
                                     return_distance=True)
    for q in range(X.shape[0]):
        qi, qd = oracle.subsequence_matches(pats[1], X[q], 4.0, metric, p["r"], p["g"], p["p"])
        if qi.size == 0:
            assert idx[q] is None
        else:
            np.testing.assert_array_equal(idx[q], qi)
            np.testing.assert_array_equal(dist[q], qd)
    gi, gd = wd.argmin_subsequence_distance(pats, X, k=3, metric=metric, metric_params=params,
                                            return_distance=True)
    oi, od = oracle.argmin_subsequence(pats, X, k=3, metric=metric, r=p["r"],
                                       g=p["p"] if metric == "adtw" else p["g"])
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


def test_fp32_long_series_accumulate_in_float64():
    # FP32 requests on series past 2048 steps run with float64 accumulation:
    # the result is the oracle's on the float32-rounded inputs, bit for bit,
    # and within 1e-5 of the float64 oracle.  With the promotion off, the
    # float32 recursion's drift at n = 4096 is stated (asserted below 1e-3,
    # reported in the assertion message).
    from wildboar_amd import _lib as lib

    X = rw(140, 3, 4096)
    Y = rw(141, 4, 4096)
    Xr = X.astype(np.float32).astype(np.float64)
    Yr = Y.astype(np.float32).astype(np.float64)
    for r in (0.1, 1.0):
        want = oracle.pairwise(X, Y, metric="dtw", r=r)
        want_r = oracle.pairwise(Xr, Yr, metric="dtw", r=r)
        got = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r}, precision="fp32")
        np.testing.assert_array_equal(got, want_r)
        assert np.max(np.abs(got - want) / want) <= 1e-5
        lib.set_acc64_length(0)
        try:
            raw = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r},
                                       precision="fp32")
        finally:
            lib.set_acc64_length(2048)
        drift = float(np.max(np.abs(raw - want) / want))
        assert drift <= 1e-3, f"float32 accumulation at n=4096, r={r}: max rel err {drift:.2e}"
        print(f"n=4096 r={r}: float32 accumulation max rel err {drift:.3e}, "
              f"float64 accumulation {np.max(np.abs(got - want) / want):.3e}")
    idx = wd.argmin_distance(X, Y, k=2, metric="dtw", metric_params={"r": 0.1}, precision="fp32")
    np.testing.assert_array_equal(idx, oracle.argmin(Xr, Yr, k=2, metric="dtw", r=0.1)[0])
