"""Pins of the oracle's z-normalised subsequence scans (ScaledEuclidean /
ScaledDtw subsequence metrics, distance/_metric.pyx:561-651,
distance/_elastic.pyx:158-482, 1928-2203) by brute force, closed forms and the
reference's LB_Kim term.  CPU only."""

import numpy as np
import pytest

import oracle


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.standard_normal((n, t)), axis=1)


def znorm_windows(x, L):
    w = np.lib.stride_tricks.sliding_window_view(x, L)
    sd = w.std(axis=1)
    sd[sd <= 0] = 1.0
    return (w - w.mean(axis=1, keepdims=True)) / sd[:, None]


def zpat(s):
    s = np.asarray(s, dtype=np.float64)
    sd = s.std()
    return (s - s.mean()) / (sd if sd > 1e-13 else 1.0)


def band_dtw(a, b, w):
    L = len(a)
    C = np.full((L + 1, L + 1), np.inf)
    C[0, 0] = 0.0
    for i in range(1, L + 1):
        for j in range(max(1, i - w), min(L, i + w) + 1):
            C[i, j] = (a[i - 1] - b[j - 1]) ** 2 + min(C[i - 1, j], C[i, j - 1], C[i - 1, j - 1])
    return C[L, L]


def test_window_stats_match_numpy_and_the_degenerate_rule():
    x = rw(0, 1, 200)[0]
    for L in (1, 3, 17, 200):
        mean, std = oracle.window_stats(x, L)
        w = np.lib.stride_tricks.sliding_window_view(x, L)
        np.testing.assert_allclose(mean, w.mean(axis=1), rtol=1e-10, atol=1e-10)
        if L > 1:
            np.testing.assert_allclose(std, w.std(axis=1), rtol=1e-7)
    mean, std = oracle.window_stats(np.zeros(30), 5)
    assert (mean == 0).all() and (std == 1.0).all()  # ex2 / L - mean^2 == 0 -> 1.0


def test_scaled_euclidean_is_the_znormalised_window_minimum():
    X = rw(1, 6, 160)
    rng = np.random.default_rng(2)
    pats = [rng.standard_normal(L) for L in (4, 9, 31, 80)] + [X[2, 40:70] * 3 - 1]
    d, i = oracle.scaled_subsequence(pats, X, "euclidean")
    for a in range(X.shape[0]):
        for b, s in enumerate(pats):
            full = ((znorm_windows(X[a], len(s)) - zpat(s)) ** 2).sum(axis=1)
            np.testing.assert_allclose(d[a, b] ** 2, full.min(), rtol=1e-9, atol=1e-12)
            assert abs(full[i[a, b]] - full.min()) <= 1e-9 * max(full.min(), 1.0)
    # an affine copy of a window is found at its own start with distance ~0
    assert i[2, 4] == 40 and d[2, 4] < 1e-6


def test_scaled_dtw_band_zero_equals_scaled_euclidean_bitwise():
    # warp width floor(L * 0) = 0 leaves the diagonal, whose sums run in the
    # euclidean order; from length 6 the first and last three points of
    # LB_Kim are distinct diagonal cells, so no window is pruned wrongly
    X = rw(3, 8, 120)
    rng = np.random.default_rng(4)
    pats = [rng.standard_normal(L) for L in (6, 7, 15, 60)] + [X[1, 5:50].copy()]
    d, i = oracle.scaled_subsequence(pats, X, "dtw", r=0.0)
    e, ei = oracle.scaled_subsequence(pats, X, "euclidean")
    np.testing.assert_array_equal(d, e)
    np.testing.assert_array_equal(i, ei)


@pytest.mark.parametrize("r", [0.1, 0.25, 1.0])
def test_scaled_dtw_exhaustive_is_brute_force_and_bounds_the_cascade(r):
    X = rw(5, 4, 70)
    rng = np.random.default_rng(6)
    pats = [rng.standard_normal(L) for L in (3, 8, 12)] + [X[3, 20:36] * 0.5]
    d, i = oracle.scaled_subsequence(pats, X, "dtw", r=r)
    e, ei = oracle.scaled_subsequence(pats, X, "dtw", r=r, cascade=False)
    for a in range(X.shape[0]):
        for b, s in enumerate(pats):
            L = len(s)
            w = oracle.warp_width(L, r)
            zs = zpat(s)
            costs = np.array([band_dtw(zs, z, w) for z in znorm_windows(X[a], L)])
            np.testing.assert_allclose(e[a, b] ** 2, costs.min(), rtol=1e-9, atol=1e-12)
    # the cascade only removes candidates: never below the exhaustive minimum
    assert (d >= e).all()
    assert oracle.warp_width(10, 1.0) == 9 and oracle.warp_width(10, 0.25) == 2
    assert oracle.warp_width(10, 0.05) == 0


def test_scaled_dtw_follows_the_reference_lb_kim_term():
    # The reference's LB_Kim (constant_lower_bound, _elastic.pyx:197-203)
    # takes min(d(T[L-2], S[L-2]), d(T[L-1], S[L-2]), d(T[L-2], S[L-2])),
    # without d(T[L-2], S[L-1]); where that exceeds the DTW of the best
    # window the reference skips the window and reports a larger minimum.
    rng = np.random.default_rng(1)
    X = np.cumsum(rng.standard_normal((40, 300)), axis=1)
    pats = []
    for _ in range(30):
        L = int(rng.integers(6, 60))
        s = int(rng.integers(0, 300 - L))
        pats.append(X[int(rng.integers(0, 40)), s:s + L] * rng.uniform(0.5, 2) + rng.normal())
    r = 0.3
    d, _ = oracle.scaled_subsequence(pats, X, "dtw", r=r)
    e, ei = oracle.scaled_subsequence(pats, X, "dtw", r=r, cascade=False)
    bad = np.argwhere(d != e)
    assert len(bad) > 0
    sq = lambda a, b: (a - b) ** 2  # noqa: E731
    for a, b in bad[:8]:
        s = np.asarray(pats[b])
        L = len(s)
        zs = zpat(s)
        mean, std = oracle.window_stats(X[a], L)
        I = ei[a, b]
        zx = (X[a, I:I + L] - mean[I]) / std[I]
        D = band_dtw(zs, zx, oracle.warp_width(L, r))
        head = sq(zx[0], zs[0]) + sq(zx[-1], zs[-1]) + min(
            sq(zx[1], zs[0]), sq(zx[0], zs[1]), sq(zx[1], zs[1]))
        k_ref = head + min(sq(zx[-2], zs[-2]), sq(zx[-1], zs[-2]))
        k_full = head + min(sq(zx[-2], zs[-2]), sq(zx[-1], zs[-2]), sq(zx[-2], zs[-1]))
        assert k_ref > D * (1 + 1e-9) and k_full <= D * (1 + 1e-9)
        assert d[a, b] > e[a, b]


def test_scaled_scans_are_affine_invariant_in_the_pattern():
    X = rw(7, 5, 140)
    rng = np.random.default_rng(8)
    pats = [rng.standard_normal(L) for L in (10, 25)]
    for metric, r in (("euclidean", 1.0), ("dtw", 0.1)):
        d, i = oracle.scaled_subsequence(pats, X, metric, r=r)
        d2, i2 = oracle.scaled_subsequence([p * 7.5 + 3.0 for p in pats], X, metric, r=r)
        np.testing.assert_allclose(d2, d, rtol=1e-7)
        np.testing.assert_array_equal(i2, i)


@pytest.mark.parametrize("metric,r", [("euclidean", 1.0), ("dtw", 0.1), ("dtw", 1.0)])
def test_scaled_matches_at_infinity_are_the_full_profile(metric, r):
    X = rw(9, 3, 150)
    s = X[1, 30:60] * 2.0 + 5.0
    d, i = oracle.scaled_subsequence([s], X, metric, r=r, cascade=False) if metric == "dtw" \
        else oracle.scaled_subsequence([s], X, metric)
    for q in range(X.shape[0]):
        idx, dist = oracle.scaled_subsequence_matches(s, X[q], np.inf, metric, r)
        np.testing.assert_array_equal(idx, np.arange(150 - 30 + 1))
        # every window recorded with its full distance: the minimum is the
        # exhaustive scan's, at its index
        assert dist.min() == d[q, 0] and np.argmin(dist) == i[q, 0]
        zs = zpat(s)
        w = oracle.warp_width(30, r)
        full = [np.sqrt(((z - zs) ** 2).sum()) if metric == "euclidean"
                else np.sqrt(band_dtw(zs, z, w)) for z in znorm_windows(X[q], 30)]
        np.testing.assert_allclose(dist, full, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("metric,r", [("euclidean", 1.0), ("dtw", 0.3)])
def test_scaled_matches_at_a_threshold_are_recorded_full_distances(metric, r):
    X = rw(10, 4, 200)
    s = X[2, 70:100].copy()
    for q in range(X.shape[0]):
        all_i, all_d = oracle.scaled_subsequence_matches(s, X[q], np.inf, metric, r)
        for th in (1.0, 2.5, 4.0):
            idx, dist = oracle.scaled_subsequence_matches(s, X[q], th, metric, r)
            assert (dist <= th).all()
            np.testing.assert_array_equal(dist, all_d[idx])
            if metric == "euclidean":
                # no bound: exactly the windows within the threshold
                np.testing.assert_array_equal(idx, all_i[all_d <= th])
            else:
                # the cascade only removes windows
                assert set(idx.tolist()) <= set(all_i[all_d <= th].tolist())


def test_scaled_wrap_is_the_first_smallest_znormalised_distance():
    # ScaledSubsequenceMetricWrap: windows normalised by incremental statistics
    # (population std), the wrapped metric's _eadistance on the normalised
    # buffers (derivative kinds: slopes with the band of length L - 2; weights
    # sized by the series, t and t - 2), the window of smallest distance
    X = rw(11, 3, 90)
    t = X.shape[1]
    rng = np.random.default_rng(12)
    pats = [X[1, 20:45] * 3.0 - 2.0, rng.standard_normal(9)]
    r = 0.3
    for metric, kw in (("wdtw", {"g": 0.1}), ("ddtw", {}), ("wddtw", {"g": 0.2}),
                       ("adtw", {"p": 0.3})):
        d, i = oracle.scaled_subsequence(pats, X, metric, r=r, **kw)
        for a in range(X.shape[0]):
            for b, s in enumerate(pats):
                L = len(s)
                zs = zpat(s)
                costs = []
                for z in znorm_windows(X[a], L):
                    if metric == "wdtw":
                        c = oracle.dtw_distance(zs, z, oracle.compute_r(L, r),
                                                weights=oracle.wdtw_weights(t, kw["g"]))
                    elif metric == "adtw":
                        c = oracle.adtw_distance(zs, z, oracle.compute_r(L, r), kw["p"])
                    else:
                        wts = oracle.wdtw_weights(t - 2, kw["g"]) if metric == "wddtw" else None
                        c = oracle.dtw_distance(oracle.average_slope(zs), oracle.average_slope(z),
                                                oracle.compute_r(L - 2, r), weights=wts)
                    costs.append(np.sqrt(c))
                costs = np.array(costs)
                np.testing.assert_allclose(d[a, b], costs.min(), rtol=1e-8, atol=1e-12)
                assert costs[i[a, b]] <= costs.min() * (1 + 1e-8) + 1e-12
        # the pattern's own (affinely changed) window is found, distance ~0
        assert i[1, 0] == 20 and d[1, 0] < 1e-6


@pytest.mark.parametrize("metric", ["euclidean", "dtw", "ddtw"])
def test_scaled_argmin_k1_matches_the_znormalised_minimum(metric):
    # argmin with k = 1 keeps the first window of smallest distance over the
    # z-normalised windows: the pattern's own affine copy at distance ~0, and
    # the distance of the best window equals brute force on numpy-normalised
    # windows
    X = rw(13, 4, 80)
    pats = [X[q, 10:10 + L] * 1.5 - 0.5 for q, L in enumerate((9, 16, 24, 31))]
    i, d = oracle.scaled_argmin_subsequence(pats, X, k=1, metric=metric, r=0.25)
    for q, s in enumerate(pats):
        assert i[q, 0] == 10 and d[q, 0] < 1e-6
    i, d = oracle.scaled_argmin_subsequence([np.sin(np.arange(12.0))] * 4, X, k=1,
                                            metric=metric, r=0.25)
    for q in range(4):
        zs = zpat(np.sin(np.arange(12.0)))
        costs = []
        for z in znorm_windows(X[q], 12):
            if metric == "euclidean":
                costs.append(np.sqrt(((z - zs) ** 2).sum()))
            elif metric == "dtw":
                costs.append(np.sqrt(oracle.dtw_distance(zs, z, oracle.compute_r(12, 0.25))))
            else:
                costs.append(np.sqrt(oracle.dtw_distance(oracle.average_slope(zs),
                                                         oracle.average_slope(z),
                                                         oracle.compute_r(10, 0.25))))
        np.testing.assert_allclose(d[q, 0], min(costs), rtol=1e-8)


@pytest.mark.parametrize("metric,kw", [("wdtw", {"g": 0.1}), ("ddtw", {}), ("adtw", {"p": 0.4})])
def test_scaled_wrap_matches_are_the_strictly_closer_windows(metric, kw):
    # the wrap records a window when its distance is below the threshold
    # (strict); at infinity every window, with the distances whose minimum
    # the wrap's scan reports
    X = rw(14, 3, 70)
    s = X[0, 10:30] * 2.0
    d, i = oracle.scaled_subsequence([s], X, metric, r=0.3, **kw)
    for q in range(X.shape[0]):
        all_i, all_d = oracle.scaled_subsequence_matches(s, X[q], np.inf, metric, 0.3, **kw)
        np.testing.assert_array_equal(all_i, np.arange(70 - 20 + 1))
        assert all_d.min() == d[q, 0] and np.argmin(all_d) == i[q, 0]
        th = float(np.median(all_d))
        idx, dist = oracle.scaled_subsequence_matches(s, X[q], th, metric, 0.3, **kw)
        np.testing.assert_array_equal(idx, all_i[all_d < th])
        np.testing.assert_array_equal(dist, all_d[all_d < th])

