"""Host logic of subsequence_match / paired_subsequence_match /
argmin_subsequence_distance without a GPU: the window profiles (the one
device step) are replaced by the oracle's, so the threshold, exclusion,
max_matches and return-shape rules of distance/_distance.py:732-1079 run as
written.  The device profiles themselves are compared with the oracle in
tests/test_gpu_match.py."""

import math

import numpy as np
import pytest

import oracle
from wildboar_amd import distance as wd


def _oracle_profiles(patterns, x_, dim, metric, paired, threshold=np.inf):
    n, _, t = x_.shape
    nwp = t - min(p.shape[0] for p in patterns) + 1
    prof = np.full((n, nwp), np.inf)
    wts = oracle.wdtw_weights(t, metric.g) if metric.name == "wdtw" else None
    for i in range(n):
        s = patterns[i if paired else 0]
        L = s.shape[0]
        # euclidean is DTW on the diagonal band (r = 0), bit for bit
        band = oracle.compute_r(L, 0.0 if metric.name == "euclidean" else metric.r)
        for w in range(t - L + 1):
            prof[i, w] = oracle.dtw_distance(s, x_[i, dim, w:w + L], band, weights=wts)
    return prof


@pytest.fixture
def cpu_profiles(monkeypatch):
    monkeypatch.setattr(wd, "_profiles", _oracle_profiles)


def test_known_answers(cpu_profiles):
    # tests/wildboar/distance/test_distance.py:1047-1060
    s = np.array([1, 2, 3, 4])
    x = np.array([1, 2, 3, 3, 3, 3, 1, 2, 3, 4])
    np.testing.assert_array_equal(wd.subsequence_match(s, x, threshold=2), [0, 1, 6])
    x2 = np.array([[1, 2, 3, 3, 3, 3, 1, 2, 3, 4], [10, 2, 3, 3, 3, 3, 1, 2, 3, 4]])
    ind = wd.subsequence_match(s, x2, threshold=2)
    np.testing.assert_array_equal(ind[0], [0, 1, 6])
    np.testing.assert_array_equal(ind[1], [1, 6])
    ind, dist = wd.subsequence_match(s, x2, threshold=2, return_distance=True)
    assert dist[0].tolist() == [1.0, math.sqrt(3.0), 0.0]


def test_default_is_top10_by_distance(cpu_profiles):
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.standard_normal((3, 60)), axis=1)
    s = x[1, 10:22].copy()
    ind, dist = wd.subsequence_match(s, x, metric="dtw", metric_params={"r": 0.2},
                                     return_distance=True)
    prof = _oracle_profiles([s], x[:, None, :], 0, wd.check_metric("dtw")(r=0.2), False)
    for i in range(3):
        order = np.argsort(np.sqrt(prof[i]))[:10]
        assert ind[i].tolist() == order.tolist()
        assert dist[i].tolist() == np.sqrt(prof[i][order]).tolist()
    assert ind[1][0] == 10 and dist[1][0] == 0.0
    # a threshold keeps occurrence order; max_matches then orders by distance
    thr = float(np.sqrt(np.median(prof[0])))
    ind = wd.subsequence_match(s, x, thr, metric="dtw", metric_params={"r": 0.2})
    assert ind[0].tolist() == np.flatnonzero(prof[0] <= thr * thr).tolist()
    ind = wd.subsequence_match(s, x, thr, metric="dtw", metric_params={"r": 0.2},
                               max_matches=3)
    hit = np.flatnonzero(prof[0] <= thr * thr)
    assert ind[0].tolist() == hit[np.argsort(prof[0][hit])[:3]].tolist()


def test_exclude_keeps_best_of_each_neighbourhood(cpu_profiles):
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.standard_normal(80))
    s = x[30:40].copy()
    for exclude, zone in ((3, 3), (0.25, math.ceil(10 * 0.25)), (0, 0)):
        ind, dist = wd.subsequence_match(s, x, np.inf, exclude=exclude, return_distance=True)
        prof = np.sqrt(_oracle_profiles([s], x[None, None, :], 0, wd.check_metric("euclidean")(),
                                        False)[0])
        kept = []
        for w in np.argsort(prof):
            if not any(w - zone < e < w + zone for e in kept):
                kept.append(w)
        if zone == 0:
            kept = list(range(prof.size))
        assert sorted(ind.tolist()) == sorted(kept)
        assert all(i < j for i, j in zip(ind, ind[1:]))
        np.testing.assert_array_equal(dist, prof[ind])


def test_threshold_forms(cpu_profiles):
    y = np.array([1.0, 2.0])
    x = np.array([1.0, 2.0, 3.0, 4.0])
    # callable and "auto" (tests/wildboar/distance/test_distance.py:1664-1678)
    assert len(wd.subsequence_match(y, x, threshold=lambda d: d < 1.5)) > 0
    ind, dist = wd.subsequence_match(y, x, threshold="auto", return_distance=True)
    d_all = np.array([0.0, math.sqrt(2.0), math.sqrt(8.0)])
    cut = max(d_all.mean() - 2.0 * d_all.std(), d_all.min())
    assert ind.tolist() == [i for i in np.argsort(d_all)[:10] if d_all[i] <= cut]
    with pytest.raises(ValueError):
        wd.subsequence_match(y, x, threshold="not-a-rule")
    with pytest.raises(TypeError):
        wd.subsequence_match(y, x, threshold=object())
    # per-sample thresholds
    X = np.array([[1.0, 2.0, 3.0, 4.0], [0.0, 1.5, 2.5, 3.5]])
    ind = wd.subsequence_match(y, X, threshold=[0.5, 1.5])
    assert ind[0].tolist() == [0] and ind[1].tolist() == [0, 1]
    with pytest.raises(ValueError, match="threshold array length"):
        wd.subsequence_match(y, X, threshold=[0.5])
    # no match gives None; a lone sample returns its own array
    none = wd.subsequence_match(y, X, threshold=0.0)
    assert none[0].tolist() == [0] and none[1] is None
    one = wd.subsequence_match(y, X[:1], threshold=0.0)
    assert isinstance(one, np.ndarray) and one.tolist() == [0]
    assert wd.subsequence_match(np.array([9.0, 9.0]), x, threshold=0.1) is None


def test_paired_match(cpu_profiles):
    rng = np.random.default_rng(5)
    X = np.cumsum(rng.standard_normal((4, 30)), axis=1)
    ys = [X[0, 2:9].copy(), X[1, 0:3].copy(), X[2, 10:30].copy(), X[3, 5:6].copy()]
    ind, dist = wd.paired_subsequence_match(ys, X, np.inf, metric="wdtw",
                                            metric_params={"r": 0.5, "g": 0.1},
                                            return_distance=True)
    m = wd.check_metric("wdtw")(r=0.5, g=0.1)
    for i, s in enumerate(ys):
        prof = _oracle_profiles([s], X[i:i + 1, None, :], 0, m, False)[0]
        assert ind[i].tolist() == list(range(30 - len(s) + 1))
        assert dist[i].tolist() == np.sqrt(prof).tolist()
    top = wd.paired_subsequence_match(ys, X, metric="wdtw", metric_params={"r": 0.5, "g": 0.1})
    assert [t[0] for t in top] == [2, 0, 10, 5]
    one = wd.paired_subsequence_match([ys[0]], X[:1], 0.0)
    assert one.shape == (1,) and one[0].tolist() == [2]
    with pytest.raises(ValueError, match="same number of samples"):
        wd.paired_subsequence_match(ys[:2], X)
    with pytest.raises(ValueError, match="invalid subsequnce shape"):
        wd.paired_subsequence_match([np.zeros(31)] * 4, X)


def test_match_validation_matches_reference():
    # tests/wildboar/distance/test_distance.py:1643-1707; all raise before any
    # device work
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([1, 2, 3]), np.array([1, 2]))
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([]), np.array([1, 2, 3]))
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([1, 2, 3]), np.array([]))
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([1, 2]), np.array([[1, 2, 3, 4]]), dim=2)
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([1, 2]), np.array([1, 2, 3, 4]), metric="invalid_metric")
    with pytest.raises(ValueError, match="single subsequence"):
        wd.subsequence_match([[1, 2], [3, 4]], np.array([1, 2, 3, 4]))
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([1, 2]), np.array([1, 2, 3, 4]), exclude=-1)
    with pytest.raises(TypeError):
        wd.subsequence_match(np.array([1, 2]), np.array([1, 2, 3, 4]), exclude="2")
    with pytest.raises(NotImplementedError):
        wd.subsequence_match(np.array([1, 2]), np.array([1, 2, 3, 4]), metric="scaled_lcss")
    with pytest.raises(ValueError):
        wd.subsequence_match(np.array([1, 2]), np.array([1, 2, 3, 4]), metric="scaled_nope")
    with pytest.warns(UserWarning, match="n_jobs"):
        with pytest.raises(TypeError):
            wd.subsequence_match(np.array([1, 2]), np.array([1, 2, 3, 4]), threshold=object(),
                                 n_jobs=2)


def test_argmin_subsequence_validation():
    X = np.zeros((3, 10))
    S = np.zeros((3, 4))
    with pytest.raises(ValueError, match="k must be less"):
        wd.argmin_subsequence_distance(S, X, k=8)
    with pytest.raises(ValueError, match="same number of samples"):
        wd.argmin_subsequence_distance(S[:2], X)
    with pytest.raises(ValueError, match="longest subsequence"):
        wd.argmin_subsequence_distance([np.zeros(11)] * 3, X)
    with pytest.raises(ValueError, match="1d-array"):
        wd.argmin_subsequence_distance([np.zeros((2, 2))] * 3, X)
    with pytest.raises(ValueError, match="dim"):
        wd.argmin_subsequence_distance(S, X, dim=1)
    with pytest.raises(ValueError):
        wd.argmin_subsequence_distance(S, X, k=0)
    with pytest.raises(ValueError):
        wd.argmin_subsequence_distance(S, X, metric="lcss_not_here")
    with pytest.raises(ValueError):
        wd.argmin_subsequence_distance(S, X, metric="scaled_lcss_not_here")
