"""GPU parity of the window-profile entry points (wbx_subsequence_profile,
wbx_argmin_subsequence and their device forms) and of the functions built on
them: subsequence_match, paired_subsequence_match and
argmin_subsequence_distance.  Float64 throughout: profiles, match indices,
distances and heap order bit-identical to the oracle."""

import numpy as np
import pytest

import oracle
from wildboar_amd import _lib
from wildboar_amd import distance as wd

pytestmark = pytest.mark.gpu

METRICS = [("euclidean", {}), ("dtw", {"r": 0.1}), ("dtw", {"r": 1.0}),
           ("wdtw", {"r": 0.3, "g": 0.2})]


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


def oracle_costs(s, x, metric, params):
    """Squared cost of every window of x against s, by the oracle's DTW
    (euclidean = the diagonal band, bit for bit)."""
    L = len(s)
    r = 0.0 if metric == "euclidean" else params.get("r", 1.0)
    band = oracle.compute_r(L, r)
    wts = oracle.wdtw_weights(len(x), params.get("g", 0.05)) if metric == "wdtw" else None
    return np.array([oracle.dtw_distance(s, x[w:w + L], band, weights=wts)
                     for w in range(len(x) - L + 1)])


def host_profile(patterns, X, paired, metric, params, dim=0):
    S, offs, lens = oracle.pack_shapelets(patterns)
    X3 = X if X.ndim == 3 else X[:, None, :]
    X3 = np.ascontiguousarray(X3, dtype=np.float64)
    n, dims, t = X3.shape
    nwp = t - int(lens.min()) + 1
    prof = np.empty((n, nwp))
    m = wd.check_metric(metric)(**params).struct(_lib.FP64)
    _lib.check(_lib.load().wbx_subsequence_profile(
        m, S.ctypes.data, offs.ctypes.data, lens.ctypes.data, len(patterns), 1 if paired else 0,
        X3.ctypes.data, n, dims, t, dim, nwp, prof.ctypes.data))
    return prof


@pytest.mark.parametrize("metric,params", METRICS)
def test_profile_bitwise(metric, params):
    X = rw(60, 6, 150)
    rng = np.random.default_rng(61)
    pats = [X[2, 10:41].copy(), rng.standard_normal(1), X[0, :150].copy(), X[5, 70:77].copy(),
            rng.standard_normal(64), X[1, 3:100].copy()]
    prof = host_profile(pats, X, True, metric, params)
    assert prof.shape == (6, 150)
    for q, s in enumerate(pats):
        want = oracle_costs(s, X[q], metric, params)
        np.testing.assert_array_equal(prof[q, :want.size], want)
        assert np.isinf(prof[q, want.size:]).all()
    # one pattern against every series
    prof = host_profile([pats[0]], X, False, metric, params)
    for q in range(6):
        np.testing.assert_array_equal(prof[q], oracle_costs(pats[0], X[q], metric, params))


@pytest.mark.parametrize("metric,params", METRICS)
def test_subsequence_match_bitwise(metric, params):
    X = rw(62, 9, 300)
    s = X[4, 100:140].copy()
    r, g = params.get("r", 1.0), params.get("g", 0.05)
    full = [oracle.subsequence_matches(s, X[q], np.inf, metric, r=r, g=g) for q in range(9)]
    thr = float(np.median(np.concatenate([d for _, d in full])))
    ind, dist = wd.subsequence_match(s, X, thr, metric=metric, metric_params=params,
                                     return_distance=True)
    for q in range(9):
        oi, od = oracle.subsequence_matches(s, X[q], thr, metric, r=r, g=g)
        if oi.size == 0:
            assert ind[q] is None and dist[q] is None
        else:
            np.testing.assert_array_equal(ind[q], oi)
            np.testing.assert_array_equal(dist[q], od)
    # no threshold: the 10 closest windows by distance
    ind, dist = wd.subsequence_match(s, X, metric=metric, metric_params=params,
                                     return_distance=True)
    for q in range(9):
        oi, od = full[q]
        order = np.argsort(od)[:10]
        np.testing.assert_array_equal(ind[q], oi[order])
        np.testing.assert_array_equal(dist[q], od[order])
    assert ind[4][0] == 100 and dist[4][0] == 0.0


def test_subsequence_match_known_answer_and_exclude():
    # tests/wildboar/distance/test_distance.py:1047-1060
    s = np.array([1, 2, 3, 4])
    x = np.array([[1, 2, 3, 3, 3, 3, 1, 2, 3, 4], [10, 2, 3, 3, 3, 3, 1, 2, 3, 4]])
    ind = wd.subsequence_match(s, x, threshold=2)
    np.testing.assert_array_equal(ind[0], [0, 1, 6])
    np.testing.assert_array_equal(ind[1], [1, 6])
    np.testing.assert_array_equal(wd.subsequence_match(s, x[0], threshold=2), [0, 1, 6])
    # euclidean and dtw at r = 0 agree, also after exclusion and max_matches
    # (tests/wildboar/distance/test_distance.py:1441-1475)
    X = rw(63, 9, 150)
    sub = X[8, 40:60].copy()
    a = wd.subsequence_match(sub, X, return_distance=True, max_matches=20, exclude=4)
    b = wd.subsequence_match(sub, X, return_distance=True, max_matches=20, exclude=4,
                             metric="dtw", metric_params={"r": 0.0})
    for p, q in zip(a[0], b[0]):
        np.testing.assert_array_equal(p, q)
    for p, q in zip(a[1], b[1]):
        np.testing.assert_array_equal(p, q)


@pytest.mark.parametrize("metric,params", METRICS)
def test_paired_subsequence_match_bitwise(metric, params):
    X = rw(64, 5, 120)
    pats = [X[0, 5:25].copy(), X[1, 0:1].copy(), X[2, :120].copy(), X[3, 60:99].copy(),
            X[4, 30:33].copy()]
    r, g = params.get("r", 1.0), params.get("g", 0.05)
    ind, dist = wd.paired_subsequence_match(pats, X, np.inf, metric=metric, metric_params=params,
                                            return_distance=True)
    for q, s in enumerate(pats):
        oi, od = oracle.subsequence_matches(s, X[q], np.inf, metric, r=r, g=g)
        np.testing.assert_array_equal(ind[q], oi)
        np.testing.assert_array_equal(dist[q], od)


@pytest.mark.parametrize("metric,params", METRICS)
@pytest.mark.parametrize("k", [1, 3, 7])
def test_argmin_subsequence_bitwise(metric, params, k):
    # the reference's own layout (tests/wildboar/distance/test_distance.py:1568-1601):
    # every length-10 window of one series against that series
    X = rw(65, 1, 200)[0]
    S = np.lib.stride_tricks.sliding_window_view(X, window_shape=10)
    XX = np.broadcast_to(X, shape=(S.shape[0], X.shape[0]))
    r, g = params.get("r", 1.0), params.get("g", 0.05)
    idx, dist = wd.argmin_subsequence_distance(S, XX, k=k, metric=metric, metric_params=params,
                                               return_distance=True)
    oi, od = oracle.argmin_subsequence(list(S), XX, k=k, metric=metric, r=r, g=g)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, od)
    assert (idx[:, 0] == np.arange(S.shape[0])).all() or k > 1


def test_argmin_subsequence_ragged_multivariate():
    rng = np.random.default_rng(66)
    X = rng.standard_normal((7, 3, 90)).cumsum(axis=2)
    pats = [rng.standard_normal(int(L)) for L in (1, 5, 90, 17, 33, 2, 60)]
    for metric, params in METRICS:
        r, g = params.get("r", 1.0), params.get("g", 0.05)
        idx, dist = wd.argmin_subsequence_distance(pats, X, dim=2, k=1, metric=metric,
                                                   metric_params=params, return_distance=True)
        oi, od = oracle.argmin_subsequence(pats, X, k=1, metric=metric, r=r, g=g, dim=2)
        np.testing.assert_array_equal(idx, oi)
        np.testing.assert_array_equal(dist, od)
    with pytest.raises(ValueError, match="k must be less"):
        wd.argmin_subsequence_distance(pats, X, k=2)


def test_device_forms_match_host():
    import torch
    from wildboar_amd import device

    X = rw(67, 8, 256)
    pats = [X[q, s:s + L].copy() for q, (s, L) in enumerate(
        [(0, 16), (3, 40), (100, 9), (200, 56), (7, 1), (30, 128), (1, 255), (64, 64)])]
    for dt in (torch.float64, torch.float32):
        Xd = torch.from_numpy(X).to(dt).cuda()
        sd = [torch.from_numpy(p).to(dt) for p in pats]
        Xh = Xd.double().cpu().numpy()
        ph = [s.double().numpy() for s in sd]
        for metric, params in METRICS:
            r, g = params.get("r", 1.0), params.get("g", 0.05)
            prof = device.subsequence_profile(sd, Xd, paired=True, metric=metric, r=r, g=g)
            np.testing.assert_array_equal(prof.cpu().numpy(), host_profile(ph, Xh, True, metric,
                                                                           params))
            idx, dist = device.argmin_subsequence(sd, Xd, k=2, metric=metric, r=r, g=g)
            oi, od = oracle.argmin_subsequence(ph, Xh, k=2, metric=metric, r=r, g=g)
            np.testing.assert_array_equal(idx.cpu().numpy(), oi)
            np.testing.assert_array_equal(dist.cpu().numpy(), od)


def test_profile_entry_validation():
    X = np.zeros((2, 1, 10))
    S, offs, lens = oracle.pack_shapelets([np.zeros(4), np.zeros(11)])
    prof = np.empty((2, 7))
    m = wd.check_metric("dtw")(r=0.5).struct(_lib.FP64)
    lib = _lib.load()
    rc = lib.wbx_subsequence_profile(m, S.ctypes.data, offs.ctypes.data, lens.ctypes.data, 2, 1,
                                     X.ctypes.data, 2, 1, 10, 0, 7, prof.ctypes.data)
    assert rc == _lib.WBX_EINVAL
    S, offs, lens = oracle.pack_shapelets([np.zeros(4), np.zeros(3)])
    rc = lib.wbx_subsequence_profile(m, S.ctypes.data, offs.ctypes.data, lens.ctypes.data, 2, 1,
                                     X.ctypes.data, 2, 1, 10, 0, 7, prof.ctypes.data)
    assert rc == _lib.WBX_EINVAL  # width must be t - min_len + 1 = 8
    idx = np.empty((2, 9), dtype=np.int64)
    dist = np.empty((2, 9))
    rc = lib.wbx_argmin_subsequence(m, S.ctypes.data, offs.ctypes.data, lens.ctypes.data, 2,
                                    X.ctypes.data, 1, 10, 0, 9, idx.ctypes.data, dist.ctypes.data)
    assert rc == _lib.WBX_EINVAL  # k > t - max_len + 1 = 7


def test_float64_lane_classes_pairs_profiles_and_scans():
    """Band widths that land in every float64 register class (B = 1 .. 55)
    and past it (block kernels), weighted and not, equal and unequal
    lengths; everything bit-identical to the oracle."""
    X = rw(70, 6, 300)
    Y = rw(71, 5, 300)
    for t, r in ((200, 0.1), (240, 0.1), (256, 0.1), (280, 0.1), (150, 0.1), (90, 0.05),
                 (300, 0.16)):
        for metric in ("dtw", "wdtw"):
            got = wd.pairwise_distance(X[:, :t], Y[:, :t], metric=metric, metric_params={"r": r})
            want = oracle.pairwise(X[:, :t], Y[:, :t], metric=metric, r=r)
            np.testing.assert_array_equal(got, want)
    # unequal lengths: B = |lx - ly| + 2 w - 1 = 41 (class 48) and 23
    for lx, ly in ((60, 90), (90, 60), (33, 45)):
        got = wd.pairwise_distance(X[:, :lx], Y[:, :ly], metric="dtw", metric_params={"r": 0.1})
        np.testing.assert_array_equal(got, oracle.pairwise(X[:, :lx], Y[:, :ly], metric="dtw",
                                                           r=0.1))
    pats = [X[0, 5:205].copy(), X[1, 0:240].copy(), X[2, 20:270].copy(), X[3, 1:2].copy(),
            X[4, 50:90].copy(), X[5, 0:300].copy()]
    for metric, params in (("dtw", {"r": 0.1}), ("wdtw", {"r": 0.1, "g": 0.05})):
        prof = host_profile(pats, X, True, metric, params)
        for q, s in enumerate(pats):
            want = oracle_costs(s, X[q], metric, params)
            np.testing.assert_array_equal(prof[q, :want.size], want)
            assert np.isinf(prof[q, want.size:]).all()
        d, i = wd.pairwise_subsequence_distance(pats, X, metric=metric, metric_params=params,
                                                return_index=True)
        od, oi = oracle.subsequence(pats, X, metric=metric, r=params["r"],
                                    g=params.get("g", 0.05))
        np.testing.assert_array_equal(i, oi)
        np.testing.assert_array_equal(d, od)


def test_wdtw_zero_weights_take_the_block_kernels():
    # g * t / 2 > 709: exp overflows, the first weights are exactly 0 and a
    # column outside the series would give inf * 0; those weight vectors are
    # served by the block kernels, still bit-identical
    X = rw(72, 4, 100)
    Y = rw(73, 3, 100)
    got = wd.pairwise_distance(X, Y, metric="wdtw", metric_params={"r": 0.1, "g": 20.0})
    np.testing.assert_array_equal(got, oracle.pairwise(X, Y, metric="wdtw", r=0.1, g=20.0))
    assert oracle.wdtw_weights(100, 20.0)[0] == 0.0
    s = X[1, 10:40].copy()
    prof = host_profile([s], X, False, "wdtw", {"r": 0.1, "g": 20.0})
    for q in range(4):
        np.testing.assert_array_equal(prof[q], oracle_costs(s, X[q], "wdtw", {"r": 0.1,
                                                                             "g": 20.0}))
