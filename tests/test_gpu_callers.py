"""GPU parity of the callers of the distance path: dtw_mapping and the
majorize-minimize dtw_average (float64 paths on the device), and the
NearestNeighbors / KNeighborsClassifier / KMeans estimators, against the
oracle and restatements of the reference's estimator loops over it.  Labels,
indices, paths, barycenters and costs bit for bit."""

import math

import numpy as np
import pytest

import oracle
from wildboar_amd import distance as wd
from wildboar_amd.dtw import dtw_average, dtw_mapping
from wildboar_amd.neighbors import KMeans, KNeighborsClassifier, NearestNeighbors

pytestmark = pytest.mark.gpu


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


@pytest.mark.parametrize("lx,ly,r", [(30, 30, 0.1), (30, 30, 1.0), (25, 40, 0.3), (40, 25, 0.3),
                                      (1, 7, 1.0), (9, 1, 0.5), (64, 64, 0.05)])
def test_dtw_mapping_matches_the_oracle_backtrack(lx, ly, r):
    rng = np.random.default_rng(lx * 100 + ly)
    for x, y in ((rng.standard_normal(lx), rng.standard_normal(ly)),
                 # integer values: many equal costs exercise the first-wins rule
                 (rng.integers(0, 3, lx).astype(float), rng.integers(0, 3, ly).astype(float))):
        ind, (pi, pj) = dtw_mapping(x, y, r=r, return_index=True)
        A = oracle.cost_matrix(x, y, oracle.warp_size(lx, r, ly))
        oi, oj = oracle.mapping(A)
        np.testing.assert_array_equal(pi, oi)
        np.testing.assert_array_equal(pj, oj)
        assert ind.shape == (lx, ly) and ind.sum() == len(oi)


@pytest.mark.parametrize("r", [0.1, 0.5, 1.0])
def test_dtw_average_mm_bitwise(r):
    X = rw(90, 12, 60)
    init = X[3].copy()
    mean, cost = dtw_average(X, r=r, init=init, return_cost=True)
    om, oc = oracle.mm_dtw_average(X, init, r=r)
    np.testing.assert_array_equal(mean, om)
    assert cost == oc
    w = np.random.default_rng(91).uniform(0.5, 2.0, 12)
    mean, cost = dtw_average(X, r=r, init=init, sample_weight=w, return_cost=True,
                             max_epoch=5)
    om, oc = oracle.mm_dtw_average(X, init, r=r, sample_weight=w, max_epoch=5)
    np.testing.assert_array_equal(mean, om)
    assert cost == oc
    # init="random" draws one series with the given RandomState
    a = dtw_average(X, r=r, random_state=7, max_epoch=2)
    b, _ = oracle.mm_dtw_average(X, X[np.random.RandomState(7).randint(12)], r=r, max_epoch=2)
    np.testing.assert_array_equal(a, b)
    # WDTW barycenters: wdtw_alignment paths, mean WDTW cost
    mean, cost = dtw_average(X, r=r, g=0.1, init=init, return_cost=True, max_epoch=6)
    om, oc = oracle.mm_dtw_average(X, init, r=r, g=0.1, max_epoch=6)
    np.testing.assert_array_equal(mean, om)
    assert cost == oc
    with pytest.raises(ValueError):
        dtw_average(X[:1])


def test_knn_and_nearest_neighbors_match_the_oracle():
    X = rw(92, 80, 64)
    y = np.random.default_rng(93).integers(0, 4, 80)
    Q = rw(94, 30, 64)
    for metric, params in (("dtw", {"r": 0.1}), ("euclidean", None), ("wdtw", {"r": 0.2})):
        p = params or {}
        clf = KNeighborsClassifier(5, metric=metric, metric_params=params).fit(X, y)
        idx, _ = oracle.argmin(Q, X, k=5, metric=metric, r=p.get("r", 1.0), g=p.get("g", 0.05))
        classes, yi = np.unique(y, return_inverse=True)
        preds = yi[idx]
        proba = np.stack([np.sum(preds == c, axis=1) / 5 for c in range(len(classes))], axis=1)
        np.testing.assert_array_equal(clf.predict_proba(Q), proba)
        np.testing.assert_array_equal(clf.predict(Q), classes[np.argmax(proba, axis=1)])
        nn = NearestNeighbors(3, metric=metric, metric_params=params).fit(X)
        d, i = nn.kneighbors(Q)
        oi, od = oracle.argmin(Q, X, k=3, metric=metric, r=p.get("r", 1.0))
        order = np.argsort(od, axis=1, kind="stable")
        np.testing.assert_array_equal(i, np.take_along_axis(oi, order, axis=1))
        np.testing.assert_array_equal(d, np.take_along_axis(od, order, axis=1))
        # query = training set: each sample's own index is dropped
        d, i = nn.kneighbors()
        assert i.shape == (80, 3) and not (i == np.arange(80)[:, None]).any()


def test_multivariate_neighbours_match_the_restated_selection():
    # (n, 3, t) inputs: the dim="mean" matrix, partial selection (kth = k - 1
    # for kneighbors, kth = k for the class shares: _neighbors.py:115-135, 270-272)
    rng = np.random.default_rng(140)
    X = np.cumsum(rng.standard_normal((40, 3, 32)), axis=2)
    Q = np.cumsum(rng.standard_normal((9, 3, 32)), axis=2)
    y = rng.integers(0, 3, 40)
    D = np.mean([oracle.pairwise(Q[:, d], X[:, d], metric="dtw", r=0.2) for d in range(3)], axis=0)
    nn = NearestNeighbors(4, metric="dtw", metric_params={"r": 0.2}).fit(X)
    d, i = nn.kneighbors(Q)
    part = np.argpartition(D, 3, axis=1)[:, :4]
    pd = np.take_along_axis(D, part, axis=1)
    order = np.argsort(pd, axis=1)
    np.testing.assert_array_equal(i, np.take_along_axis(part, order, axis=1))
    np.testing.assert_array_equal(d, np.take_along_axis(pd, order, axis=1))
    clf = KNeighborsClassifier(4, metric="dtw", metric_params={"r": 0.2}).fit(X, y)
    votes = np.unique(y, return_inverse=True)[1][np.argpartition(D, 4, axis=1)[:, :4]]
    want = np.stack([np.sum(votes == c, axis=1) / 4 for c in range(3)], axis=1)
    np.testing.assert_array_equal(clf.predict_proba(Q), want)


def _kmeans_restated(x, k, metric, r, random_state, max_iter=300, tol=1e-3, g=None):
    """_neighbors.py:483-564 over the oracle (one init)."""
    rs = (random_state if isinstance(random_state, np.random.RandomState)
          else np.random.RandomState(random_state))
    if g is not None:
        metric = "wdtw"
    C = x[rs.choice(x.shape[0], size=k, replace=False)].copy()
    cost, reassign = -np.inf, True
    for it in range(max_iter):
        a, dist = oracle.argmin(x, C, k=1, metric=metric, r=r, g=g or 0.05)
        a = a.ravel()
        prev, cost = cost, oracle.paired(x, C[a], metric=metric, r=r,
                                         g=g or 0.05).sum() / x.shape[0]
        if math.isclose(cost, prev, rel_tol=tol):
            reassign = False
            break
        for c in range(k):
            grp = x[a == c]
            if grp.shape[0] == 0:
                far = dist.min(axis=1).argmax()
                a[far] = c
                C[c] = x[far]
            elif grp.shape[0] == 1:
                C[c] = grp
            elif metric == "euclidean":
                C[c] = grp.mean(axis=0)
            else:
                rs.randint(np.iinfo(np.int32).max)
                C[c] = oracle.mm_dtw_average(grp, C[c], r=r, g=g)[0]
    if reassign:
        a = oracle.argmin(x, C, k=1, metric=metric, r=r, g=g or 0.05)[0].ravel()
    return a, cost, C, it


@pytest.mark.parametrize("metric,g", [("euclidean", None), ("dtw", None), ("dtw", 0.2)])
def test_kmeans_matches_the_restated_loop(metric, g):
    X = np.concatenate([rw(95, 20, 48), rw(96, 20, 48) + 2.0, rw(97, 20, 48)[:, ::-1] - 1.0])
    km = KMeans(3, metric=metric, r=0.2, g=g, random_state=11, max_iter=10).fit(X)
    a, cost, C, it = _kmeans_restated(X, 3, metric, 0.2, 11, max_iter=10, g=g)
    np.testing.assert_array_equal(km.labels_, a)
    np.testing.assert_array_equal(km.cluster_centers_, C)
    assert km.inertia_ == cost and km.n_iter_ == it
    pm = "wdtw" if g is not None else metric
    np.testing.assert_array_equal(km.predict(X[:7]), oracle.pairwise(
        X[:7], C, metric=pm, r=0.2, g=g or 0.05).argmin(axis=1))


def test_kmeans_several_inits_keep_the_best_run():
    X = np.concatenate([rw(101, 15, 40), rw(102, 15, 40) + 2.5, rw(103, 15, 40) - 2.5])
    km = KMeans(3, metric="dtw", r=0.3, n_init=3, random_state=4, max_iter=8).fit(X)
    rs = np.random.RandomState(4)
    runs = [_kmeans_restated(X, 3, "dtw", 0.3, rs, max_iter=8) for _ in range(3)]
    best = min(range(3), key=lambda q: (runs[q][1], q))
    a, cost, C, it = runs[best]
    np.testing.assert_array_equal(km.labels_, a)
    np.testing.assert_array_equal(km.cluster_centers_, C)
    assert km.inertia_ == cost and km.n_iter_ == it


@pytest.mark.parametrize("r,g,weighted", [(0.2, None, False), (1.0, None, True), (0.3, 0.1, False)])
def test_ssg_average_bitwise(r, g, weighted):
    from wildboar_amd.dtw import dtw_average

    rng = np.random.default_rng(31)
    X = np.cumsum(rng.standard_normal((7, 40)), axis=1)
    sw = rng.uniform(0.5, 2.0, 7) if weighted else None
    init = X[3].copy()
    got, cost = dtw_average(X, r=r, g=g, sample_weight=sw, init=init, method="ssg",
                            max_epoch=6, max_stable=2, learning_rate=0.2, decay=0.85,
                            return_cost=True, random_state=5)
    want, wcost = oracle.ssg_dtw_average(X, init, r=r, max_epoch=6, max_stable=2,
                                         learning_rate=0.2, decay=0.85, sample_weight=sw, g=g,
                                         random_state=5)
    np.testing.assert_array_equal(got, want)
    assert cost == wcost
