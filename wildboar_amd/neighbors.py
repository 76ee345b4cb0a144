"""Nearest-neighbour search and k-means on the GPU distance path.

Public estimators with the reference's parameters, fitted attributes and
results (``distance/_neighbors.py``):

- ``NearestNeighbors``      (``_neighbors.py:18-169``)
- ``KNeighborsClassifier``  (``_neighbors.py:172-305``)
- ``KMeans``                (``_neighbors.py:308-612``)

Layout of this module: one neighbour search (``_nearest``) serves both
neighbour estimators; k-means is one Lloyd loop (``_lloyd``) parameterised
by a centre rule (``_mean_rule`` for euclidean, ``_barycenter_rule`` for the
DTW family).  Every distance comes from ``wildboar_amd.distance`` in float64
(bit-identical to the reference's loops); the elastic centres are the
majorize-minimize barycenters of ``wildboar_amd.dtw.dtw_average``.
"""

import math
import numbers

import numpy as np
from sklearn.base import BaseEstimator, ClassifierMixin, ClusterMixin, TransformerMixin
from sklearn.utils import check_random_state
from sklearn.utils.multiclass import check_classification_targets
from sklearn.utils.validation import check_is_fitted

from . import distance as wd

__all__ = ["NearestNeighbors", "KNeighborsClassifier", "KMeans"]

# seeds handed to each barycenter are drawn below this bound (int32 max)
_SEED_BOUND = 2**31 - 1


# ------------------------------------------------------------- validation
def _require(ok, name, what, value):
    if not ok:
        raise ValueError(f"The '{name}' parameter must be {what}, got {value!r}.")


def _is_int(v):
    return isinstance(v, numbers.Integral) and not isinstance(v, bool)


def _is_real(v):
    return isinstance(v, numbers.Real) and not isinstance(v, bool)


def _series(x, multivariate):
    """Validated input; a multivariate array of one dimension is univariate
    (so the k-best search runs instead of the full matrix)."""
    x = wd.check_array(x, allow_3d=multivariate, ensure_2d=True)
    if x.ndim == 3 and x.shape[1] == 1:
        x = x.reshape(x.shape[0], -1)
    return x


# ------------------------------------------------------------- neighbours
def _nearest(query, index, k, metric, params, skip_self=False):
    """Distances and indices of the ``k`` nearest ``index`` rows of each
    ``query`` row, closest first.

    Univariate queries use ``argmin_distance`` (the reference's k-best Heap,
    stable-sorted); multivariate ones the full ``dim="mean"`` matrix with a
    partial selection, as ``_neighbors.py:115-135`` does.  ``skip_self``: the
    query is the index itself and a row is not its own neighbour.
    """
    n = query.shape[0]
    if query.ndim == 3:
        D = wd.pairwise_distance(query, index, dim="mean", metric=metric, metric_params=params)
        if skip_self:
            np.fill_diagonal(D, np.inf)
        cand = np.argpartition(D, k - 1, axis=1)[:, :k]
        cand_d = np.take_along_axis(D, cand, axis=1)
        order = np.argsort(cand_d, axis=1)
        return np.take_along_axis(cand_d, order, axis=1), np.take_along_axis(cand, order, axis=1)
    want = k + 1 if skip_self else k
    idx, dist = wd.argmin_distance(query, index, k=want, metric=metric, metric_params=params,
                                   return_distance=True, sorted=True)
    if skip_self:
        drop = idx == np.arange(n)[:, None]
        # Deviation: where exact copies rank before a row's own index, the
        # reference's reshape fails; here the row drops its farthest entry.
        drop[~drop.any(axis=1), -1] = True
        keep = ~drop
        idx = idx[keep].reshape(n, idx.shape[1] - 1)
        dist = dist[keep].reshape(n, dist.shape[1] - 1)
    return dist, idx


class _NeighborParams(BaseEstimator):
    """Parameters and training storage shared by the neighbour estimators."""

    def __init__(self, n_neighbors=5, *, metric="euclidean", metric_params=None, n_jobs=None):
        self.n_neighbors = n_neighbors
        self.metric = metric
        self.metric_params = metric_params
        self.n_jobs = n_jobs

    def _check_params(self):
        _require(_is_int(self.n_neighbors) and self.n_neighbors >= 1, "n_neighbors",
                 "an int in the range [1, inf)", self.n_neighbors)
        _require(isinstance(self.metric, str) and self.metric in wd.METRICS, "metric",
                 f"a str among {sorted(wd.METRICS)}", self.metric)
        _require(self.metric_params is None or isinstance(self.metric_params, dict),
                 "metric_params", "a dict or None", self.metric_params)

    def _remember(self, x):
        self._check_params()
        self._fit_X = wd.check_array(x, allow_3d=True, ensure_2d=True).copy()


class NearestNeighbors(TransformerMixin, _NeighborParams):
    """Unsupervised neighbour search (``_neighbors.py:18-169``)."""

    def fit(self, x, y=None):
        self._remember(x)
        return self

    def kneighbors(self, x=None, n_neighbors=None, return_distance=True):
        """The ``n_neighbors`` closest training series of each query, closest
        first; without ``x`` the training series query themselves and skip
        their own index."""
        check_is_fitted(self, "_fit_X")
        k = self.n_neighbors if n_neighbors is None else n_neighbors
        own = x is None
        q = _series(self._fit_X if own else x, True)
        dist, idx = _nearest(q, self._fit_X, k, self.metric, self.metric_params, skip_self=own)
        return (dist, idx) if return_distance else idx


class KNeighborsClassifier(ClassifierMixin, _NeighborParams):
    """Majority of the ``n_neighbors`` closest training series
    (``_neighbors.py:172-305``)."""

    def fit(self, x, y):
        self._remember(x)
        y = np.asarray(y)
        if y.shape[0] != self._fit_X.shape[0]:
            raise ValueError("Found input variables with inconsistent numbers of samples: "
                             f"[{self._fit_X.shape[0]}, {y.shape[0]}]")
        check_classification_targets(y)
        self.classes_, self._codes = np.unique(y, return_inverse=True)
        return self

    def predict_proba(self, x):
        """Share of each class among the neighbours of each query."""
        check_is_fitted(self, "_fit_X")
        q = _series(x, True)
        k = self.n_neighbors
        if q.ndim == 3:
            # the reference selects with kth = k here (_neighbors.py:270-272)
            D = wd.pairwise_distance(q, self._fit_X, dim="mean", metric=self.metric,
                                     metric_params=self.metric_params)
            members = np.argpartition(D, k, axis=1)[:, :k]
        else:
            members = wd.argmin_distance(q, self._fit_X, k=k, metric=self.metric,
                                         metric_params=self.metric_params)
        votes = self._codes[members]
        n_cls = len(self.classes_)
        counts = np.stack([np.count_nonzero(votes == c, axis=1) for c in range(n_cls)], axis=1)
        return counts / k

    def predict(self, x):
        return self.classes_[np.argmax(self.predict_proba(x), axis=1)]


# ------------------------------------------------------------------ k-means
def _mean_rule(members, current, rng):
    return members.mean(axis=0)


def _barycenter_rule(r, g):
    """DTW (g None) or WDTW centre: a majorize-minimize barycenter started at
    the current centre, seeded from the run's RandomState."""

    def rule(members, current, rng):
        from .dtw import dtw_average

        seed = rng.randint(_SEED_BOUND)
        return dtw_average(members, r=r, g=g, init=current, method="mm", random_state=seed)

    return rule


def _lloyd(x, centres, metric, params, rule, rng, max_iter, tol, verbose):
    """One k-means run from ``centres`` (updated in place).

    Each iteration assigns every series to its closest centre (first index on
    ties), scores the assignment by the mean distance to it, stops when the
    score repeats within ``tol`` (relative), and otherwise moves each centre:
    an empty cluster takes the series farthest from its nearest centre, a
    single member becomes the centre, larger clusters go through ``rule`` --
    the order of ``_neighbors.py:308-362, 537-564``.

    Returns (last iteration, score, labels, converged).
    """
    score = -np.inf
    converged = False
    labels = None
    it = 0
    for it in range(max_iter):
        labels, nearest = wd.argmin_distance(x, centres, k=1, metric=metric,
                                             metric_params=params, return_distance=True)
        labels = labels.ravel()
        previous = score
        score = wd.paired_distance(x, centres[labels], dim="mean", metric=metric,
                                   metric_params=params).sum() / x.shape[0]
        if verbose > 0:
            print(f"Iteration {it}, {score} (prev_cost = {previous})")
        if math.isclose(score, previous, rel_tol=tol):
            converged = True
            break
        for c in range(centres.shape[0]):
            members = x[labels == c]
            if members.shape[0] == 0:
                far = int(nearest.min(axis=1).argmax())
                labels[far] = c
                centres[c] = x[far]
            elif members.shape[0] == 1:
                centres[c] = members[0]
            else:
                centres[c] = rule(members, centres[c], rng)
    return it, score, labels, converged


class KMeans(ClusterMixin, TransformerMixin, BaseEstimator):
    """k-means with euclidean, DTW or, with ``g``, WDTW distances
    (``_neighbors.py:404-612``)."""

    def __init__(self, n_clusters=8, *, metric="euclidean", r=1.0, g=None, init="random",
                 n_init="auto", max_iter=300, tol=1e-3, verbose=0, random_state=None):
        self.n_clusters = n_clusters
        self.metric = metric
        self.r = r
        self.g = g
        self.init = init
        self.n_init = n_init
        self.max_iter = max_iter
        self.tol = tol
        self.verbose = verbose
        self.random_state = random_state

    def _check_params(self):
        _require(_is_int(self.n_clusters) and self.n_clusters >= 1, "n_clusters",
                 "an int in the range [1, inf)", self.n_clusters)
        _require(self.metric in ("euclidean", "dtw"), "metric", "a str among ['dtw', 'euclidean']",
                 self.metric)
        _require(_is_real(self.r) and 0 <= self.r <= 1, "r", "a float in the range [0, 1]", self.r)
        _require(self.g is None or (_is_real(self.g) and self.g > 0), "g",
                 "None or a float in the range (0, inf)", self.g)
        _require(self.init == "random", "init", "'random'", self.init)
        _require(self.n_init == "auto" or (_is_int(self.n_init) and self.n_init >= 1), "n_init",
                 "'auto' or an int >= 1", self.n_init)
        _require(_is_int(self.max_iter) and self.max_iter >= 1, "max_iter",
                 "an int in the range [1, inf)", self.max_iter)
        _require(isinstance(self.tol, float), "tol", "a float", self.tol)

    def _distance(self):
        """Metric name and parameters the clustering measures with."""
        if self.metric == "euclidean":
            return "euclidean", None
        if self.g is None:
            return "dtw", {"r": self.r}
        return "wdtw", {"r": self.r, "g": self.g}

    def fit(self, x, y=None):
        self._check_params()
        x = wd.check_array(x, allow_3d=False, ensure_2d=True)
        metric, params = self._distance()
        rule = _mean_rule if metric == "euclidean" else _barycenter_rule(self.r, self.g)
        rng = check_random_state(self.random_state)
        best = None
        for _ in range(1 if self.n_init == "auto" else self.n_init):
            centres = x[rng.choice(x.shape[0], size=self.n_clusters, replace=False)]
            run = _lloyd(x, centres, metric, params, rule, rng, self.max_iter, self.tol,
                         self.verbose)
            if best is None or run[1] < best[1][1]:
                best = (centres, run)
        centres, (it, score, labels, converged) = best
        if not converged:
            # the last move of the centres has not been assigned yet
            labels = wd.argmin_distance(x, centres, k=1, metric=metric,
                                        metric_params=params).ravel()
        self.cluster_centers_ = centres
        self.labels_ = labels
        self.inertia_ = score
        self.n_iter_ = it
        return self

    def transform(self, x):
        """Distance of each series to each centre."""
        check_is_fitted(self, "cluster_centers_")
        x = wd.check_array(x, allow_3d=False, ensure_2d=True)
        metric, params = self._distance()
        return wd.pairwise_distance(x, self.cluster_centers_, dim="mean", metric=metric,
                                    metric_params=params or {})

    def predict(self, x):
        return self.transform(x).argmin(axis=1)
