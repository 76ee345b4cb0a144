"""Nearest-neighbour and k-means estimators on the GPU distance path.

Mirrors of wildboar's ``distance/_neighbors.py``:

- ``NearestNeighbors``      (``_neighbors.py:18-169``)
- ``KNeighborsClassifier``  (``_neighbors.py:172-305``)
- ``KMeans``                (``_neighbors.py:308-612``)

Same parameters, fitted attributes and results; the distances come from
``wildboar_amd.distance`` (``argmin_distance`` / ``pairwise_distance`` /
``paired_distance`` in float64, bit-identical to the reference) and the DTW
centroids of ``KMeans`` from ``wildboar_amd.dtw.dtw_average``.  The
bookkeeping around them (votes, label maps, centroid means) is numpy, as in
the reference.
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


def _check_int(v, name, low):
    if isinstance(v, bool) or not isinstance(v, numbers.Integral) or v < low:
        raise ValueError(f"The '{name}' parameter must be an int in the range [{low}, inf), "
                         f"got {v!r}.")


def _check_metric_name(metric, allowed):
    if not isinstance(metric, str) or metric not in allowed:
        raise ValueError(f"The 'metric' parameter must be a str among {sorted(allowed)}, "
                         f"got {metric!r}.")


def _validate_x(x, allow_3d):
    x = wd.check_array(x, allow_3d=allow_3d, ensure_2d=True)
    return x


class NearestNeighbors(TransformerMixin, BaseEstimator):
    """Unsupervised nearest neighbours (``_neighbors.py:18-169``)."""

    def __init__(self, n_neighbors=5, *, metric="euclidean", metric_params=None, n_jobs=None):
        self.n_neighbors = n_neighbors
        self.metric = metric
        self.metric_params = metric_params
        self.n_jobs = n_jobs

    def _validate_params(self):
        _check_int(self.n_neighbors, "n_neighbors", 1)
        _check_metric_name(self.metric, wd.METRICS)
        if self.metric_params is not None and not isinstance(self.metric_params, dict):
            raise ValueError("The 'metric_params' parameter must be a dict or None.")

    def fit(self, x, y=None):
        self._validate_params()
        self._fit_X = _validate_x(x, allow_3d=True).copy()
        return self

    def kneighbors(self, x=None, n_neighbors=None, return_distance=True):
        check_is_fitted(self, "_fit_X")
        if n_neighbors is None:
            n_neighbors = self.n_neighbors
        if x is not None:
            query_is_train = False
            x = _validate_x(x, allow_3d=True)
        else:
            query_is_train = True
            x = self._fit_X
        # a multivariate array with one dimension is univariate
        if x.ndim == 3 and x.shape[1] == 1:
            x = x.reshape(x.shape[0], -1)
        if x.ndim == 3:
            dists = wd.pairwise_distance(x, self._fit_X, dim="mean", metric=self.metric,
                                         metric_params=self.metric_params)
            if query_is_train:
                np.fill_diagonal(dists, np.inf)
            sample_range = np.arange(x.shape[0])[:, None]
            neigh_ind = np.argpartition(dists, n_neighbors - 1, axis=1)[:, :n_neighbors]
            neigh_dist = dists[sample_range, neigh_ind]
            sort_inds = np.argsort(neigh_dist, axis=1)
            neigh_ind = neigh_ind[sample_range, sort_inds]
            if return_distance:
                return neigh_dist[sample_range, sort_inds], neigh_ind
            return neigh_ind
        k = n_neighbors + 1 if query_is_train else n_neighbors
        indices, distances = wd.argmin_distance(x, self._fit_X, metric=self.metric,
                                                metric_params=self.metric_params, k=k,
                                                return_distance=True, sorted=True)
        if query_is_train:
            mask = indices != np.arange(x.shape[0])[:, None]
            indices = indices[mask].reshape(x.shape[0], -1)
            distances = distances[mask].reshape(x.shape[0], -1)
        if return_distance:
            return distances, indices
        return indices


class KNeighborsClassifier(ClassifierMixin, BaseEstimator):
    """k-nearest-neighbour classifier (``_neighbors.py:172-305``): the class
    shares of the ``n_neighbors`` closest training series."""

    def __init__(self, n_neighbors=5, *, metric="euclidean", metric_params=None, n_jobs=None):
        self.n_neighbors = n_neighbors
        self.metric = metric
        self.metric_params = metric_params
        self.n_jobs = n_jobs

    def _validate_params(self):
        _check_int(self.n_neighbors, "n_neighbors", 1)
        _check_metric_name(self.metric, wd.METRICS)
        if self.metric_params is not None and not isinstance(self.metric_params, dict):
            raise ValueError("The 'metric_params' parameter must be a dict or None.")

    def fit(self, x, y):
        self._validate_params()
        x = _validate_x(x, allow_3d=True)
        y = np.asarray(y)
        if y.shape[0] != x.shape[0]:
            raise ValueError(f"Found input variables with inconsistent numbers of samples: "
                             f"[{x.shape[0]}, {y.shape[0]}]")
        check_classification_targets(y)
        self._fit_X = x.copy()
        self.classes_, self._y = np.unique(y, return_inverse=True)
        return self

    def predict_proba(self, x):
        check_is_fitted(self, "_fit_X")
        x = _validate_x(x, allow_3d=True)
        if x.ndim == 3 and x.shape[1] == 1:
            x = x.reshape(x.shape[0], -1)
        if x.ndim == 3:
            dists = wd.pairwise_distance(x, self._fit_X, dim="mean", metric=self.metric,
                                         metric_params=self.metric_params)
            closest = np.argpartition(dists, self.n_neighbors, axis=1)[:, :self.n_neighbors]
        else:
            closest = wd.argmin_distance(x, self._fit_X, metric=self.metric,
                                         metric_params=self.metric_params, k=self.n_neighbors)
        preds = self._y[closest]
        probs = np.empty((x.shape[0], len(self.classes_)), dtype=float)
        for i in range(len(self.classes_)):
            probs[:, i] = np.sum(preds == i, axis=1) / self.n_neighbors
        return probs

    def predict(self, x):
        proba = np.argmax(self.predict_proba(x), axis=1)
        return np.take(self.classes_, proba)


class _KMeansCluster:
    """Assignment, cost and centroid update of one k-means run
    (``_neighbors.py:308-362``)."""

    def __init__(self, init, metric, metric_params, random_state):
        self.metric = metric
        self.metric_params = metric_params
        self.centroids_ = init
        self.random_state = random_state

    def cost(self, x):
        return wd.paired_distance(x, self.centroids_[self.assigned_], dim="mean",
                                  metric=self.metric,
                                  metric_params=self.metric_params).sum() / x.shape[0]

    def assign(self, x):
        self.assigned_, self.distance_ = wd.argmin_distance(
            x, self.centroids_, k=1, metric=self.metric, metric_params=self.metric_params,
            return_distance=True)
        self.assigned_ = np.ravel(self.assigned_)

    def update(self, x):
        for c in range(self.centroids_.shape[0]):
            cluster = x[self.assigned_ == c]
            if cluster.shape[0] == 0:
                far_from_center = self.distance_.min(axis=1).argmax()
                self.assigned_[far_from_center] = c
                self.centroids_[c] = x[far_from_center]
            elif cluster.shape[0] == 1:
                self.centroids_[c] = cluster
            else:
                self.centroids_[c] = self._update_centroid(cluster, self.centroids_[c])


class _EuclideanCluster(_KMeansCluster):
    def __init__(self, centroids, random_state):
        super().__init__(centroids, "euclidean", None, random_state)

    def _update_centroid(self, group, _centroid):
        return group.mean(axis=0)


class _DtwCluster(_KMeansCluster):
    def __init__(self, centroids, r, random_state):
        super().__init__(centroids, "dtw", {"r": r}, random_state)
        self.r = r

    def _update_centroid(self, group, current_center):
        from .dtw import dtw_average

        return dtw_average(group, r=self.r, init=current_center, method="mm",
                           random_state=self.random_state.randint(np.iinfo(np.int32).max))


class _WDtwCluster(_KMeansCluster):
    def __init__(self, centroids, r, g, random_state):
        super().__init__(centroids, "wdtw", {"r": r, "g": g}, random_state)
        self.r = r
        self.g = g

    def _update_centroid(self, group, current_center):
        from .dtw import dtw_average

        return dtw_average(group, r=self.r, g=self.g, init=current_center, method="mm",
                           random_state=self.random_state.randint(np.iinfo(np.int32).max))


class KMeans(ClusterMixin, TransformerMixin, BaseEstimator):
    """k-means with euclidean, DTW or (``g`` given) WDTW distances
    (``_neighbors.py:404-612``); the elastic centroids are majorize-minimize
    barycenters."""

    def __init__(self, n_clusters=8, *, metric="euclidean", r=1.0, g=None, init="random",
                 n_init="auto", max_iter=300, tol=1e-3, verbose=0, random_state=None):
        self.metric = metric
        self.r = r
        self.g = g
        self.n_clusters = n_clusters
        self.init = init
        self.n_init = n_init
        self.max_iter = max_iter
        self.tol = tol
        self.verbose = verbose
        self.random_state = random_state

    def _validate_params(self):
        _check_int(self.n_clusters, "n_clusters", 1)
        _check_metric_name(self.metric, {"euclidean", "dtw"})
        if (isinstance(self.r, bool) or not isinstance(self.r, numbers.Real)
                or not 0 <= self.r <= 1):
            raise ValueError(f"The 'r' parameter must be a float in the range [0, 1], "
                             f"got {self.r!r}.")
        if self.g is not None:
            if isinstance(self.g, bool) or not isinstance(self.g, numbers.Real) or not self.g > 0:
                raise ValueError(f"The 'g' parameter must be None or a float in the range "
                                 f"(0, inf), got {self.g!r}.")
        if self.init != "random":
            raise ValueError(f"The 'init' parameter must be 'random', got {self.init!r}.")
        if not (self.n_init == "auto" or (isinstance(self.n_init, numbers.Integral)
                                          and not isinstance(self.n_init, bool)
                                          and self.n_init >= 1)):
            raise ValueError(f"The 'n_init' parameter must be 'auto' or an int >= 1, "
                             f"got {self.n_init!r}.")
        _check_int(self.max_iter, "max_iter", 1)
        if not isinstance(self.tol, float):
            raise ValueError(f"The 'tol' parameter must be a float, got {self.tol!r}.")

    def fit(self, x, y=None):
        self._validate_params()
        x = _validate_x(x, allow_3d=False)
        n_init = 1 if self.n_init == "auto" else self.n_init
        random_state = check_random_state(self.random_state)
        best_cost = np.inf
        best_cluster = None
        best_reassign = True
        best_iter = None
        for _ in range(n_init):
            it, cost, cluster, reassign = self._fit_one_init(x, random_state)
            if cost < best_cost:
                best_cost = cost
                best_cluster = cluster
                best_reassign = reassign
                best_iter = it
        self.n_iter_ = best_iter
        self.inertia_ = best_cost
        self.cluster_centers_ = best_cluster.centroids_
        if best_reassign:
            best_cluster.assign(x)
        self.labels_ = best_cluster.assigned_
        return self

    def _fit_one_init(self, x, random_state):
        centroids = x[random_state.choice(x.shape[0], size=self.n_clusters, replace=False)]
        if self.metric == "euclidean":
            cluster = _EuclideanCluster(centroids, random_state=random_state)
        elif self.g is None:
            cluster = _DtwCluster(centroids, r=self.r, random_state=random_state)
        else:
            cluster = _WDtwCluster(centroids, r=self.r, g=self.g, random_state=random_state)
        prev_cost = np.inf
        cost = -np.inf
        reassign = True
        it = 0
        for it in range(self.max_iter):
            cluster.assign(x)
            prev_cost, cost = cost, cluster.cost(x)
            if self.verbose > 0:
                print(f"Iteration {it}, {cost} (prev_cost = {prev_cost})")
            if math.isclose(cost, prev_cost, rel_tol=self.tol):
                reassign = False
                break
            cluster.update(x)
        return it, cost, cluster, reassign

    def predict(self, x):
        return self.transform(x).argmin(axis=1)

    def transform(self, x):
        check_is_fitted(self, "cluster_centers_")
        x = _validate_x(x, allow_3d=False)
        metric_params = {}
        metric = self.metric
        if self.metric == "dtw":
            metric_params["r"] = self.r
            if self.g is not None:
                metric_params["g"] = self.g
                metric = "wdtw"
        return wd.pairwise_distance(x, self.cluster_centers_, dim="mean", metric=metric,
                                    metric_params=metric_params)
