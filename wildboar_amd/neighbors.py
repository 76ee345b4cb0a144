"""Nearest-neighbour and k-means estimators on the GPU distance path.

Mirrors of wildboar's ``distance/_neighbors.py``:

- ``NearestNeighbors``      (``_neighbors.py:18-169``)
- ``KNeighborsClassifier``  (``_neighbors.py:172-305``)
- ``KMeans``                (``_neighbors.py:308-612``)
- ``KMedoids``              (``_neighbors.py:615-930``, PAM helpers ``_cneighbors.pyx``)

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

__all__ = ["NearestNeighbors", "KNeighborsClassifier", "KMeans", "KMedoids"]


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


# ------------------------------------------------------------------ k-medoids
def _pam_optimal_swap(D, medoid_idxs, not_medoid_idxs, Djs, Ejs, block=128):
    """The best (medoid, non-medoid) swap of PAM (``_cneighbors.pyx:13-67``):
    the change in cost of every swap, each summed over the non-medoids in
    order (a running sum, as the reference's loop), the first smallest change
    if it is negative, else None.  Candidates are taken ``block`` at a time."""
    idx_j = not_medoid_idxs
    k, m = medoid_idxs.size, idx_j.size
    Dij = D[np.ix_(medoid_idxs, idx_j)]                  # (k, m): D[id_i, id_j]
    dj, ej = Djs[idx_j], Ejs[idx_j]
    in_i = Dij == dj                                     # cluster_i_bool
    ei = Ejs[medoid_idxs][None, :]
    change = np.empty((m, k))
    for h0 in range(0, m, block):
        hs = not_medoid_idxs[h0:h0 + block]
        Dhj = D[np.ix_(hs, idx_j)]                       # D[id_h, id_j]
        Djh = D[np.ix_(idx_j, hs)].T                     # D[id_j, id_h]
        second = Dhj < ej                                # second_best_medoid
        t_a = Djh - dj
        t_b = np.broadcast_to(ej - dj, t_a.shape)
        t_c = np.where(Djh < dj, Djh - dj, 0.0)
        terms = np.where(in_i[None, :, :], np.where(second[:, None, :], t_a[:, None, :],
                                                    t_b[:, None, :]), t_c[:, None, :])
        run = np.cumsum(terms, axis=2)[:, :, -1] if m else np.zeros((hs.size, k))
        Dhi = D[np.ix_(hs, medoid_idxs)]                 # D[id_h, id_i]
        Dih = D[np.ix_(medoid_idxs, hs)].T               # D[id_i, id_h]
        change[h0:h0 + block] = run + np.where(Dhi < ei, Dih, ei)
    flat = change.ravel()                                # h outer, i inner
    if flat.size == 0:
        return None
    best = int(np.argmin(flat))
    if flat[best] < 0.0:
        h, i = divmod(best, k)
        return medoid_idxs[i], not_medoid_idxs[h], flat[best]
    return None


def _pam_build(D, n_clusters):
    """PAM's greedy BUILD initialisation (``_cneighbors.pyx:70-100``)."""
    n = len(D)
    medoids = np.zeros(n_clusters, dtype=int)
    medoids[0] = np.argmin(np.sum(D, axis=0))
    not_medoids = np.delete(np.arange(n, dtype=int), medoids[0])
    Dj = D[medoids[0]].copy()
    for c in range(1, n_clusters):
        gains = np.maximum(0.0, Dj[not_medoids][None, :] - D[np.ix_(not_medoids, not_medoids)])
        cost = np.cumsum(gains, axis=1)[:, -1]
        # `>=` keeps the last of equal gains; a gain of 0 beats the start of 0
        best = int(np.flatnonzero(cost >= max(0.0, cost.max()))[-1])
        medoids[c] = not_medoids[best]
        not_medoids = np.delete(not_medoids, best)
        Dj = np.minimum(Dj, D[:, medoids[c]])
    return medoids


class _KMedoidsCluster:
    """``_neighbors.py:615-632``."""

    def __init__(self, dist, cluster_idx):
        self._dist = dist
        self._cluster_idx = cluster_idx
        self._prev_cluster_idx = cluster_idx.copy()

    def assign(self):
        self._prev_cluster_idx[:] = self._cluster_idx
        self.labels_ = self._dist[self._cluster_idx].argmin(axis=0)

    def cost(self):
        # flat indices into the distance matrix, as the reference takes them
        self.cost_ = (np.take(self._dist, self._cluster_idx[self.labels_]).sum()
                      / self._dist.shape[0])

    def update(self):
        pass


class _FastKMedoidsCluster(_KMedoidsCluster):
    """``_neighbors.py:635-647``."""

    def update(self):
        for idx in range(self._cluster_idx.shape[0]):
            cluster_idx = np.where(self.labels_ == idx)[0]
            if cluster_idx.shape[0] == 0:
                continue
            cost = self._dist[cluster_idx, cluster_idx.reshape(-1, 1)].sum(axis=1)
            min_idx = cost.argmin()
            min_cost = cost[min_idx]
            curr_cost = cost[(cluster_idx == self._cluster_idx[idx]).argmax()]
            if min_cost < curr_cost:
                self._cluster_idx[idx] = cluster_idx[min_idx]


class _PamKMedoidsCluster(_KMedoidsCluster):
    """``_neighbors.py:650-674``."""

    def __init__(self, dist, cluster_idx):
        super().__init__(dist, cluster_idx)
        self._djs, self._ejs = np.sort(self._dist[self._cluster_idx], axis=0)[[0, 1]]

    def update(self):
        not_cluster_idx = np.delete(np.arange(self._dist.shape[0]), self._cluster_idx)
        swap = _pam_optimal_swap(self._dist, self._cluster_idx, not_cluster_idx, self._djs,
                                 self._ejs)
        if swap is not None:
            i, j, _ = swap
            self._cluster_idx[self._cluster_idx == i] = j
            self._djs, self._ejs = np.sort(self._dist[self._cluster_idx], axis=0)[[0, 1]]


class KMedoids(ClusterMixin, TransformerMixin, BaseEstimator):
    """k-medoids over the GPU distance matrix (``_neighbors.py:677-930``):
    ``"fast"`` (alternate) or ``"pam"`` (BUILD + SWAP) updates."""

    def __init__(self, n_clusters=8, metric="euclidean", metric_params=None, init="random",
                 n_init="auto", algorithm="fast", max_iter=30, tol=1e-4, verbose=0,
                 n_jobs=None, random_state=None):
        self.n_clusters = n_clusters
        self.metric = metric
        self.metric_params = metric_params
        self.init = init
        self.n_init = n_init
        self.algorithm = algorithm
        self.max_iter = max_iter
        self.tol = tol
        self.verbose = verbose
        self.n_jobs = n_jobs
        self.random_state = random_state

    def _validate_params(self):
        _check_int(self.n_clusters, "n_clusters", 1)
        _check_metric_name(self.metric, set(wd.METRICS) | {"precomputed"})
        if self.metric_params is not None and not isinstance(self.metric_params, dict):
            raise ValueError("The 'metric_params' parameter must be None or a dict.")
        if self.init not in ("random", "auto", "min"):
            raise ValueError(f"The 'init' parameter must be a str among "
                             f"{{'auto', 'min', 'random'}}, got {self.init!r}.")
        if not (self.n_init == "auto" or (isinstance(self.n_init, numbers.Integral)
                                          and not isinstance(self.n_init, bool)
                                          and self.n_init >= 1)):
            raise ValueError(f"The 'n_init' parameter must be 'auto' or an int >= 1, "
                             f"got {self.n_init!r}.")
        if self.algorithm not in ("fast", "pam"):
            raise ValueError(f"The 'algorithm' parameter must be a str among "
                             f"{{'fast', 'pam'}}, got {self.algorithm!r}.")
        _check_int(self.max_iter, "max_iter", 1)
        if isinstance(self.tol, bool) or not isinstance(self.tol, numbers.Real) or self.tol < 0:
            raise ValueError(f"The 'tol' parameter must be a float >= 0, got {self.tol!r}.")

    def fit(self, x, y=None):
        import warnings

        from sklearn.exceptions import ConvergenceWarning

        x = (np.asarray(x, dtype=float) if self.metric == "precomputed"
             else _validate_x(x, allow_3d=True))
        self._validate_params()
        if self.n_init == "auto":
            n_init = 1 if self.init == "auto" else 10
        else:
            n_init = self.n_init
        if n_init > 1 and self.init != "random":
            n_init = 1
        max_iter = self.max_iter
        if self.algorithm == "pam" and self.n_clusters == 1 and self.max_iter != 0:
            warnings.warn("n_clusters must be larger than 1 if max_iter larger than 0")
            max_iter = 0
        random_state = check_random_state(self.random_state)
        if self.metric == "precomputed":
            dist = x
        else:
            dist = wd.pairwise_distance(x, dim="mean", metric=self.metric,
                                        metric_params=self.metric_params)
        best_iter, best_cost, best_clusterer, best_reassign = 0, np.inf, None, None
        clusterer = None
        for i in range(n_init):
            if self.verbose:
                print(f"Running initialization {i}/{n_init}")
            it, clusterer, reassign = self._fit_one_init(dist, random_state=random_state,
                                                         max_iter=max_iter,
                                                         warn=warnings.warn,
                                                         warning=ConvergenceWarning)
            if clusterer.cost_ < best_cost:
                best_cost = clusterer.cost_
                best_clusterer = clusterer
                best_iter = it
                best_reassign = reassign
        if best_reassign:
            clusterer.assign()  # the reference reassigns the last run's clusterer
        self.cluster_centers_ = (None if self.metric == "precomputed"
                                 else x[best_clusterer._cluster_idx])
        self.inertia_ = best_clusterer.cost_
        self.medoid_indices_ = best_clusterer._cluster_idx
        self.n_iter_ = best_iter
        self.labels_ = best_clusterer.labels_
        return self

    def _fit_one_init(self, dist, *, max_iter, random_state, warn, warning):
        centers = self._init_centers(dist, self.n_clusters, random_state=random_state)
        cls = _PamKMedoidsCluster if self.algorithm == "pam" else _FastKMedoidsCluster
        clusterer = cls(dist, centers)
        clusterer.assign()
        reassign = False
        prev_cost = np.inf
        it = 0
        for it in range(max_iter):
            clusterer.update()
            clusterer.cost()
            if self.verbose:
                print(f"Iteration {it}/{self.max_iter}: current_cost={clusterer.cost_}")
            if math.isclose(clusterer.cost_, prev_cost, rel_tol=self.tol):
                reassign = True
                break
            prev_cost = clusterer.cost_
            clusterer.assign()
        if it + 1 == self.max_iter:
            warn("Maximum number of iterations reached before convergence. Consider "
                 "increasing max_iter to improve the fit", warning)
        return it, clusterer, reassign

    def _init_centers(self, dist, n_clusters, random_state):
        if self.init == "random":
            return random_state.choice(dist.shape[0], n_clusters, replace=False)
        if self.init == "min" or (self.algorithm == "fast" and self.init == "auto"):
            return np.argpartition(dist.sum(axis=1), n_clusters)[:n_clusters]
        return _pam_build(dist, n_clusters)

    def predict(self, x):
        return self.transform(x).argmin(axis=1)

    def transform(self, x):
        check_is_fitted(self, "medoid_indices_")
        if self.metric == "precomputed":
            return np.asarray(x, dtype=float)[:, self.medoid_indices_]
        x = _validate_x(x, allow_3d=True)
        return wd.pairwise_distance(x, self.cluster_centers_, dim="mean", metric=self.metric,
                                    metric_params=self.metric_params)
