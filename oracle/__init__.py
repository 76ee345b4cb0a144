"""Float64 CPU oracle for wildboar's distance hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product (``wildboar_amd``) never imports it.

The arithmetic lives in ``oracle.c`` (a restatement of
``distance/_elastic.pyx:869-940`` and friends; every function there cites the
reference lines it follows).  This module is a thin ctypes layer over it with
the reference's driver semantics:

- ``pairwise``   : ``_pairwise_distance``   (``_cdistance.pyx:1184-1205``)
- ``singleton``  : ``_singleton_pairwise_distance`` (``:1248-1267``)
- ``paired``     : ``_paired_distance``     (``:1632-1652``, D(y[i], x[i]))
- ``argmin``     : ``_argmin_distance``     (``:1348-1378``, heap order)
- ``subsequence``: ``_pairwise_subsequence_distance`` (``:1018-1037``)
- ``argmin_subsequence``: ``_ArgminSubsequenceDistance`` (``:1408-1434``,
  heap order)
- ``subsequence_matches``: the per-series window matches behind
  ``_subsequence_match`` (``:1065-1100``; ``_elastic.pyx:658-698``,
  ``_metric.pyx:762-795``)
- ``dtw_grad``   : path-backtrack gradient (``_elastic.pyx:1011-1073`` +
  ``dtw.py:395-423``)
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

EUCLIDEAN, DTW, WDTW, DDTW, WDDTW, ADTW = 0, 1, 2, 3, 4, 5
_KINDS = {"euclidean": EUCLIDEAN, "dtw": DTW, "wdtw": WDTW, "ddtw": DDTW, "wddtw": WDDTW,
          "adtw": ADTW}


def _gp(metric, g, p):
    """The metric's second parameter: the penalty p for adtw, else g."""
    return float(p) if metric == "adtw" else float(g)

_lib = None


def build():
    """Compile oracle.c into oracle/lib/liboracle.so (gcc, -O2, no FMA)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    d, i64, i32 = ctypes.c_double, ctypes.c_int64, ctypes.c_int
    p = ctypes.c_void_p
    lib.orc_compute_r.restype = i64
    lib.orc_compute_r.argtypes = [i64, d]
    lib.orc_wdtw_weights.restype = None
    lib.orc_wdtw_weights.argtypes = [i64, d, p]
    lib.orc_dtw_distance.restype = d
    lib.orc_dtw_distance.argtypes = [p, i64, p, i64, i64, p, p, p, d]
    lib.orc_euclidean_distance.restype = d
    lib.orc_euclidean_distance.argtypes = [p, i64, p, i64, d, p]
    lib.orc_metric_distance.restype = d
    lib.orc_metric_distance.argtypes = [i32, d, d, p, i64, p, i64, i64]
    lib.orc_pairwise.restype = i32
    lib.orc_pairwise.argtypes = [i32, d, d, p, i64, i64, i64, p, i64, i64, i64, p, i32]
    lib.orc_singleton.restype = i32
    lib.orc_singleton.argtypes = [i32, d, d, p, i64, i64, i64, i64, p, i32]
    lib.orc_paired.restype = i32
    lib.orc_paired.argtypes = [i32, d, d, p, p, i64, i64, i64, i64, p, i32]
    lib.orc_argmin.restype = i32
    lib.orc_argmin.argtypes = [i32, d, d, p, i64, i64, i64, p, i64, i64, i64, i64, p, p, p, i32]
    lib.orc_subsequence_pairwise.restype = i32
    lib.orc_subsequence_pairwise.argtypes = [i32, d, d, p, p, p, i64, p, i64, i64, i64, i64, p, p, i32]
    lib.orc_window_stats.restype = None
    lib.orc_window_stats.argtypes = [p, i64, i64, p, p]
    lib.orc_warp_width.restype = i64
    lib.orc_warp_width.argtypes = [i64, d]
    lib.orc_scaled_euclidean_matches.restype = i64
    lib.orc_scaled_euclidean_matches.argtypes = [p, i64, d, d, p, i64, d, p, p]
    lib.orc_scaled_dtw_matches.restype = i64
    lib.orc_scaled_dtw_matches.argtypes = [p, i64, d, d, p, i64, d, d, p, p]
    lib.orc_scaled_wrap_pairwise.restype = i32
    lib.orc_scaled_wrap_pairwise.argtypes = [i32, d, d, p, p, p, p, p, i64, p, i64, i64, p, p]
    lib.orc_scaled_argmin_subsequence.restype = i32
    lib.orc_scaled_argmin_subsequence.argtypes = [i32, d, d, p, p, p, i64, p, i64, i64, p, p]
    lib.orc_scaled_wrap_matches.restype = i64
    lib.orc_scaled_wrap_matches.argtypes = [i32, d, d, p, i64, d, d, p, i64, d, p, p]
    lib.orc_scaled_subsequence_pairwise.restype = i32
    lib.orc_scaled_subsequence_pairwise.argtypes = [i32, d, p, p, p, p, p, i64, p, i64, i64, i32,
                                                    p, p]
    lib.orc_argmin_replay.restype = None
    lib.orc_argmin_replay.argtypes = [p, i64, i64, p, p, p]
    lib.orc_dtw_cost_matrix.restype = i32
    lib.orc_dtw_cost_matrix.argtypes = [p, i64, p, i64, i64, p, p]
    lib.orc_dtw_mapping.restype = i64
    lib.orc_dtw_mapping.argtypes = [p, i64, i64, p, p]
    lib.orc_dtw_grad.restype = d
    lib.orc_dtw_grad.argtypes = [p, i64, p, i64, d, p, p, p]
    lib.orc_argmin_subsequence.restype = i32
    lib.orc_argmin_subsequence.argtypes = [i32, d, d, p, p, p, i64, p, i64, i64, i64, i64, p, p]
    lib.orc_dtw_alignment.restype = i32
    lib.orc_dtw_alignment.argtypes = [p, i64, p, i64, i64, p, p]
    lib.orc_average_slope.restype = None
    lib.orc_average_slope.argtypes = [p, i64, p]
    lib.orc_adtw_distance.restype = d
    lib.orc_adtw_distance.argtypes = [p, i64, p, i64, i64, p, p, d, d]
    lib.orc_dtw_envelop.restype = i32
    lib.orc_dtw_envelop.argtypes = [p, i64, i64, p, p]
    lib.orc_lb_keogh.restype = d
    lib.orc_lb_keogh.argtypes = [p, i64, p, p]
    lib.orc_subsequence_matches.restype = i64
    lib.orc_subsequence_matches.argtypes = [i32, d, d, p, i64, p, i64, d, p, p]
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _ts(a):
    """(n, dims, t) C-contiguous float64, as utils/validation.py:_check_ts_array."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(1, 1, -1)
    elif a.ndim == 2:
        a = a.reshape(a.shape[0], 1, a.shape[1])
    return np.ascontiguousarray(a)


def _kind(metric):
    return _KINDS[metric]


def default_threads():
    return os.cpu_count() or 1


def compute_r(length, r):
    return _load().orc_compute_r(int(length), float(r))


def wdtw_weights(n, g):
    out = np.empty(n, dtype=np.float64)
    _load().orc_wdtw_weights(n, float(g), out.ctypes.data)
    return out


def dtw_distance(x, y, band, weights=None, min_dist=np.inf):
    """Raw dtw_distance (squared cost) with an integer band ``band``."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = max(x.shape[0], y.shape[0]) + 1
    cost = np.empty(n)
    cost_prev = np.empty(n)
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=np.float64)
    return _load().orc_dtw_distance(
        x.ctypes.data, x.shape[0], y.ctypes.data, y.shape[0], int(band),
        cost.ctypes.data, cost_prev.ctypes.data, _ptr(weights), float(min_dist))


def average_slope(q):
    """average_slope of a 1-D series: len - 2 values."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.empty(max(q.shape[0] - 2, 0))
    _load().orc_average_slope(q.ctypes.data, q.shape[0], out.ctypes.data)
    return out


def adtw_distance(x, y, band, p, min_dist=np.inf):
    """Raw adtw_distance (squared cost) with an integer band."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = max(x.shape[0], y.shape[0]) + 1
    cost = np.empty(n)
    cost_prev = np.empty(n)
    return _load().orc_adtw_distance(x.ctypes.data, x.shape[0], y.ctypes.data, y.shape[0],
                                     int(band), cost.ctypes.data, cost_prev.ctypes.data,
                                     float(p), float(min_dist))


def metric_distance(x, y, metric="dtw", r=1.0, g=0.05, weight_n=0, p=1.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return _load().orc_metric_distance(_kind(metric), float(r), _gp(metric, g, p),
                                       x.ctypes.data, x.shape[0],
                                       y.ctypes.data, y.shape[0], int(weight_n))


def pairwise(X, Y=None, metric="dtw", r=1.0, g=0.05, dim=0, n_threads=None, p=1.0):
    n_threads = n_threads or default_threads()
    g = _gp(metric, g, p)
    X = _ts(X)
    lib = _load()
    if Y is None:
        out = np.empty((X.shape[0], X.shape[0]))
        rc = lib.orc_singleton(_kind(metric), r, g, X.ctypes.data, X.shape[0],
                               X.shape[1], X.shape[2], dim, out.ctypes.data,
                               n_threads)
    else:
        Y = _ts(Y)
        out = np.empty((X.shape[0], Y.shape[0]))
        rc = lib.orc_pairwise(_kind(metric), r, g, X.ctypes.data, X.shape[0],
                              X.shape[1], X.shape[2], Y.ctypes.data, Y.shape[0],
                              Y.shape[2], dim, out.ctypes.data, n_threads)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out


def paired(X, Y, metric="dtw", r=1.0, g=0.05, dim=0, n_threads=None, p=1.0):
    """out[i] = D(Y[i], X[i]) -- the reference's argument swap."""
    n_threads = n_threads or default_threads()
    g = _gp(metric, g, p)
    X = _ts(X)
    Y = _ts(Y)
    out = np.empty(X.shape[0])
    rc = _load().orc_paired(_kind(metric), r, g, X.ctypes.data, Y.ctypes.data,
                            X.shape[0], X.shape[1], X.shape[2], dim,
                            out.ctypes.data, n_threads)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out


def argmin(X, Y, k=1, metric="dtw", r=1.0, g=0.05, dim=0, lower_bound=None,
           n_threads=None, p=1.0):
    n_threads = n_threads or default_threads()
    g = _gp(metric, g, p)
    X = _ts(X)
    Y = _ts(Y)
    k = min(k, Y.shape[0])
    idx = np.zeros((X.shape[0], k), dtype=np.int64)
    dist = np.zeros((X.shape[0], k))
    if lower_bound is not None:
        lower_bound = np.ascontiguousarray(lower_bound, dtype=np.float64)
    rc = _load().orc_argmin(_kind(metric), r, g, X.ctypes.data, X.shape[0],
                            X.shape[1], X.shape[2], Y.ctypes.data, Y.shape[0],
                            Y.shape[2], dim, k, _ptr(lower_bound),
                            idx.ctypes.data, dist.ctypes.data, n_threads)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return idx, dist


def argmin_replay(row, k, lb_row=None):
    row = np.ascontiguousarray(row, dtype=np.float64)
    idx = np.zeros(k, dtype=np.int64)
    dist = np.zeros(k)
    if lb_row is not None:
        lb_row = np.ascontiguousarray(lb_row, dtype=np.float64)
    _load().orc_argmin_replay(row.ctypes.data, row.shape[0], k, _ptr(lb_row),
                              idx.ctypes.data, dist.ctypes.data)
    return idx, dist


def pack_shapelets(shapelets):
    lens = np.array([len(s) for s in shapelets], dtype=np.int64)
    offs = np.zeros(len(shapelets), dtype=np.int64)
    if len(shapelets) > 1:
        offs[1:] = np.cumsum(lens)[:-1]
    packed = np.concatenate([np.asarray(s, dtype=np.float64) for s in shapelets])
    return np.ascontiguousarray(packed), offs, lens


def subsequence(shapelets, X, metric="dtw", r=1.0, g=0.05, dim=0, n_threads=None, p=1.0):
    """(dist, idx), each (n_samples, n_shapelets)."""
    n_threads = n_threads or default_threads()
    g = _gp(metric, g, p)
    X = _ts(X)
    S, offs, lens = pack_shapelets(shapelets)
    n = X.shape[0]
    dist = np.empty((n, len(shapelets)))
    idx = np.empty((n, len(shapelets)), dtype=np.int64)
    rc = _load().orc_subsequence_pairwise(
        _kind(metric), r, g, S.ctypes.data, offs.ctypes.data, lens.ctypes.data,
        len(shapelets), X.ctypes.data, n, X.shape[1], X.shape[2], dim,
        dist.ctypes.data, idx.ctypes.data, n_threads)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return dist, idx


def pattern_stats(shapelets):
    """(mean, std) of each pattern as ScaledSubsequenceMetric.from_array
    (distance/_cdistance.pyx:453-467): np.mean, np.std, std <= 1e-13 -> 0."""
    mean = np.array([np.mean(np.asarray(s, dtype=np.float64)) for s in shapelets])
    std = np.array([np.std(np.asarray(s, dtype=np.float64)) for s in shapelets])
    std[std <= 1e-13] = 0.0
    return mean, std


def window_stats(x, L):
    """Running (mean, std) of every length-L window of the 1-D series x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nw = x.shape[0] - L + 1
    mean, std = np.empty(nw), np.empty(nw)
    _load().orc_window_stats(x.ctypes.data, x.shape[0], L, mean.ctypes.data, std.ctypes.data)
    return mean, std


def warp_width(L, r):
    return int(_load().orc_warp_width(L, r))


def scaled_subsequence(shapelets, X, metric="dtw", r=1.0, dim=0, cascade=True, g=0.05, p=1.0):
    """(dist, idx) of pairwise_subsequence_distance(scale=True) for
    metric "euclidean" (ScaledEuclideanSubsequenceMetric), "dtw"
    (ScaledDtwSubsequenceMetric, the UCR-suite cascade; cascade=False:
    every window's full scaled DTW, the exhaustive minimum) or "wdtw",
    "ddtw", "wddtw", "adtw" (ScaledSubsequenceMetricWrap)."""
    X = np.ascontiguousarray(_ts(X)[:, dim, :])
    S, offs, lens = pack_shapelets(shapelets)
    mean, std = pattern_stats(shapelets)
    n, t = X.shape
    dist = np.empty((n, len(shapelets)))
    idx = np.empty((n, len(shapelets)), dtype=np.int64)
    if metric in ("wdtw", "ddtw", "wddtw", "adtw"):
        # ScaledSubsequenceMetricWrap around the metric (incremental statistics)
        rc = _load().orc_scaled_wrap_pairwise(
            _kind(metric), r, _gp(metric, g, p), S.ctypes.data, offs.ctypes.data,
            lens.ctypes.data, mean.ctypes.data, std.ctypes.data, len(shapelets), X.ctypes.data,
            n, t, dist.ctypes.data, idx.ctypes.data)
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        return dist, idx
    rc = _load().orc_scaled_subsequence_pairwise(
        {"euclidean": 0, "dtw": 1}[metric], r, S.ctypes.data, offs.ctypes.data,
        lens.ctypes.data, mean.ctypes.data, std.ctypes.data, len(shapelets), X.ctypes.data, n,
        t, int(cascade), dist.ctypes.data, idx.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return dist, idx


def scaled_subsequence_matches(s, x, threshold, metric="dtw", r=1.0, g=0.05, p=1.0):
    """(indices, distances) of the windows of the 1-D series x that the
    reference's scaled matching records for pattern s (scaled_euclidean /
    scaled_dtw ``_matches``), in window order."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    mean, std = pattern_stats([s])
    nw = x.shape[0] - s.shape[0] + 1
    idx = np.zeros(max(nw, 1), dtype=np.int64)
    dist = np.zeros(max(nw, 1))
    lib = _load()
    if metric in ("wdtw", "ddtw", "wddtw", "adtw"):
        n = lib.orc_scaled_wrap_matches(_kind(metric), float(r), _gp(metric, g, p), s.ctypes.data,
                                        s.shape[0], mean[0], std[0], x.ctypes.data, x.shape[0],
                                        float(threshold), dist.ctypes.data, idx.ctypes.data)
    elif metric == "euclidean":
        n = lib.orc_scaled_euclidean_matches(s.ctypes.data, s.shape[0], mean[0], std[0],
                                             x.ctypes.data, x.shape[0], float(threshold),
                                             dist.ctypes.data, idx.ctypes.data)
    else:
        n = lib.orc_scaled_dtw_matches(s.ctypes.data, s.shape[0], mean[0], std[0], x.ctypes.data,
                                       x.shape[0], float(r), float(threshold), dist.ctypes.data,
                                       idx.ctypes.data)
    if n < 0:
        raise MemoryError("oracle allocation failed")
    return idx[:n].copy(), dist[:n].copy()


def scaled_argmin_subsequence(shapelets, X, k=1, metric="dtw", r=1.0, g=0.05, p=1.0):
    """(idx, dist), each (n_samples, k) in heap order: argmin_subsequence_distance
    with scale=True (pattern i against series i, both z-normalised)."""
    X = np.ascontiguousarray(_ts(X)[:, 0, :])
    S, offs, lens = pack_shapelets(shapelets)
    n, t = X.shape
    idx = np.zeros((n, k), dtype=np.int64)
    dist = np.zeros((n, k))
    rc = _load().orc_scaled_argmin_subsequence(
        _kind(metric), r, _gp(metric, g, p), S.ctypes.data, offs.ctypes.data, lens.ctypes.data,
        n, X.ctypes.data, t, k, idx.ctypes.data, dist.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return idx, dist


def argmin_subsequence(shapelets, X, k=1, metric="dtw", r=1.0, g=0.05, dim=0):
    """(idx, dist), each (n_samples, k) in heap order: shapelet i against
    series i."""
    X = _ts(X)
    S, offs, lens = pack_shapelets(shapelets)
    n = X.shape[0]
    idx = np.zeros((n, k), dtype=np.int64)
    dist = np.zeros((n, k))
    rc = _load().orc_argmin_subsequence(
        _kind(metric), r, g, S.ctypes.data, offs.ctypes.data, lens.ctypes.data, n,
        X.ctypes.data, X.shape[1], X.shape[2], dim, k, idx.ctypes.data, dist.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return idx, dist


def subsequence_matches(s, x, threshold, metric="dtw", r=1.0, g=0.05, p=1.0):
    """(indices, distances) of the windows of the 1-D series x whose squared
    cost against s is <= threshold ** 2, in window order."""
    g = _gp(metric, g, p)
    s = np.ascontiguousarray(s, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nw = x.shape[0] - s.shape[0] + 1
    idx = np.zeros(max(nw, 1), dtype=np.int64)
    dist = np.zeros(max(nw, 1))
    m = _load().orc_subsequence_matches(_kind(metric), r, g, s.ctypes.data, s.shape[0],
                                        x.ctypes.data, x.shape[0], float(threshold),
                                        dist.ctypes.data, idx.ctypes.data)
    if m < 0:
        raise MemoryError("oracle allocation failed")
    return idx[:m].copy(), dist[:m].copy()


def cost_matrix(x, y, band, weights=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty((x.shape[0], y.shape[0]))
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=np.float64)
    _load().orc_dtw_cost_matrix(x.ctypes.data, x.shape[0], y.ctypes.data,
                                y.shape[0], int(band), _ptr(weights),
                                out.ctypes.data)
    return out


def mapping(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0] + A.shape[1]
    pi = np.empty(n, dtype=np.int64)
    pj = np.empty(n, dtype=np.int64)
    m = _load().orc_dtw_mapping(A.ctypes.data, A.shape[0], A.shape[1],
                                pi.ctypes.data, pj.ctypes.data)
    return pi[:m][::-1].copy(), pj[:m][::-1].copy()


def dtw_grad(x, y, r, weights=None):
    """(D, dD/dx, dD/dy) along the backtracked optimal path."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    gx = np.zeros(x.shape[0])
    gy = np.zeros(y.shape[0])
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=np.float64)
    D = _load().orc_dtw_grad(x.ctypes.data, x.shape[0], y.ctypes.data,
                             y.shape[0], float(r), _ptr(weights),
                             gx.ctypes.data, gy.ctypes.data)
    return D, gx, gy


def warp_size(x_size, r, y_size=0):
    """_compute_warp_size, distance/dtw.py:38-40."""
    return max(int(np.floor(max(x_size, y_size) * r)), 1)


def jeong_weight(n, g):
    """distance/dtw.py:330-359 (numpy's exp, as the reference)."""
    return 1.0 / (1.0 + np.exp(-g * (np.arange(n, dtype=float) - n / 2.0)))


def mm_dtw_average(X, mean, r=1.0, tol=1e-5, max_epoch=50, sample_weight=None, g=None):
    """_mm_dtw_average, distance/dtw.py:640-666, with dtw_alignment's warp
    size (dtw.py:246-290) over the alignment matrix and backtrack above
    (wdtw_alignment's weights when g is given); the cost is the mean DTW or
    WDTW distance (dtw.py:505-525).  Returns (mean, cost)."""
    X = np.asarray(X, dtype=np.float64)
    mean = np.array(mean, dtype=np.float64)

    def costfn(m):
        if g is None:
            d = pairwise(m[None, :], X, metric="dtw", r=r)[0]
        else:
            d = pairwise(m[None, :], X, metric="wdtw", r=r, g=g)[0]
        return np.mean(d) if sample_weight is None else np.average(d, weights=sample_weight)

    cost = costfn(mean)
    for _ in range(max_epoch):
        z = np.zeros(mean.shape[0])
        V = np.zeros(mean.shape[0])
        w_band = warp_size(mean.shape[0], r, X.shape[1])
        cw = None if g is None else jeong_weight(max(mean.shape[0], X.shape[1]), g)
        for i in range(X.shape[0]):
            pm, px = mapping(dtw_alignment(mean, X[i], w_band, cw))
            wgt = 1.0 if sample_weight is None else sample_weight[i]
            for m, x in zip(pm, px):
                V[m] += wgt
                z[m] += X[i, x] * wgt
        mean = z / V
        prev, cost = cost, costfn(mean)
        if abs(prev - cost) < tol:
            break
    return mean, cost


def ssg_dtw_average(X, mean, r=1.0, max_epoch=50, max_stable=5, learning_rate=0.1, decay=0.9,
                    sample_weight=None, g=None, random_state=0):
    """_ssg_dtw_average, distance/dtw.py:597-647, over the alignment matrix
    and backtrack above (dtw_alignment's warp size; wdtw_alignment's weights
    when g is given) with numpy's RandomState shuffling the series order.
    Returns (best mean, its cost)."""
    X = np.asarray(X, dtype=np.float64)
    mean = np.array(mean, dtype=np.float64)
    rs = np.random.RandomState(random_state) if not hasattr(random_state, "shuffle") \
        else random_state

    def costfn(m):
        if g is None:
            d = pairwise(m[None, :], X, metric="dtw", r=r)[0]
        else:
            d = pairwise(m[None, :], X, metric="wdtw", r=r, g=g)[0]
        return np.mean(d) if sample_weight is None else np.average(d, weights=sample_weight)

    best_mean, min_cost, n_stable = None, np.inf, 0
    order = np.arange(X.shape[0])
    z = np.empty(mean.shape[0])
    w_band = warp_size(mean.shape[0], r, X.shape[1])
    cw = None if g is None else jeong_weight(max(mean.shape[0], X.shape[1]), g)
    for epoch in range(max_epoch):
        if n_stable > max_stable:
            break
        rs.shuffle(order)
        for i, o in enumerate(order):
            z.fill(0)
            pm, px = mapping(dtw_alignment(mean, X[o], w_band, cw))
            w = 1.0 if sample_weight is None else sample_weight[o]
            for m, x in zip(pm, px):
                z[m] += mean[m] - X[o, x] * w
            mean -= learning_rate * z
            if epoch == 0:
                learning_rate = decay**i * learning_rate
        cost = costfn(mean)
        if cost < min_cost:
            min_cost, n_stable, best_mean = cost, 0, mean.copy()
        else:
            n_stable += 1
    return best_mean, min_cost


def dtw_envelop(x, r):
    """(lower, upper) of a 1-D series for an integer warp r (the deque pass)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    lower = np.empty(x.shape[0])
    upper = np.empty(x.shape[0])
    if _load().orc_dtw_envelop(x.ctypes.data, x.shape[0], int(r), lower.ctypes.data,
                               upper.ctypes.data) != 0:
        raise MemoryError("oracle allocation failed")
    return lower, upper


def lb_keogh(x, lower, upper):
    x = np.ascontiguousarray(x, dtype=np.float64)
    lower = np.ascontiguousarray(lower, dtype=np.float64)
    upper = np.ascontiguousarray(upper, dtype=np.float64)
    return _load().orc_lb_keogh(x.ctypes.data, x.shape[0], lower.ctypes.data, upper.ctypes.data)


def lb_kim(X_fit, Y):
    """DtwKimLowerBound.transform, distance/lb.py:206-312 (numpy, as the
    reference): (len(Y), len(X_fit))."""
    X_fit = np.asarray(X_fit, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)

    def dist(a, b):
        v = a - b
        return v * v

    out = np.empty((Y.shape[0], X_fit.shape[0]))
    t = X_fit.shape[-1]
    for i in range(Y.shape[0]):
        x0, x0_, y0, y0_ = X_fit[:, 0], X_fit[:, -1], Y[i, 0], Y[i, -1]
        d = dist(x0, y0) + dist(x0_, y0_)
        if t > 1:
            x1, x1_, y1, y1_ = X_fit[:, 1], X_fit[:, -2], Y[i, 1], Y[i, -2]
            d += np.minimum(dist(x1, y0), np.minimum(dist(x0, y1), dist(x1, y1)))
            d += np.minimum(dist(x1_, y1), np.minimum(dist(x0_, y1_), dist(x1_, y1_)))
            if t > 2:
                x2, x2_, y2, y2_ = X_fit[:, 2], X_fit[:, -3], Y[i, 2], Y[i, -3]
                d += np.minimum(dist(x0, y2), np.minimum(dist(x1, y2), np.minimum(
                    dist(x2, y2), np.minimum(dist(x2, y1), dist(x2, y0)))))
                d += np.minimum(dist(x0_, y2_), np.minimum(dist(x1_, y2_), np.minimum(
                    dist(x2_, y2_), np.minimum(dist(x2_, y1_), dist(x2_, y0_)))))
        out[i, :] = d
    return out


def dtw_alignment(x, y, band, weights=None):
    """_dtw_alignment's matrix (+inf off the band), weights without the row-0
    quirk."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty((x.shape[0], y.shape[0]))
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=np.float64)
    _load().orc_dtw_alignment(x.ctypes.data, x.shape[0], y.ctypes.data, y.shape[0], int(band),
                              _ptr(weights), out.ctypes.data)
    return out
