"""Drop-in distance API for metric = "euclidean", "dtw", "wdtw", "ddtw",
"wddtw" and "adtw" (subsequence functions: euclidean, dtw, wdtw).

Same names, signatures, metric names, ``metric_params`` and return shapes as
wildboar's ``distance/_distance.py``:

- ``pairwise_distance``            (``_distance.py:1178-1304``)
- ``paired_distance``              (``_distance.py:1082-1175``)
- ``argmin_distance``              (``_distance.py:1307-1457``)
- ``pairwise_subsequence_distance`` (``_distance.py:543-636``)
- ``subsequence_match``            (``_distance.py:732-924``)
- ``paired_subsequence_match``     (``_distance.py:927-1079``)
- ``argmin_subsequence_distance``  (``_distance.py:1636-1764``)

Every computation runs in ``libwbx.so`` on the GPU; these functions only
validate, reshape and dispatch.  One keyword is added: ``precision`` --
``"fp64"`` (default) is bit-identical to the reference; ``"fp32"`` computes
cells in float32 (argmin indices and distances stay exact through a certified
float64 re-rank).
"""

import math
import numbers
import warnings

import numpy as np

from . import _lib
from .variable_len import is_variable_length

__all__ = [
    "pairwise_distance",
    "paired_distance",
    "argmin_distance",
    "pairwise_subsequence_distance",
    "subsequence_match",
    "paired_subsequence_match",
    "argmin_subsequence_distance",
    "check_metric",
    "METRICS",
]


# --------------------------------------------------------------------- metrics
def _check_r(r):
    # DtwMetric.__init__: check_scalar(r, "r", float, min_val=0.0, max_val=1.0)
    # (distance/_elastic.pyx:3133-3135); the Cython signature converts to double.
    if isinstance(r, bool) or not isinstance(r, numbers.Real):
        raise TypeError(f"r must be an instance of float, not {type(r).__name__}.")
    r = float(r)
    if not (0.0 <= r <= 1.0):
        raise ValueError(f"r == {r}, must be >= 0.0 and <= 1.0." if r == r else "r is NaN")
    return r


def _check_g(g):
    # WeightedDtwMetric.__init__: check_scalar(g, "g", float, min_val=0.0)
    if isinstance(g, bool) or not isinstance(g, numbers.Real):
        raise TypeError(f"g must be an instance of float, not {type(g).__name__}.")
    g = float(g)
    if not g >= 0.0:
        raise ValueError(f"g == {g}, must be >= 0.0.")
    return g


class _Metric:
    """A resolved metric: the kind, its parameters and elasticity."""

    def __init__(self, name, kind, r=1.0, g=0.05, p=1.0):
        self.name, self.kind, self.r, self.g, self.p = name, kind, r, g, p

    @property
    def is_elastic(self):
        return self.kind != _lib.EUCLIDEAN

    def struct(self, precision):
        return _lib.metric_struct(self.kind, precision, self.r, self.g, self.p)

    def __reduce__(self):
        return _make_metric, (self.name, self.r, self.g, self.p)


def _make_metric(name, r, g, p=1.0):
    return _Metric(name, _lib.KINDS[name], r, g, p)


def _euclidean(**params):
    if params:
        raise TypeError(f"__init__() got an unexpected keyword argument '{next(iter(params))}'")
    return _Metric("euclidean", _lib.EUCLIDEAN)


def _dtw(r=1.0, **params):
    if params:
        raise TypeError(f"__init__() got an unexpected keyword argument '{next(iter(params))}'")
    return _Metric("dtw", _lib.DTW, _check_r(r))


def _wdtw(r=1.0, g=0.05, **params):
    if params:
        raise TypeError(f"__init__() got an unexpected keyword argument '{next(iter(params))}'")
    return _Metric("wdtw", _lib.WDTW, _check_r(r), _check_g(g))


def _real(v, name):
    # the Cython signatures take `double` (_elastic.pyx:3236, 3346, 3406)
    if isinstance(v, bool) or not isinstance(v, numbers.Real):
        raise TypeError(f"{name} must be an instance of float, not {type(v).__name__}.")
    return float(v)


def _ddtw(r=1.0, **params):
    # DerivativeDtwMetric.__init__ (_elastic.pyx:3234-3235): DtwMetric's r check
    if params:
        raise TypeError(f"__init__() got an unexpected keyword argument '{next(iter(params))}'")
    return _Metric("ddtw", _lib.DDTW, _check_r(r))


def _wddtw(r=1.0, g=0.05, **params):
    # WeightedDerivativeDtwMetric.__init__ (_elastic.pyx:3406-3408) checks r
    # only; g is any float
    if params:
        raise TypeError(f"__init__() got an unexpected keyword argument '{next(iter(params))}'")
    return _Metric("wddtw", _lib.WDDTW, _check_r(r), _real(g, "g"))


def _adtw(r=1.0, p=1.0, **params):
    # AmercingDtwMetric.__init__ (_elastic.pyx:3344-3348): check_scalar(p, float)
    if params:
        raise TypeError(f"__init__() got an unexpected keyword argument '{next(iter(params))}'")
    p = _real(p, "p")
    if not p > -np.inf:
        # -inf + inf (a move from outside the band) would be NaN
        raise ValueError(f"p == {p}, must be a number above -inf.")
    return _Metric("adtw", _lib.ADTW, _check_r(r), p=p)


# The subset of the reference's _METRICS registry (_distance.py:185-204)
# served here, by every distance and subsequence function; with scale=True
# (or a "scaled_" name) the subsequence functions serve the scaled forms of
# all six.
METRICS = {"euclidean": _euclidean, "dtw": _dtw, "wdtw": _wdtw, "ddtw": _ddtw,
           "wddtw": _wddtw, "adtw": _adtw}


def check_metric(metric):
    """Resolve a metric name (``_distance.py:231-260``)."""
    if isinstance(metric, str) and metric in METRICS:
        return METRICS[metric]
    raise ValueError(
        "unsupported metric {}, 'metric' must be a str among {}".format(metric, set(METRICS)))


def _precision(precision):
    try:
        return _lib.PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {set(_lib.PRECISIONS)}, got {precision!r}")


# ------------------------------------------------------------------ validation
def check_array(x, *, allow_3d=True, ensure_2d=False, input_name=""):
    """float64 cast + the reference's checks (``utils/validation.py:404-591``)."""
    a = np.asarray(x)
    if a.dtype == object:
        a = a.astype(np.float64)
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 0:
        raise ValueError(f"Expected 2D array, got scalar array instead:\narray={a}.")
    if ensure_2d and a.ndim == 1:
        raise ValueError(f"Expected 2D array, got 1D array instead:\narray={a}.")
    if a.ndim >= 3 and not allow_3d:
        raise ValueError(f"Found array with dim {a.ndim}. expected <= 2.")
    if a.ndim >= 4:
        raise ValueError(f"Found array with dim {a.ndim}. expected <= 3.")
    if a.ndim >= 2 and a.shape[0] < 1:
        raise ValueError(f"Found array with 0 sample(s) (shape={a.shape}) while a minimum of 1 is required.")
    if a.shape[-1] < 1:
        raise ValueError(f"Found array with 0 feature(s) (shape={a.shape}) while a minimum of 1 is required.")
    if a.ndim == 3 and a.shape[1] < 1:
        raise ValueError(f"Found array with 0 dimension(s) (shape={a.shape}) while a minimum of 1 is required.")
    name = input_name + " " if input_name else ""
    if is_variable_length(a):
        raise ValueError(f"Input {name}expected time series of equal length.")
    if np.isnan(a).any():
        raise ValueError(f"Input {name}contains NaN.")
    if np.isinf(a).any():
        raise ValueError(f"Input {name}contains infinity.")
    return np.ascontiguousarray(a)


def _ts(a):
    """(n, dims, t) C-contiguous view (``utils/validation.py:594-620``)."""
    if a.ndim == 1:
        a = a.reshape(1, 1, a.shape[0])
    elif a.ndim == 2:
        a = a.reshape(a.shape[0], 1, a.shape[1])
    return np.ascontiguousarray(a, dtype=np.float64)


def _format_return(x, y_dims, x_dims):
    """``_distance.py:514-540``."""
    if x_dims == 1 and y_dims == 1 and x.size == 1:
        return x.item()
    elif x_dims == 1 or y_dims == 1:
        return np.squeeze(x)
    return x


# --------------------------------------------------------------------- drivers
def _pairwise(x_, y_, dim, metric, precision):
    lib = _lib.load()
    m = metric.struct(precision)
    if y_ is None:
        out = np.empty((x_.shape[0], x_.shape[0]))
        _lib.check(lib.wbx_pairwise(m, x_.ctypes.data, x_.shape[0], x_.shape[1], x_.shape[2],
                                    None, 0, 0, dim, out.ctypes.data))
    else:
        out = np.empty((x_.shape[0], y_.shape[0]))
        _lib.check(lib.wbx_pairwise(m, x_.ctypes.data, x_.shape[0], x_.shape[1], x_.shape[2],
                                    y_.ctypes.data, y_.shape[0], y_.shape[2], dim,
                                    out.ctypes.data))
    return out


def _paired(x_, y_, dim, metric, precision):
    out = np.empty(x_.shape[0])
    _lib.check(_lib.load().wbx_paired(metric.struct(precision), x_.ctypes.data, y_.ctypes.data,
                                      x_.shape[0], x_.shape[1], x_.shape[2], dim,
                                      out.ctypes.data))
    return out


# ------------------------------------------------------------------ public API
def pairwise_distance(x, y=None, *, dim="mean", metric="euclidean", metric_params=None,
                      n_jobs=None, precision="fp64"):
    """Distance between every series of ``x`` and ``y`` (``_distance.py:1178-1304``).

    Returns a scalar for two 1-D inputs, ``(x_samples, y_samples)`` otherwise
    (squeezed when either input is 1-D), ``(n_dims, ...)`` for ``dim="full"``.
    With ``y=None`` the result is the symmetric ``(n, n)`` matrix with a zero
    diagonal, computed for ``i < j`` as ``D(x[i], x[j])``.
    """
    Metric = check_metric(metric)
    metric = Metric(**(metric_params if metric_params is not None else {}))
    prec = _precision(precision)
    if y is None:
        y = x
    if x is y:
        x = check_array(x)
        if x.ndim == 1:
            return 0.0
        x_ = _ts(x)
        n_dims = x_.shape[1]
        if n_dims == 1 and dim == "mean":
            dim = 0
        if dim in ("mean", "full"):
            d = [_pairwise(x_, None, k, metric, prec) for k in range(n_dims)]
            return np.mean(d, axis=0) if dim == "mean" else np.stack(d, axis=0)
        if isinstance(dim, numbers.Integral) and 0 <= dim < n_dims:
            return _pairwise(x_, None, int(dim), metric, prec)
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")

    x = check_array(x)
    y = check_array(y)
    if x.ndim != 1 and y.ndim != 1 and x.ndim != y.ndim:
        raise ValueError("x (%dD-array) and y (%dD-array) are not compatible" % (x.ndim, y.ndim))
    if x.ndim == 3 and y.ndim == 3 and x.shape[1] != y.shape[1]:
        raise ValueError("x and y must have the same number of dimensions.")
    if x.shape[-1] != y.shape[-1] and not metric.is_elastic:
        raise ValueError("Illegal n_timestep (%r != %r) for non-elastic distance measure"
                         % (x.shape[-1], y.shape[-1]))
    x_ = _ts(x)
    y_ = _ts(y)
    n_dims = x.shape[1] if x.ndim == 3 else 1
    if n_dims == 1 and dim == "mean":
        dim = 0
    if dim in ("mean", "full"):
        d = [_pairwise(x_, y_, k, metric, prec) for k in range(n_dims)]
        distances = np.mean(d, axis=0) if dim == "mean" else np.stack(d, axis=0)
    elif isinstance(dim, numbers.Integral) and 0 <= dim < n_dims:
        distances = _pairwise(x_, y_, int(dim), metric, prec)
    else:
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")
    return _format_return(distances, y.ndim, x.ndim)


def paired_distance(x, y, *, dim="mean", metric="euclidean", metric_params=None, n_jobs=None,
                    precision="fp64"):
    """Distance between ``x[i]`` and ``y[i]`` (``_distance.py:1082-1175``).

    Like the reference's driver (``_cdistance.pyx:1632-1652``) each value is
    ``D(y[i], x[i])`` -- the order matters for the asymmetric ``wdtw``.
    """
    x = check_array(x)
    y = check_array(y)
    x, y = np.broadcast_arrays(x, y)
    if x.ndim != y.ndim:
        raise ValueError("x (%dD-array) and y (%dD-array) are not compatible." % (x.ndim, y.ndim))
    if x.ndim == 3 and x.shape[1] != y.shape[1]:
        raise ValueError("x and y must have the same number of dimensions.")
    if x.ndim > 1 and y.ndim > 1 and x.shape[0] != y.shape[0]:
        raise ValueError("x and y must have the same number of samples.")
    if n_jobs is not None:
        warnings.warn("n_jobs is not yet supported.", UserWarning)
    Metric = check_metric(metric)
    metric = Metric(**(metric_params if metric_params is not None else {}))
    prec = _precision(precision)
    n_dims = x.shape[1] if x.ndim == 3 else 1
    if n_dims == 1 and dim == "mean":
        dim = 0
    x_ = _ts(x)
    y_ = _ts(y)
    if dim in ("mean", "full"):
        d = [_paired(x_, y_, k, metric, prec) for k in range(n_dims)]
        distances = np.mean(d, axis=0) if dim == "mean" else np.stack(d, axis=0)
    elif isinstance(dim, numbers.Integral) and 0 <= dim < n_dims:
        distances = _paired(x_, y_, int(dim), metric, prec)
    else:
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")
    return _format_return(distances, y.ndim, x.ndim)


def argmin_distance(x, y=None, *, dim=0, k=1, metric="euclidean", metric_params=None,
                    lower_bound=None, sorted=False, return_distance=False, n_jobs=None,
                    precision="fp64"):
    """Indices of the ``k`` nearest ``y`` rows per ``x`` row (``_distance.py:1307-1457``).

    Output order is the reference's heap order (``utils/_misc.pyx:18-107``)
    unless ``sorted=True`` (stable sort by distance).
    """
    if isinstance(k, bool) or not isinstance(k, numbers.Integral):
        raise TypeError(f"The 'k' parameter must be an int, got {k!r}.")
    if k < 1:
        raise ValueError(f"The 'k' parameter must be an int in the range [1, inf), got {k!r}.")
    if metric_params is not None and not isinstance(metric_params, dict):
        raise TypeError("The 'metric_params' parameter must be None or a dict.")
    metric = check_metric(metric)(**(metric_params if metric_params is not None else {}))
    prec = _precision(precision)
    x = _ts(check_array(x))
    y = x if y is None else _ts(check_array(y))
    if x.shape[-1] != y.shape[-1] and not metric.is_elastic:
        raise ValueError("Illegal n_timestep (%d != %d) for non-elastic distance measure."
                         % (x.shape[-1], y.shape[-1]))
    lb = None
    if lower_bound is not None:
        lb = np.atleast_2d(check_array(lower_bound))
        if x.shape[0] != lb.shape[0] or y.shape[0] != lb.shape[1]:
            raise ValueError(
                "The lower bound must be of shape (x.shape[0], y.shape[0]), got ({}, {})".format(
                    *lb.shape))
        lb = np.ascontiguousarray(lb, dtype=np.float64)
    n_dims = x.shape[1]
    k = min(k, y.shape[0])
    if not (0 <= dim < 1):
        raise ValueError(f"The parameter dim must be dim ({dim}) < n_dims ({n_dims})")
    idx = np.zeros((x.shape[0], k), dtype=np.intp)
    dist = np.zeros((x.shape[0], k))
    _lib.check(_lib.load().wbx_argmin(
        metric.struct(prec), x.ctypes.data, x.shape[0], x.shape[1], x.shape[2], y.ctypes.data,
        y.shape[0], y.shape[2], int(dim), int(k), None if lb is None else lb.ctypes.data,
        idx.ctypes.data, dist.ctypes.data))
    if sorted:
        order = np.argsort(dist, axis=1, kind="stable")
        idx = np.take_along_axis(idx, order, axis=1)
        if return_distance:
            dist = np.take_along_axis(dist, order, axis=1)
    return (idx, dist) if return_distance else idx


def _validate_subsequence(y):
    """``_distance.py:336-373``: a list of 1-D float64 arrays."""
    if len(y) == 0:
        raise ValueError("Subsequence y cannot be empty.")
    if isinstance(y, np.ndarray) and y.dtype != object:
        if y.ndim == 1:
            return [y.astype(float)]
        if y.ndim == 2:
            return list(y.astype(float))
        raise ValueError("Expected 2D array, got {}D array instead:\narray={}.\n".format(y.ndim, y))
    if any(isinstance(e, (list, tuple, np.ndarray)) or hasattr(e, "__array__") for e in y):
        return [np.array(e, dtype=np.double) for e in y]
    return [np.array(y, dtype=np.double)]


def pairwise_subsequence_distance(y, x, *, dim=0, metric="euclidean", metric_params=None,
                                  scale=False, return_index=False, n_jobs=None,
                                  precision="fp64"):
    """Minimum subsequence distance between shapelets ``y`` and series ``x``
    (``_distance.py:543-636``): ``(n_samples, n_subsequences)``, squeezed as
    the reference does; with ``return_index`` also the first window index of
    the minimum."""
    y = _validate_subsequence(y)
    x = check_array(x)
    for s in y:
        if s.shape[0] > x.shape[-1]:
            raise ValueError("Invalid subsequnce shape (%d > %d)" % (s.shape[0], x.shape[-1]))
    scaled = _scaled_name(metric, scale)
    if scaled is not None:
        return _scaled_pairwise_subsequence(y, x, scaled, metric_params, dim, return_index,
                                            precision)
    metric_ = check_metric(metric)(**(metric_params if metric_params is not None else {}))
    prec = _precision(precision)
    x_ = _ts(x)
    if not (isinstance(dim, numbers.Integral) and 0 <= dim < x_.shape[1]):
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")
    lens = np.array([s.shape[0] for s in y], dtype=np.int64)
    offs = np.zeros(len(y), dtype=np.int64)
    if len(y) > 1:
        offs[1:] = np.cumsum(lens)[:-1]
    S = np.ascontiguousarray(np.concatenate(y).astype(np.float64))
    for s in y:
        if np.isnan(s).any() or np.isinf(s).any():
            raise ValueError("Input contains NaN or infinity.")
    dist = np.empty((x_.shape[0], len(y)))
    idx = np.empty((x_.shape[0], len(y)), dtype=np.intp)
    _lib.check(_lib.load().wbx_subsequence_pairwise(
        metric_.struct(prec), S.ctypes.data, offs.ctypes.data, lens.ctypes.data, len(y),
        x_.ctypes.data, x_.shape[0], x_.shape[1], x_.shape[2], int(dim), dist.ctypes.data,
        idx.ctypes.data))
    if return_index:
        return _format_return(dist, len(y), x.ndim), _format_return(idx, len(y), x.ndim)
    return _format_return(dist, len(y), x.ndim)


# Scaled subsequence metrics served: "scaled_" + these (_distance.py:151-183);
# the registry's other scaled names raise NotImplementedError.
_SCALED_SERVED = ("euclidean", "dtw", "wdtw", "ddtw", "wddtw", "adtw")
_SCALED_MATCH_SERVED = _SCALED_SERVED
_REFERENCE_ONLY = ("lcss", "edr", "twe", "msm", "erp", "manhattan", "minkowski", "chebyshev",
                   "cosine", "angular")


def _scaled_name(metric, scale):
    """The base metric of a scaled request (``_infer_scaled_metric``,
    ``_distance.py:208-228``, applied when ``scale`` is set or the name starts
    with ``scaled_``), or None for an unscaled one."""
    if callable(metric):
        if scale:
            raise NotImplementedError("scaled callable metrics are not served by this library")
        return None
    if not isinstance(metric, str) or not (scale or metric.startswith("scaled_")):
        return None
    base = metric[len("scaled_"):] if metric.startswith("scaled_") else metric
    if base not in _SCALED_SERVED:
        if base in METRICS or base in _REFERENCE_ONLY:
            raise NotImplementedError("metric 'scaled_%s' is not served by this library" % base)
        raise ValueError("unsupported metric 'scaled_{}'".format(base))
    return base


def _pattern_stats(y):
    """Mean and std of each pattern as ``ScaledSubsequenceMetric.from_array``
    computes them (``_cdistance.pyx:453-467``): ``np.mean``, ``np.std``, and a
    std of at most 1e-13 stored as 0."""
    mean = np.array([np.mean(s) for s in y], dtype=np.float64)
    std = np.array([np.std(s) for s in y], dtype=np.float64)
    std[std <= 1e-13] = 0.0
    return mean, std


def _scaled_pairwise_subsequence(y, x, base, metric_params, dim, return_index, precision):
    """``pairwise_subsequence_distance`` with ``ScaledEuclideanSubsequenceMetric``
    (``_metric.pyx:96-160``) or ``ScaledDtwSubsequenceMetric``
    (``_elastic.pyx:1928-2203``)."""
    params = metric_params if metric_params is not None else {}
    if metric_params is not None and not isinstance(metric_params, dict):
        raise TypeError("The 'metric_params' parameter must be None or a dict.")
    metric_ = check_metric(base)(**params)
    prec = _precision(precision)
    x_ = _ts(x)
    if not (isinstance(dim, numbers.Integral) and 0 <= dim < x_.shape[1]):
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")
    for s in y:
        if np.isnan(s).any() or np.isinf(s).any():
            raise ValueError("Input contains NaN or infinity.")
    lens = np.array([s.shape[0] for s in y], dtype=np.int64)
    offs = np.zeros(len(y), dtype=np.int64)
    if len(y) > 1:
        offs[1:] = np.cumsum(lens)[:-1]
    S = np.ascontiguousarray(np.concatenate(y).astype(np.float64))
    mean, std = _pattern_stats(y)
    dist = np.empty((x_.shape[0], len(y)))
    idx = np.empty((x_.shape[0], len(y)), dtype=np.intp)
    _lib.check(_lib.load().wbx_scaled_subsequence_pairwise(
        metric_.struct(prec), S.ctypes.data, offs.ctypes.data, lens.ctypes.data,
        mean.ctypes.data, std.ctypes.data, len(y), x_.ctypes.data, x_.shape[0], x_.shape[1],
        x_.shape[2], int(dim), dist.ctypes.data, idx.ctypes.data))
    if return_index:
        return _format_return(dist, len(y), x.ndim), _format_return(idx, len(y), x.ndim)
    return _format_return(dist, len(y), x.ndim)


# ------------------------------------------------------- subsequence matching
def _is_arraylike(x):
    # sklearn.utils.validation._is_arraylike, as the reference uses it
    return hasattr(x, "__len__") or hasattr(x, "shape") or hasattr(x, "__array__")


def _safe_jagged_array(lst):
    """``utils/__init__.py:147-150``: a 1-D object array of arrays (or None)."""
    arr = np.empty(len(lst), dtype=object)
    arr[:] = lst
    return arr


def _std_below_mean(s):
    """``_distance.py:310-330``."""

    def f(x):
        return max(np.mean(x) - s * np.std(x), np.min(x))

    return f


_THRESHOLD = {"auto": _std_below_mean(2.0)}


def _resolve_threshold(threshold, max_matches, n_samples, per_sample):
    """The threshold branches of ``subsequence_match`` (``_distance.py:836-881``)
    and ``paired_subsequence_match`` (:1033-1058).  Returns (bound, max_dist,
    max_matches): ``bound`` is what the GPU scan compares against,
    ``max_dist(i, d)`` the optional host mask."""
    if threshold is None:
        threshold = np.inf
        if max_matches is None:
            max_matches = 10
    if callable(threshold):
        fn = threshold
        return np.inf, lambda _, d: d <= fn(d), max_matches
    if isinstance(threshold, str):
        if threshold not in _THRESHOLD:
            raise ValueError("The parameter threshold must be one of {}, got {!r}".format(
                set(_THRESHOLD), threshold))
        fn = _THRESHOLD[threshold]
        return np.inf, lambda _, d: d <= fn(d), max_matches
    if per_sample and _is_arraylike(threshold):
        arr = np.asarray(threshold)
        if len(arr) != n_samples:
            raise ValueError(f"threshold array length ({len(arr)}) must match "
                             f"the number of samples ({n_samples})")
        return np.inf, lambda i, d: d <= arr[i], max_matches
    if not isinstance(threshold, numbers.Real):
        kinds = "str, callable, array-like or float" if per_sample else "str, callable or float"
        raise TypeError("threshold must be %s, not %s" % (kinds, type(threshold).__qualname__))
    return float(threshold), None, max_matches


def _pack(patterns):
    lens = np.array([p.shape[0] for p in patterns], dtype=np.int64)
    offs = np.zeros(len(patterns), dtype=np.int64)
    if len(patterns) > 1:
        offs[1:] = np.cumsum(lens)[:-1]
    S = np.ascontiguousarray(np.concatenate(patterns).astype(np.float64))
    if not np.isfinite(S).all():
        raise ValueError("Input contains NaN or infinity.")
    return S, offs, lens


class _ScaledMetric:
    """A scaled subsequence metric (``scaled_euclidean`` / ``scaled_dtw``)."""

    def __init__(self, base):
        self.base = base


def _profiles(patterns, x_, dim, metric, paired, threshold=np.inf):
    """Squared cost of every window, one row per series: ``(n, t - min_len + 1)``
    (windows past a pattern's last are +inf).  Scaled metrics: the squared
    distance the reference's scaled matching records at ``threshold``, +inf
    for windows it does not record."""
    S, offs, lens = _pack(patterns)
    n, dims, t = x_.shape
    nwp = t - int(lens.min()) + 1
    prof = np.empty((n, nwp))
    if isinstance(metric, _ScaledMetric):
        mean, std = _pattern_stats(patterns)
        _lib.check(_lib.load().wbx_scaled_subsequence_profile(
            metric.base.struct(_lib.FP64), S.ctypes.data, offs.ctypes.data, lens.ctypes.data,
            mean.ctypes.data, std.ctypes.data, len(patterns), 1 if paired else 0, x_.ctypes.data,
            n, dims, t, int(dim), nwp, float(threshold), prof.ctypes.data))
        return prof
    _lib.check(_lib.load().wbx_subsequence_profile(
        metric.struct(_lib.FP64), S.ctypes.data, offs.ctypes.data, lens.ctypes.data, len(patterns),
        1 if paired else 0, x_.ctypes.data, n, dims, t, int(dim), nwp, prof.ctypes.data))
    return prof


def _matches_from_profiles(prof, threshold, n_windows):
    """The windows with cost <= threshold^2, in order, with their distances
    (``_metric.pyx:762-795`` / ``_elastic.pyx:658-698``; the early-abandoned
    sums of the reference never pass that test).  Row i has ``n_windows[i]``
    windows (``t - len + 1`` of its pattern).  None for no match
    (``_cdistance.pyx:936-953``)."""
    bound = threshold * threshold
    indices, distances = [], []
    for row, nw in zip(prof, n_windows):
        row = row[:nw]
        hit = np.flatnonzero(row <= bound)
        if hit.size == 0:
            indices.append(None)
            distances.append(None)
        else:
            indices.append(hit.astype(np.intp))
            distances.append(np.sqrt(row[hit]))
    return indices, distances


def _any_in_exclude(lst, i, exclude):
    """``_distance.py:376-397``."""
    for e in lst:
        if i - exclude < e < i + exclude:
            return True
    return False


def _exclude_trivial_matches(indices, distances, exclude):
    """``_distance.py:400-443``: greedy by distance, dropping matches within
    ``exclude`` of a kept one."""
    out_i, out_d = [], []
    for index, distance in zip(indices, distances):
        if index is None:
            out_i.append(None)
            out_d.append(None)
            continue
        order = np.argsort(distance)
        keep = np.zeros(order.size, dtype=bool)
        kept = []
        for j in range(index.size):
            if not _any_in_exclude(kept, index[order[j]], exclude):
                kept.append(index[order[j]])
                keep[order[j]] = True
        out_i.append(index[keep])
        out_d.append(distance[keep])
    return out_i, out_d


def _filter_by_max_matches(indices, distances, max_matches):
    """``_distance.py:446-477``."""
    out_i, out_d = [], []
    for index, distance in zip(indices, distances):
        if index is None:
            out_i.append(None)
            out_d.append(None)
        else:
            order = np.argsort(distance)[:max_matches]
            out_i.append(index[order])
            out_d.append(distance[order])
    return out_i, out_d


def _filter_by_max_dist(indices, distances, max_dist):
    """``_distance.py:480-511``."""
    out_i, out_d = [], []
    for sample, (index, distance) in enumerate(zip(indices, distances)):
        if index is None:
            out_i.append(None)
            out_d.append(None)
        else:
            mask = max_dist(sample, distance)
            out_i.append(index[mask])
            out_d.append(distance[mask])
    return out_i, out_d


def _subsequence_metric(metric, metric_params, scale, scaled_ok=False):
    if metric_params is not None and not isinstance(metric_params, dict):
        raise TypeError("The 'metric_params' parameter must be None or a dict.")
    params = metric_params if metric_params is not None else {}
    if scale or (isinstance(metric, str) and metric.startswith("scaled_")):
        base = _scaled_name(metric, scale)
        if not scaled_ok or base not in _SCALED_MATCH_SERVED:
            raise NotImplementedError("scaled subsequence metrics are not served here")
        return _ScaledMetric(check_metric(base)(**params))
    return check_metric(metric)(**params)


def subsequence_match(y, x, threshold=None, *, dim=0, metric="euclidean", metric_params=None,
                      scale=False, max_matches=None, exclude=None, return_distance=False,
                      n_jobs=None):
    """Start indices of the windows of each ``x`` row within ``threshold`` of
    the subsequence ``y`` (``_distance.py:732-924``).

    Without a threshold the 10 closest windows are returned, ordered by
    distance; with one, every match in order of occurrence unless
    ``max_matches`` asks for the closest few.  ``exclude`` drops trivial
    matches near a closer one.  A sample without matches gives None.
    """
    y = _validate_subsequence(y)
    if len(y) > 1:
        raise ValueError("A single subsequence expected, got %d" % len(y))
    y = y[0]
    x = check_array(x)
    if y.shape[0] > x.shape[-1]:
        raise ValueError("Invalid subsequnce shape (%d > %d)" % (y.shape[0], x.shape[-1]))
    n_dims = x.shape[1] if x.ndim == 3 else 1
    # the reference's test reads `not 0 >= dim < n_dims` (:821), which lets only
    # dim <= 0 through; negative dims index from the end there and are
    # refused here
    if not (isinstance(dim, numbers.Integral) and 0 <= dim < n_dims):
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")
    metric_ = _subsequence_metric(metric, metric_params, scale, scaled_ok=True)
    if n_jobs is not None:
        warnings.warn("n_jobs is not yet supported.", UserWarning)
    n_samples = x.shape[0] if x.ndim > 1 else 1
    bound, max_dist, max_matches = _resolve_threshold(threshold, max_matches, n_samples, True)
    if exclude is not None:
        if isinstance(exclude, numbers.Integral):
            if exclude < 0:
                raise ValueError(f"exclude == {exclude}, must be >= 0.")
        elif isinstance(exclude, numbers.Real):
            if not exclude >= 0:
                raise ValueError(f"exclude == {exclude}, must be >= 0.")
            exclude = math.ceil(y.size * exclude)
        else:
            raise TypeError("exclude must be an instance of {int, float}, not %s."
                            % type(exclude).__qualname__)
    prof = _profiles([y], _ts(x), dim, metric_, paired=False, threshold=bound)
    indices, distances = _matches_from_profiles(prof, bound,
                                                [x.shape[-1] - y.shape[0] + 1] * prof.shape[0])
    if max_dist is not None:
        indices, distances = _filter_by_max_dist(indices, distances, max_dist)
    if exclude:
        indices, distances = _exclude_trivial_matches(indices, distances, exclude)
    if max_matches:
        indices, distances = _filter_by_max_matches(indices, distances, max_matches)
    indices = _format_return(_safe_jagged_array(indices), len(y), x.ndim)
    if indices.size == 1:
        indices = indices.item()
    if return_distance:
        distances = _format_return(_safe_jagged_array(distances), len(y), x.ndim)
        if distances.size == 1:
            distances = distances.item()
        return indices, distances
    return indices


def paired_subsequence_match(y, x, threshold=None, *, dim=0, metric="euclidean",
                             metric_params=None, scale=False, max_matches=None,
                             return_distance=False, n_jobs=None):
    """The matches of the i-th subsequence in the i-th series
    (``_distance.py:927-1079``); same threshold rules as ``subsequence_match``
    except per-sample threshold arrays."""
    y = _validate_subsequence(y)
    x = check_array(x, ensure_2d=True)
    if len(y) != x.shape[0]:
        raise ValueError("x and y must have the same number of samples")
    n_dims = x.shape[1] if x.ndim == 3 else 1
    if not (isinstance(dim, numbers.Integral) and 0 <= dim < n_dims):
        raise ValueError("The parameter dim must be 0 <= dim < n_dims")
    for s in y:
        if s.shape[0] > x.shape[-1]:
            raise ValueError("invalid subsequnce shape (%d > %d)" % (s.shape[0], x.shape[-1]))
    metric_ = _subsequence_metric(metric, metric_params, scale, scaled_ok=True)
    if n_jobs is not None:
        warnings.warn("n_jobs is not yet supported.", UserWarning)
    # the reference's one-argument max_dist (:1039) would fail inside
    # _filter_by_max_dist, which passes (sample, distances); the mask here
    # takes the distances as intended
    bound, max_dist, max_matches = _resolve_threshold(threshold, max_matches, len(y), False)
    if any(s.shape[0] < 1 for s in y):
        raise ValueError("Subsequence y cannot be empty.")
    prof = _profiles(y, _ts(x), dim, metric_, paired=True, threshold=bound)
    indices, distances = _matches_from_profiles(prof, bound,
                                                [x.shape[-1] - s.shape[0] + 1 for s in y])
    if max_dist is not None:
        indices, distances = _filter_by_max_dist(indices, distances, max_dist)
    if max_matches:
        indices, distances = _filter_by_max_matches(indices, distances, max_matches)
    indices = _format_return(_safe_jagged_array(indices), len(y), x.ndim)
    if indices.size == 1:
        indices = indices.reshape(1)
    if return_distance:
        distances = _format_return(_safe_jagged_array(distances), len(y), x.ndim)
        if distances.size == 1:
            distances = distances.reshape(1)
        return indices, distances
    return indices


def argmin_subsequence_distance(y, x, *, dim=0, k=1, metric="euclidean", metric_params=None,
                                scale=False, return_distance=False, n_jobs=None):
    """The ``k`` closest windows of ``x[i]`` to subsequence ``y[i]``
    (``_distance.py:1636-1764``): indices ``(n_samples, k)`` in the
    reference's heap order (``_cdistance.pyx:1408-1434``) and, with
    ``return_distance``, their distances."""
    if isinstance(k, bool) or not isinstance(k, numbers.Integral):
        raise TypeError(f"The 'k' parameter must be an int, got {k!r}.")
    if k < 1:
        raise ValueError(f"The 'k' parameter must be an int in the range [1, inf), got {k!r}.")
    if isinstance(y, np.ndarray) and y.dtype == float:
        y2 = check_array(y, allow_3d=False, ensure_2d=True)
        patterns = list(y2)
    else:
        patterns = []
        for s in y:
            s = np.asarray(s, dtype=np.float64)
            if s.ndim > 1:
                raise ValueError("shapelet must be 1d-array")
            patterns.append(s)
    if len(patterns) == 0:
        raise ValueError("Subsequence y cannot be empty.")
    lens = np.array([p.shape[0] for p in patterns], dtype=np.int64)
    if (lens < 1).any():
        raise ValueError("Subsequence y cannot be empty.")
    max_len = int(lens.max())
    x = np.asarray(x)
    if x.ndim == 1:
        x = np.broadcast_to(x, shape=(len(patterns), x.shape[0]))
    x_ = _ts(check_array(x, ensure_2d=True))
    if not max_len <= x_.shape[2]:
        raise ValueError("the longest subsequence must be shorter than samples.")
    if len(patterns) != x_.shape[0]:
        raise ValueError("both arrays must have the same number of samples.")
    if not (isinstance(dim, numbers.Integral) and 0 <= dim < x_.shape[1]):
        raise ValueError(f"The parameter dim must be dim ({dim}) < n_dims ({x_.shape[1]})")
    if not k <= (x_.shape[2] - max_len + 1):
        raise ValueError("k must be less x.shape[-1] - y.shape[-1] + 1.")
    if isinstance(metric, str) and (scale or metric.startswith("scaled_")):
        # _distance.py:1738-1756: "scaled_" + metric or scale=True runs the
        # plain metric on z-normalised buffers (_ScaledArgminSubsequenceDistance)
        if metric_params is not None and not isinstance(metric_params, dict):
            raise TypeError("The 'metric_params' parameter must be None or a dict.")
        base = metric[len("scaled_"):] if metric.startswith("scaled_") else metric
        metric_ = check_metric(base)(**(metric_params if metric_params is not None else {}))
        S, offs, lens = _pack(patterns)
        n = x_.shape[0]
        idx = np.empty((n, k), dtype=np.intp)
        dist = np.empty((n, k))
        _lib.check(_lib.load().wbx_scaled_argmin_subsequence(
            metric_.struct(_lib.FP64), S.ctypes.data, offs.ctypes.data, lens.ctypes.data, n,
            x_.ctypes.data, x_.shape[1], x_.shape[2], int(dim), int(k), idx.ctypes.data,
            dist.ctypes.data))
        return (idx, dist) if return_distance else idx
    metric_ = _subsequence_metric(metric, metric_params, scale)
    S, offs, lens = _pack(patterns)
    n = x_.shape[0]
    idx = np.empty((n, k), dtype=np.intp)
    dist = np.empty((n, k))
    _lib.check(_lib.load().wbx_argmin_subsequence(
        metric_.struct(_lib.FP64), S.ctypes.data, offs.ctypes.data, lens.ctypes.data, n,
        x_.ctypes.data, x_.shape[1], x_.shape[2], int(dim), int(k), idx.ctypes.data,
        dist.ctypes.data))
    return (idx, dist) if return_distance else idx
