"""Drop-in distance API for metric = "euclidean", "dtw" and "wdtw".

Same names, signatures, metric names, ``metric_params`` and return shapes as
wildboar's ``distance/_distance.py``:

- ``pairwise_distance``            (``_distance.py:1178-1304``)
- ``paired_distance``              (``_distance.py:1082-1175``)
- ``argmin_distance``              (``_distance.py:1307-1457``)
- ``pairwise_subsequence_distance`` (``_distance.py:543-636``)

Every computation runs in ``libwbx.so`` on the GPU; these functions only
validate, reshape and dispatch.  One keyword is added: ``precision`` --
``"fp64"`` (default) is bit-identical to the reference; ``"fp32"`` computes
cells in float32 (argmin indices and distances stay exact through a certified
float64 re-rank).
"""

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

    def __init__(self, name, kind, r=1.0, g=0.05):
        self.name, self.kind, self.r, self.g = name, kind, r, g

    @property
    def is_elastic(self):
        return self.kind != _lib.EUCLIDEAN

    def struct(self, precision):
        return _lib.metric_struct(self.kind, precision, self.r, self.g)

    def __reduce__(self):
        return _make_metric, (self.name, self.r, self.g)


def _make_metric(name, r, g):
    return _Metric(name, _lib.KINDS[name], r, g)


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


# The subset of the reference's _METRICS registry (_distance.py:185-204)
# served here; the same names resolve for subsequence metrics (:148-182).
METRICS = {"euclidean": _euclidean, "dtw": _dtw, "wdtw": _wdtw}


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
    if scale:
        raise NotImplementedError("scaled subsequence metrics are not served by this library")
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
