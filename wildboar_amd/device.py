"""Device-tensor entry points: torch CUDA (ROCm) tensors in, torch tensors out.

Each function calls the matching ``wbx_dev_*`` C entry point with the tensors'
device pointers and enqueues the work on torch's current stream, so results
are ordered with surrounding torch work.  Inputs are (n, t) float32 or float64
tensors whose last dimension is contiguous.  No host copies, no CPU fallback.
"""

import numpy as np
import torch

from . import _lib
from .distance import _check_g, _check_r


def _dtype(t):
    if t.dtype == torch.float32:
        return _lib.F32
    if t.dtype == torch.float64:
        return _lib.F64
    raise TypeError(f"expected float32 or float64 tensor, got {t.dtype}")


def _rows(t, name):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D (n, t), got shape {tuple(t.shape)}")
    if t.stride(1) != 1:
        t = t.contiguous()
    return t


def _metric(metric, r, g, precision, p=1.0):
    kind = _lib.KINDS[metric]
    prec = _lib.PRECISIONS[precision]
    if kind != _lib.EUCLIDEAN:
        r = _check_r(r)
    if kind == _lib.WDTW:
        g = _check_g(g)
    return _lib.metric_struct(kind, prec, r, g, p)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def pairwise(X, Y=None, *, metric="dtw", r=1.0, g=0.05, precision="fp32", out=None, p=1.0):
    """(nx, ny) float64 distances; ``Y=None`` gives the symmetric singleton matrix.
    ``p``: the amercing penalty of metric="adtw"."""
    X = _rows(X, "X")
    if Y is not None:
        Y = _rows(Y, "Y")
        if Y.dtype != X.dtype:
            raise TypeError("X and Y must have the same dtype")
    nx = X.shape[0]
    ny = nx if Y is None else Y.shape[0]
    if out is None:
        out = torch.empty((nx, ny), dtype=torch.float64, device=X.device)
    m = _metric(metric, r, g, precision, p)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_pairwise(
            m, _dtype(X), X.data_ptr(), nx, X.stride(0), X.shape[1],
            None if Y is None else Y.data_ptr(), ny, 0 if Y is None else Y.stride(0),
            X.shape[1] if Y is None else Y.shape[1], out.data_ptr(), _stream()))
    return out


def paired(X, Y, *, metric="dtw", r=1.0, g=0.05, precision="fp32", out=None):
    """out[i] = D(Y[i], X[i]) (the reference's paired orientation)."""
    X = _rows(X, "X")
    Y = _rows(Y, "Y")
    if X.shape != Y.shape:
        raise ValueError("X and Y must have the same shape")
    if out is None:
        out = torch.empty(X.shape[0], dtype=torch.float64, device=X.device)
    m = _metric(metric, r, g, precision)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_paired(
            m, _dtype(X), X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0), X.shape[0],
            X.shape[1], out.data_ptr(), _stream()))
    return out


def argmin(X, Y, k=1, *, metric="dtw", r=1.0, g=0.05, precision="fp32", lower_bound=None):
    """(idx int64 (nx, k), dist float64 (nx, k)) in the reference's heap order."""
    X = _rows(X, "X")
    Y = _rows(Y, "Y")
    k = min(int(k), Y.shape[0])
    idx = torch.empty((X.shape[0], k), dtype=torch.int64, device=X.device)
    dist = torch.empty((X.shape[0], k), dtype=torch.float64, device=X.device)
    lb = None
    if lower_bound is not None:
        lb = lower_bound.to(device=X.device, dtype=torch.float64).contiguous()
    m = _metric(metric, r, g, precision)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_argmin(
            m, _dtype(X), X.data_ptr(), X.shape[0], X.stride(0), X.shape[1], Y.data_ptr(),
            Y.shape[0], Y.stride(0), Y.shape[1], k, None if lb is None else lb.data_ptr(),
            idx.data_ptr(), dist.data_ptr(), _stream()))
    return idx, dist


def pack_shapelets(shapelets, device=None, dtype=None):
    """Pack a list of 1-D shapelet tensors into (values, offsets, lengths) on
    the device: the layout the subsequence entry point reads.  Packing once
    lets repeated scans skip the per-shapelet copies."""
    if not isinstance(shapelets, (list, tuple)) or not shapelets:
        raise TypeError("shapelets must be a non-empty list of 1-D tensors")
    device = device or shapelets[0].device
    dtype = dtype or shapelets[0].dtype
    lens = torch.tensor([int(s.shape[0]) for s in shapelets], dtype=torch.int64)
    packed = torch.cat([s.reshape(-1).to(device=device, dtype=dtype) for s in shapelets])
    offs = torch.zeros_like(lens)
    offs[1:] = torch.cumsum(lens, 0)[:-1]
    return packed, offs.to(device), lens.to(device)


def subsequence(shapelets, X, *, metric="dtw", r=1.0, g=0.05, precision="fp32"):
    """(dist (n, n_shapelets) float64, idx (n, n_shapelets) int64).

    shapelets: a list of 1-D tensors, or the (values, offsets, lengths)
    triple of pack_shapelets."""
    X = _rows(X, "X")
    if isinstance(shapelets, tuple) and len(shapelets) == 3 and torch.is_tensor(shapelets[0]):
        packed, offs, lens = shapelets
        if packed.dtype != X.dtype or packed.device != X.device:
            raise TypeError("packed shapelets must have the dtype and device of X")
    else:
        packed, offs, lens = pack_shapelets(shapelets, X.device, X.dtype)
    n, ns = X.shape[0], int(lens.shape[0])
    dist = torch.empty((n, ns), dtype=torch.float64, device=X.device)
    idx = torch.empty((n, ns), dtype=torch.int64, device=X.device)
    m = _metric(metric, r, g, precision)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_subsequence_pairwise(
            m, _dtype(X), packed.data_ptr(), offs.data_ptr(), lens.data_ptr(), ns, X.data_ptr(),
            n, X.stride(0), X.shape[1], dist.data_ptr(), idx.data_ptr(), _stream()))
    return dist, idx


def scaled_subsequence(shapelets, X, *, metric="dtw", r=1.0, precision="fp64"):
    """pairwise_subsequence_distance(scale=True) on device tensors: (dist, idx)
    of ScaledEuclideanSubsequenceMetric (metric="euclidean") or
    ScaledDtwSubsequenceMetric (metric="dtw", band fraction r).  shapelets: a
    list of 1-D tensors; their mean / std are taken on the host with numpy as
    the reference's from_array takes them (std <= 1e-13 -> 0)."""
    X = _rows(X, "X")
    if not isinstance(shapelets, (list, tuple)) or not shapelets:
        raise TypeError("shapelets must be a non-empty list of 1-D tensors")
    host = [s.detach().reshape(-1).cpu().double().numpy() for s in shapelets]
    mean = torch.tensor([float(np.mean(h)) for h in host], dtype=torch.float64)
    std = torch.tensor([float(np.std(h)) for h in host], dtype=torch.float64)
    std[std <= 1e-13] = 0.0
    packed, offs, lens = pack_shapelets(shapelets, X.device, X.dtype)
    mean, std = mean.to(X.device), std.to(X.device)
    n, ns = X.shape[0], len(shapelets)
    dist = torch.empty((n, ns), dtype=torch.float64, device=X.device)
    idx = torch.empty((n, ns), dtype=torch.int64, device=X.device)
    m = _metric(metric, r, 0.05, precision)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_scaled_subsequence_pairwise(
            m, _dtype(X), packed.data_ptr(), offs.data_ptr(), lens.data_ptr(), mean.data_ptr(),
            std.data_ptr(), ns, X.data_ptr(), n, X.stride(0), X.shape[1], dist.data_ptr(),
            idx.data_ptr(), _stream()))
    return dist, idx


def _patterns(shapelets, X):
    if isinstance(shapelets, tuple) and len(shapelets) == 3 and torch.is_tensor(shapelets[0]):
        packed, offs, lens = shapelets
        if packed.dtype != X.dtype or packed.device != X.device:
            raise TypeError("packed shapelets must have the dtype and device of X")
        return packed, offs, lens
    return pack_shapelets(shapelets, X.device, X.dtype)


def subsequence_profile(shapelets, X, *, paired=False, metric="dtw", r=1.0, g=0.05):
    """Squared cost of every window: (n, t - min_len + 1) float64, exact in
    float64 whatever X's dtype.  Row q scores pattern ``q if paired else 0``
    against X[q]; windows past a pattern's last are +inf."""
    X = _rows(X, "X")
    packed, offs, lens = _patterns(shapelets, X)
    n, t, ns = X.shape[0], X.shape[1], int(lens.shape[0])
    nwp = t - int(lens.min().item()) + 1
    prof = torch.empty((n, max(nwp, 1)), dtype=torch.float64, device=X.device)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_subsequence_profile(
            _metric(metric, r, g, "fp64"), _dtype(X), packed.data_ptr(), offs.data_ptr(),
            lens.data_ptr(), ns, 1 if paired else 0, X.data_ptr(), n, X.stride(0), t, nwp,
            prof.data_ptr(), _stream()))
    return prof


def argmin_subsequence(shapelets, X, k=1, *, metric="dtw", r=1.0, g=0.05):
    """(idx (n, k) int64, dist (n, k) float64): the k closest windows of X[q]
    to pattern q, in the heap order of argmin_subsequence_distance."""
    X = _rows(X, "X")
    packed, offs, lens = _patterns(shapelets, X)
    n, t = X.shape[0], X.shape[1]
    idx = torch.empty((n, k), dtype=torch.int64, device=X.device)
    dist = torch.empty((n, k), dtype=torch.float64, device=X.device)
    with torch.cuda.device(X.device):
        _lib.check(_lib.load().wbx_dev_argmin_subsequence(
            _metric(metric, r, g, "fp64"), _dtype(X), packed.data_ptr(), offs.data_ptr(),
            lens.data_ptr(), X.data_ptr(), n, X.stride(0), t, int(k), idx.data_ptr(),
            dist.data_ptr(), _stream()))
    return idx, dist
