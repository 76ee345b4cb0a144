"""Differentiable DTW and DTW barycenter averaging on the GPU.

``dtw(X, Y, r)`` returns the (nx, ny) DTW distances of all pairs and
``paired_dtw(X, Y, r)`` the (n,) distances of X[i] vs Y[i]; both are
``torch.autograd`` functions whose gradients follow the optimal warping path
of wildboar's ``dtw_mapping`` backtrack (``distance/dtw.py:362-427``, ties
resolved diag > up > left):

    dD/dx_i = sum over path cells (i, j) of (x_i - y_j) / D,
    dD/dy_j = -sum over path cells (i, j) of (x_i - y_j) / D,  (0 at D = 0).

For DTW bands up to 112 slots the forward records the backtrack decision of
every band cell (two bits per cell) and the backward walks them.  WDTW
(``metric="wdtw"``, the cell's weight multiplies its term), wider bands, and
calls with ``checkpoint=True`` keep only the cost rows at checkpoints 32
anti-diagonal steps apart; the backward recomputes each band of 32 steps from
its checkpoint, with its decisions, and walks the path through it.  Nothing
is recomputed on the CPU.

``dtw_barycenter`` averages series by gradient descent on the mean DTW
distance (Adam), the gradient-based counterpart of ``dtw_average``
(``distance/dtw.py:429-690``), which uses alignment (MM/SSG) updates.

``dtw_mapping`` and ``dtw_average`` (majorize-minimize) mirror the
reference's functions of the same names with float64 paths on the GPU.
"""

import ctypes

import torch

from . import _lib
from .distance import _check_g, _check_r


def _metric(r, metric="dtw", g=0.05):
    if metric not in ("dtw", "wdtw"):
        raise ValueError(f"differentiable DTW supports metric 'dtw' or 'wdtw', got {metric!r}")
    return _lib.metric_struct(_lib.KINDS[metric], _lib.FP32, _check_r(r),
                              _check_g(g) if metric == "wdtw" else 0.05)


def _prep(t, name):
    if not t.is_cuda or t.dim() != 2:
        raise ValueError(f"{name} must be a 2-D device tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32")
    return t if t.stride(1) == 1 else t.contiguous()


def _sizes(m, mode, nx, ny, tx, ty):
    sb = ctypes.c_int64()
    cb = ctypes.c_int64()
    _lib.check(_lib.load().wbx_dtw_train_sizes(m, mode, nx, ny, tx, ty, ctypes.byref(sb),
                                               ctypes.byref(cb)))
    return sb.value, cb.value


class _DtwFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, Y, r, mode, metric, g):
        X = _prep(X, "X")
        Y = _prep(Y, "Y")
        nx, tx = X.shape
        ny, ty = Y.shape
        if (mode & 1) and nx != ny:
            raise ValueError("paired DTW needs the same number of x and y series")
        m = _metric(r, metric, g)
        state_bytes, _ = _sizes(m, mode, nx, ny, tx, ty)
        state = torch.empty(state_bytes, dtype=torch.uint8, device=X.device)
        D = torch.empty((nx,) if mode & 1 else (nx, ny), dtype=torch.float64, device=X.device)
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().wbx_dev_dtw_train_forward(
            m, mode, X.data_ptr(), nx, X.stride(0), tx, Y.data_ptr(), ny, Y.stride(0), ty,
            D.data_ptr(), state.data_ptr(), stream))
        ctx.save_for_backward(X, Y)
        ctx.state = state
        ctx.r = r
        ctx.mode = mode
        ctx.metric = metric
        ctx.g = g
        return D

    @staticmethod
    def backward(ctx, G):
        X, Y = ctx.saved_tensors
        nx, tx = X.shape
        ny, ty = Y.shape
        m = _metric(ctx.r, ctx.metric, ctx.g)
        _, scratch_bytes = _sizes(m, ctx.mode, nx, ny, tx, ty)
        scratch = torch.empty(scratch_bytes, dtype=torch.uint8, device=X.device)
        G = G.to(torch.float64).contiguous()
        dX = torch.empty_like(X)
        dY = torch.empty_like(Y)
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().wbx_dev_dtw_train_backward(
            m, ctx.mode, X.data_ptr(), nx, X.stride(0), tx, Y.data_ptr(), ny, Y.stride(0), ty,
            G.data_ptr(), ctx.state.data_ptr(), dX.data_ptr(), dX.stride(0), dY.data_ptr(),
            dY.stride(0), scratch.data_ptr(), stream))
        ctx.state = None
        return dX, dY, None, None, None, None


def dtw_state_bytes(nx, ny, tx, ty, r=1.0, *, metric="dtw", g=0.05, checkpoint=False,
                    paired=False):
    """(state, scratch) bytes of one differentiable call: what the forward
    keeps for the backward, and the backward's working memory."""
    mode = (1 if paired else 0) | (2 if checkpoint else 0)
    return _sizes(_metric(r, metric, g), mode, nx, ny, tx, ty)


def dtw(X, Y, r=1.0, *, metric="dtw", g=0.05, checkpoint=False):
    """(nx, ny) float64 DTW (or WDTW, ``metric="wdtw"`` with weight ``g``)
    distances of every X row against every Y row.

    ``checkpoint=True`` keeps the cost rows at anti-diagonal band checkpoints
    and recomputes each band in the backward (memory ~ B (t + n) / 32 floats
    per pair) instead of recording two decision bits per band cell; WDTW and
    bands wider than 112 slots always use it.
    """
    return _DtwFunction.apply(X, Y, float(r), 2 if checkpoint else 0, metric, float(g))


def paired_dtw(X, Y, r=1.0, *, metric="dtw", g=0.05, checkpoint=False):
    """(n,) float64 DTW (or WDTW) distances of X[i] against Y[i]; see ``dtw``."""
    return _DtwFunction.apply(X, Y, float(r), 3 if checkpoint else 1, metric, float(g))


def dtw_barycenter(X, r=1.0, *, steps=200, lr=1e-2, init=None, return_history=False):
    """Barycenter of the rows of X under DTW, by Adam on mean_i DTW(mean, X_i).

    X is an (n, t) float32 device tensor; init (default X[0]) the starting mean.
    Returns the mean (t,) and, with return_history, the loss per step.
    """
    X = _prep(X, "X")
    mean = (X[0] if init is None else init).detach().clone().to(X.device, torch.float32)
    mean.requires_grad_(True)
    opt = torch.optim.Adam([mean], lr=lr)
    history = []
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = dtw(mean.unsqueeze(0), X, r).mean()
        loss.backward()
        opt.step()
        if return_history:
            history.append(loss.detach())
    mean = mean.detach()
    if return_history:
        return mean, torch.stack(history).cpu().numpy()
    return mean


# ------------------------------------------------- alignment paths and MM DBA
def _warp_size(x_size, r, y_size=0):
    """_compute_warp_size (distance/dtw.py:38-40)."""
    import math
    import numbers

    if isinstance(r, bool) or not isinstance(r, numbers.Real) or not (0 <= r <= 1):
        raise ValueError(f"r == {r!r}, must be a float in [0, 1].")
    return max(math.floor(max(x_size, y_size) * r), 1)


def _series_1d(a, name):
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    if a.ndim > 1 and a.size == max(a.shape):
        a = a.ravel()
    if a.ndim != 1 or a.shape[0] < 1:
        raise ValueError(f"{name} must be a non-empty 1-D array, got shape {a.shape}")
    if not np.isfinite(a).all():
        raise ValueError(f"Input {name} contains NaN or infinity.")
    return np.ascontiguousarray(a)


def _path_rows(x_d, y_d, w, weights_d=None):
    """Device (lx, 2) int32 column ranges of the optimal path, per row of x;
    weights_d: w[|i - j|] per cell (wdtw_alignment), or None."""
    rows = torch.empty((x_d.shape[-1], 2), dtype=torch.int32, device=x_d.device)
    with torch.cuda.device(x_d.device):
        _lib.check(_lib.load().wbx_dev_dtw_path(
            x_d.data_ptr(), 0, x_d.shape[-1], y_d.data_ptr(), 0, y_d.shape[-1], 1, int(w),
            None if weights_d is None else weights_d.data_ptr(), rows.data_ptr(),
            torch.cuda.current_stream().cuda_stream))
    return rows


def dtw_mapping(x=None, y=None, *, alignment=None, r=1, return_index=False):
    """Boolean (len(x), len(y)) indicator of the optimal warping path
    (``distance/dtw.py:362-427``), computed on the GPU in float64: the cost
    cells of ``dtw_alignment`` and the backtrack that prefers the diagonal,
    then up, then left move.  With ``return_index`` also ``indicator.nonzero()``.
    ``alignment`` (a precomputed matrix) is not served here."""
    import numpy as np

    if alignment is not None:
        raise NotImplementedError("dtw_mapping from a precomputed alignment is not served")
    if x is None or y is None:
        raise ValueError("if alignment=None, neither x or y can be None")
    x = _series_1d(x, "x")
    y = _series_1d(y, "y")
    w = _warp_size(x.shape[0], r, y_size=y.shape[0])
    dev = torch.device("cuda", torch.cuda.current_device())
    rows = _path_rows(torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev), w).cpu().numpy()
    indicator = np.zeros((x.shape[0], y.shape[0]), dtype=bool)
    for i, (lo, hi) in enumerate(rows):
        indicator[i, lo:hi + 1] = True
    if return_index:
        return indicator, indicator.nonzero()
    return indicator


def dtw_average(X, *, r=1.0, g=None, sample_weight=None, init="random", method="mm",
                max_stable=5, learning_rate=0.1, decay=0.9, tol=1e-5, max_epoch=50,
                return_cost=False, verbose=False, random_state=None):
    """DTW barycenter by majorize-minimize (``distance/dtw.py:429-690``,
    ``_mm_dtw_average`` :640-666): each epoch aligns the mean with every
    series on the GPU (float64 paths), sets each mean value to the average of
    the values aligned to it -- in the reference's order of additions -- and
    stops when the mean DTW distance changes by less than ``tol``.  Results
    equal the reference's.  With ``g`` the paths are those of
    ``wdtw_alignment`` and the cost is the mean WDTW distance.  ``"ssg"`` is
    the stochastic subgradient mean (``_ssg_dtw_average`` :597-647): series in
    the order ``random_state`` shuffles them, one GPU path and in-place step
    per series, the epoch-0 learning-rate decay as the reference compounds it,
    and the lowest-cost mean kept."""
    import numpy as np
    from sklearn.utils import check_random_state
    from sklearn.utils.validation import _check_sample_weight

    from . import distance as wd

    X = wd.check_array(X, ensure_2d=True, allow_3d=False)
    if X.shape[0] < 2:
        raise ValueError(f"Found array with {X.shape[0]} sample(s) (shape={X.shape}) while a "
                         "minimum of 2 is required.")
    _warp_size(X.shape[1], r)
    random_state = check_random_state(random_state)
    if isinstance(init, str) and init == "random":
        mean = X[random_state.randint(X.shape[0])].copy()
    elif hasattr(init, "__len__") or hasattr(init, "__array__"):
        mean = _series_1d(np.array(init, copy=True), "init").copy()
    else:
        raise ValueError("init must be array-like or 'random', not %r" % type(init).__qualname__)
    if sample_weight is not None:
        sample_weight = np.ascontiguousarray(_check_sample_weight(sample_weight, X),
                                             dtype=np.float64)
    if g is not None:
        import numbers

        if isinstance(g, bool) or not isinstance(g, numbers.Real) or not g > 0:
            raise ValueError(f"g == {g!r}, must be a float in (0, inf).")
    if method not in ("mm", "ssg"):
        raise ValueError("method must be 'mm' or 'ssg', got %r" % method)

    dev = torch.device("cuda", torch.cuda.current_device())
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(dev)
    wd_ = None if sample_weight is None else torch.from_numpy(sample_weight).to(dev)
    lib = _lib.load()

    metric, params = ("dtw", {"r": r}) if g is None else ("wdtw", {"r": r, "g": g})

    def costfn(mean):
        d = wd.pairwise_distance(mean, X, metric=metric, metric_params=params)
        return np.mean(d) if sample_weight is None else np.average(d, weights=sample_weight)

    w = _warp_size(mean.shape[0], r, y_size=X.shape[1])
    cell_w = None
    if g is not None:
        from .counterfactual import jeong_weight

        cell_w = torch.from_numpy(jeong_weight(max(mean.shape[0], X.shape[1]), g)).to(dev)
    if method == "ssg":
        md = torch.from_numpy(np.ascontiguousarray(mean, dtype=np.float64)).to(dev)
        order = np.arange(X.shape[0])
        best_mean, min_cost, n_stable = None, np.inf, 0
        stream = torch.cuda.current_stream().cuda_stream
        for epoch in range(max_epoch):
            if n_stable > max_stable:
                if verbose:
                    print(f"Completed at epoch={epoch} with cost={min_cost}.")
                break
            random_state.shuffle(order)
            for i, o in enumerate(order):
                rows = _path_rows(md, Xd[o], w, cell_w)
                wgt = 1.0 if sample_weight is None else float(sample_weight[o])
                with torch.cuda.device(dev):
                    _lib.check(lib.wbx_dev_dtw_ssg_update(
                        md.data_ptr(), md.shape[0], rows.data_ptr(), Xd[o].data_ptr(),
                        Xd.shape[1], wgt, float(learning_rate), stream))
                if epoch == 0:
                    learning_rate = decay**i * learning_rate
            mean = md.cpu().numpy()
            cost = costfn(mean)
            if cost < min_cost:
                if verbose:
                    print(f"New min cost={cost} at epoch={epoch} with "
                          f"learning_rate={learning_rate}.")
                min_cost, n_stable, best_mean = cost, 0, mean.copy()
            else:
                n_stable += 1
        return (best_mean, min_cost) if return_cost else best_mean
    cost = costfn(mean)
    for epoch in range(max_epoch):
        md = torch.from_numpy(mean).to(dev)
        out = torch.empty_like(md)
        with torch.cuda.device(dev):
            _lib.check(lib.wbx_dev_dtw_mm_update(
                md.data_ptr(), md.shape[0], Xd.data_ptr(), Xd.stride(0), Xd.shape[1],
                Xd.shape[0], int(w), None if cell_w is None else cell_w.data_ptr(),
                None if wd_ is None else wd_.data_ptr(), out.data_ptr(),
                torch.cuda.current_stream().cuda_stream))
        mean = out.cpu().numpy()
        prev_cost, cost = cost, costfn(mean)
        if abs(prev_cost - cost) < tol:
            if verbose:
                print(f"Complete at epoch={epoch} with cost={cost}.")
            break
    return (mean, cost) if return_cost else mean


def dtw_envelop(x, *, r=1.0):
    """(lower, upper) warping envelope of a series (``distance/dtw.py:155-184``),
    on the GPU."""
    from .lb import _envelope, envelope_warp

    x = _series_1d(x, "x")
    dev = torch.device("cuda", torch.cuda.current_device())
    lower, upper = _envelope(torch.from_numpy(x).to(dev)[None, :], envelope_warp(x.shape[0], r))
    return lower[0].cpu().numpy(), upper[0].cpu().numpy()


def dtw_lb_keogh(x, y=None, *, lower=None, upper=None, r=1.0):
    """(LB_Keogh, per-timestep terms) of x against the envelope of y or the
    given one (``distance/dtw.py:187-243``)."""
    import numpy as np

    x = _series_1d(x, "x")
    if y is not None:
        y = _series_1d(y, "y")
        if y.shape[0] != x.shape[0]:
            raise ValueError("x (%d) and y (%d) must have the same number of timesteps"
                             % (x.shape[0], y.shape[0]))
        lower, upper = dtw_envelop(y, r=r)
    elif lower is None or upper is None:
        raise ValueError("both y, lower and upper can't be None")
    lower = _series_1d(lower, "lower")
    upper = _series_1d(upper, "upper")
    if lower.shape[0] != upper.shape[0] or lower.shape[0] != x.shape[0]:
        raise ValueError("lower (%d), upper (%d) and x (%d) have the same number of timesteps"
                         % (lower.shape[0], upper.shape[0], x.shape[0]))
    from .lb import _lb_keogh

    dev = torch.device("cuda", torch.cuda.current_device())
    d = _lb_keogh(torch.from_numpy(x).to(dev)[None, :], torch.from_numpy(lower).to(dev)[None, :],
                  torch.from_numpy(upper).to(dev)[None, :])
    # the per-timestep terms of the bound, as cumulative_bound's cb
    cb = np.where(x > upper, (x - upper) ** 2, np.where(x < lower, (x - lower) ** 2, 0.0))
    return float(d[0, 0].item()), cb
