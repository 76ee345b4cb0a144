"""Differentiable DTW and DTW barycenter averaging on the GPU.

``dtw(X, Y, r)`` returns the (nx, ny) DTW distances of all pairs and
``paired_dtw(X, Y, r)`` the (n,) distances of X[i] vs Y[i]; both are
``torch.autograd`` functions whose gradients follow the optimal warping path
of wildboar's ``dtw_mapping`` backtrack (``distance/dtw.py:362-427``, ties
resolved diag > up > left):

    dD/dx_i = sum over path cells (i, j) of (x_i - y_j) / D,
    dD/dy_j = -sum over path cells (i, j) of (x_i - y_j) / D,  (0 at D = 0).

The forward records the backtrack decision of every band cell (two bits per
cell) and the backward walks them; nothing is recomputed on the CPU.

``dtw_barycenter`` averages series by gradient descent on the mean DTW
distance (Adam), the gradient-based counterpart of ``dtw_average``
(``distance/dtw.py:429-690``), which uses alignment (MM/SSG) updates.
"""

import ctypes

import torch

from . import _lib
from .distance import _check_r


def _metric(r):
    return _lib.metric_struct(_lib.DTW, _lib.FP32, _check_r(r), 0.05)


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
    def forward(ctx, X, Y, r, mode):
        X = _prep(X, "X")
        Y = _prep(Y, "Y")
        nx, tx = X.shape
        ny, ty = Y.shape
        if mode == 1 and nx != ny:
            raise ValueError("paired DTW needs the same number of x and y series")
        m = _metric(r)
        state_bytes, _ = _sizes(m, mode, nx, ny, tx, ty)
        state = torch.empty(state_bytes, dtype=torch.uint8, device=X.device)
        D = torch.empty((nx, ny) if mode == 0 else (nx,), dtype=torch.float64, device=X.device)
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().wbx_dev_dtw_train_forward(
            m, mode, X.data_ptr(), nx, X.stride(0), tx, Y.data_ptr(), ny, Y.stride(0), ty,
            D.data_ptr(), state.data_ptr(), stream))
        ctx.save_for_backward(X, Y)
        ctx.state = state
        ctx.r = r
        ctx.mode = mode
        return D

    @staticmethod
    def backward(ctx, G):
        X, Y = ctx.saved_tensors
        nx, tx = X.shape
        ny, ty = Y.shape
        m = _metric(ctx.r)
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
        return dX, dY, None, None


def dtw(X, Y, r=1.0):
    """(nx, ny) float64 DTW distances of every X row against every Y row."""
    return _DtwFunction.apply(X, Y, float(r), 0)


def paired_dtw(X, Y, r=1.0):
    """(n,) float64 DTW distances of X[i] against Y[i]."""
    return _DtwFunction.apply(X, Y, float(r), 1)


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
