"""DTW lower bounds on the GPU: mirrors of wildboar's ``distance/lb.py``
``DtwKeoghLowerBound`` (:314-431) and ``DtwKimLowerBound`` (:198-312).

The envelopes (``dtw_envelop``), the LB_Keogh sums and the LB_Kim terms are
computed by ``libwbx.so`` in float64 with the reference's order of
operations; ``transform`` returns the same (n_queries, n_fitted) matrix.
Univariate data (the reference reads dimension 0 of a 3-D array).
"""

import numbers

import numpy as np
import torch
from sklearn.base import BaseEstimator, TransformerMixin
from sklearn.utils.validation import check_is_fitted

from . import _lib
from . import distance as wd

__all__ = ["DtwKeoghLowerBound", "DtwKimLowerBound"]


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _rows(X):
    """(n, t) float64 rows of dimension 0, and the device copy."""
    X = wd.check_array(X, allow_3d=True, ensure_2d=True)
    if X.ndim == 3:
        X = np.ascontiguousarray(X[:, 0, :])
    dev = torch.device("cuda", torch.cuda.current_device())
    return X, torch.from_numpy(np.ascontiguousarray(X)).to(dev)


def _envelope(Xd, r_warp):
    n, t = Xd.shape
    lower = torch.empty_like(Xd)
    upper = torch.empty_like(Xd)
    _lib.check(_lib.load().wbx_dev_dtw_envelope(Xd.data_ptr(), n, Xd.stride(0), t, int(r_warp),
                                                lower.data_ptr(), upper.data_ptr(), _stream()))
    return lower, upper


def _lb_keogh(Xd, lower, upper):
    out = torch.empty((Xd.shape[0], lower.shape[0]), dtype=torch.float64, device=Xd.device)
    _lib.check(_lib.load().wbx_dev_lb_keogh(Xd.data_ptr(), Xd.shape[0], Xd.stride(0),
                                            lower.data_ptr(), upper.data_ptr(), lower.shape[0],
                                            Xd.shape[1], out.data_ptr(), _stream()))
    return out


def envelope_warp(n_timesteps, r):
    """dtw_envelop's warp (distance/dtw.py:179-183): max(floor(t r), 1), minus
    one when it reaches t."""
    from .dtw import _warp_size

    w = _warp_size(n_timesteps, r)
    return w - 1 if w == n_timesteps else w


class DtwKeoghLowerBound(TransformerMixin, BaseEstimator):
    """LB_Keogh of every query against every fitted series
    (``distance/lb.py:314-431``): ``kind`` "left" bounds the query by the
    fitted series' envelopes, "right" the fitted series by the query's,
    "both" takes the larger."""

    def __init__(self, r=1.0, *, kind="both"):
        self.r = r
        self.kind = kind

    def _validate_params(self):
        if isinstance(self.r, bool) or not isinstance(self.r, numbers.Real) or not (
                0 <= self.r <= 1):
            raise ValueError(f"The 'r' parameter must be a float in the range [0, 1], "
                             f"got {self.r!r}.")
        if self.kind not in ("both", "left", "right"):
            raise ValueError(f"The 'kind' parameter must be a str among "
                             f"{{'both', 'left', 'right'}}, got {self.kind!r}.")

    def fit(self, X, y=None):
        self._validate_params()
        self.X_, Xd = _rows(X)
        lower, upper = _envelope(Xd, envelope_warp(self.X_.shape[1], self.r))
        self.lower_ = lower.cpu().numpy()
        self.upper_ = upper.cpu().numpy()
        self._dev = (Xd, lower, upper)
        return self

    def transform(self, X):
        check_is_fitted(self, "X_")
        Y, Yd = _rows(X)
        if Y.shape[1] != self.X_.shape[1]:
            raise ValueError(f"X has {Y.shape[1]} timesteps, but DtwKeoghLowerBound is "
                             f"expecting {self.X_.shape[1]} timesteps as input.")
        Xd, lower, upper = self._dev
        if Xd.device != Yd.device:
            Xd, lower, upper = Xd.to(Yd.device), lower.to(Yd.device), upper.to(Yd.device)
        out = np.full((Y.shape[0], self.X_.shape[0]), -np.inf)
        if self.kind in ("both", "left"):
            out = _lb_keogh(Yd, lower, upper).cpu().numpy()
        if self.kind in ("both", "right"):
            ly, uy = _envelope(Yd, envelope_warp(Y.shape[1], self.r))
            right = _lb_keogh(Xd, ly, uy).cpu().numpy().T
            out = np.maximum(out, right)
        return out


class DtwKimLowerBound(TransformerMixin, BaseEstimator):
    """LB_Kim from the first and last three points (``distance/lb.py:198-312``);
    returns the squared bound, as the reference does."""

    def fit(self, X, y=None):
        self.X_, self._Xd = _rows(X)
        return self

    def transform(self, X):
        check_is_fitted(self, "X_")
        Y, Yd = _rows(X)
        if Y.shape[1] != self.X_.shape[1]:
            raise ValueError(f"X has {Y.shape[1]} timesteps, but DtwKimLowerBound is "
                             f"expecting {self.X_.shape[1]} timesteps as input.")
        Xd = self._Xd.to(Yd.device)
        out = torch.empty((Y.shape[0], self.X_.shape[0]), dtype=torch.float64, device=Yd.device)
        _lib.check(_lib.load().wbx_dev_lb_kim(Yd.data_ptr(), Yd.shape[0], Yd.stride(0),
                                              Xd.data_ptr(), Xd.shape[0], Xd.stride(0),
                                              Yd.shape[1], out.data_ptr(), _stream()))
        return out.cpu().numpy()
