"""Batched DTW over uneven-length pairs (device tensors).

Series are stored padded to a common width, with wildboar's end-of-series
marker (``variable_len.EOS``) after the last value; pair ``p`` is
``(X[p][:len_x[p]], Y[p][:len_y[p]])`` and has its own window fraction
``r[p]``.  The result is ``DTW(X[p], Y[p])`` exactly as ``DtwMetric``
computes it for the unpadded pair (``distance/_elastic.pyx:3167-3186``).

``DtwBatch`` plans once (host: geometry, routing, one queue sorted by strip
size and cost) and then runs on device pointers: one persistent strip launch
on the caller's stream (float32, or float64 for ``fp64`` / ``fp32_acc64``).
With ``graph=True`` the run is captured into a hipGraph on first use and
replayed afterwards.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .distance import _check_g, _check_r


def eos_lengths(X):
    """int32 lengths (index of the first EOS) of each row of a device tensor."""
    if X.dim() != 2 or X.stride(1) != 1:
        X = X.reshape(X.shape[0], -1).contiguous()
    dtype = _lib.F32 if X.dtype == torch.float32 else _lib.F64
    out = torch.empty(X.shape[0], dtype=torch.int32, device=X.device)
    _lib.check(_lib.load().wbx_dev_eos_lengths(dtype, X.data_ptr(), X.shape[0], X.stride(0),
                                               X.shape[1], out.data_ptr(),
                                               torch.cuda.current_stream().cuda_stream))
    return out


class DtwBatch:
    """Plan of a batched DTW run.

    Parameters
    ----------
    len_x, len_y : array-like of int, lengths of the x and y series per pair.
    t_pad : int, padded width of the arrays.
    r : float or array-like of float, window fraction per pair.
    metric : "dtw" or "wdtw" (WDTW weights are sized by ``t_pad``, the array
        width, as the reference sizes them by the array's n_timestep).
    precision : "fp32", "fp64", or "fp32_acc64" (inputs rounded to float32,
        float64 recursion).
    plain : bool or "padded", use the plain one-pair-per-block kernel for
        every pair; "padded" walks every pair over the whole padded length
        (the baseline that does not use per-pair lengths in its loops; same
        results, for measurements).
    """

    def __init__(self, len_x, len_y, t_pad, r=1.0, *, metric="dtw", g=0.05,
                 precision="fp32", plain=False):
        lx = np.ascontiguousarray(len_x, dtype=np.int32)
        ly = np.ascontiguousarray(len_y, dtype=np.int32)
        if lx.shape != ly.shape or lx.ndim != 1:
            raise ValueError("len_x and len_y must be 1-D of equal size")
        self.n_pairs = lx.shape[0]
        self.t_pad = int(t_pad)
        if np.ndim(r) == 0:
            rp = None
            r0 = _check_r(r)
        else:
            rp = np.ascontiguousarray(r, dtype=np.float64)
            if rp.shape != lx.shape:
                raise ValueError("r must be a scalar or one value per pair")
            r0 = 1.0
        kind = _lib.KINDS[metric]
        if kind == _lib.EUCLIDEAN:
            raise ValueError("batched pairs need metric 'dtw' or 'wdtw'")
        m = _lib.metric_struct(kind, _lib.PRECISIONS[precision], r0,
                               _check_g(g) if kind == _lib.WDTW else 0.05)
        if plain == "padded":
            policy = 2
        elif plain in (True, False):
            policy = 1 if plain else 0
        else:
            raise ValueError(f"plain must be a bool or 'padded', got {plain!r}")
        self._lib = _lib.load()
        plan = ctypes.c_void_p()
        _lib.check(self._lib.wbx_batch_create(
            m, self.n_pairs, self.t_pad, lx.ctypes.data, ly.ctypes.data,
            None if rp is None else rp.ctypes.data, policy, ctypes.byref(plan)))
        self._plan = plan
        cells = ctypes.c_int64()
        ns = ctypes.c_int64()
        npl = ctypes.c_int64()
        _lib.check(self._lib.wbx_batch_info(plan, ctypes.byref(cells), ctypes.byref(ns),
                                            ctypes.byref(npl)))
        self.cells = cells.value
        self.n_strip = ns.value
        self.n_plain = npl.value

    def set_launch(self, waves_per_cu=12, n_streams=1):
        """Persistent waves per CU of the strip launch (results do not depend
        on it); ``n_streams`` is accepted for compatibility and ignored."""
        _lib.check(self._lib.wbx_batch_set_launch(self._plan, int(waves_per_cu), int(n_streams)))

    def autotune(self, X, Y, reps=2):
        """Time the strip launch at several waves per CU on (X, Y) and keep
        the fastest.

        Returns the best time in ms (None when the choice came from the
        process-wide cache, or there is nothing to tune).
        """
        self._check_inputs(X, Y)
        ms = ctypes.c_float()
        dtype = _lib.F32 if X.dtype == torch.float32 else _lib.F64
        _lib.check(self._lib.wbx_batch_autotune(
            self._plan, dtype, X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0),
            torch.cuda.current_stream().cuda_stream, int(reps), ctypes.byref(ms)))
        return None if ms.value < 0 else float(ms.value)

    def buckets(self):
        """Per strip size: dict(R, pairs, waves, stream, ms) (waves and ms of
        the one strip launch)."""
        nb = ctypes.c_int()
        _lib.check(self._lib.wbx_batch_bucket_info(self._plan, -1, ctypes.byref(nb), None, None,
                                                   None, None, None))
        out = []
        for k in range(nb.value):
            v = [ctypes.c_int() for _ in range(4)]
            ms = ctypes.c_float()
            _lib.check(self._lib.wbx_batch_bucket_info(self._plan, k, None, *[ctypes.byref(x) for x in v],
                                                       ctypes.byref(ms)))
            out.append({"R": v[0].value, "pairs": v[1].value, "waves": v[2].value,
                        "stream": v[3].value, "ms": None if ms.value < 0 else float(ms.value)})
        return out

    def _check_inputs(self, X, Y):
        for t, name in ((X, "X"), (Y, "Y")):
            if not t.is_cuda or t.dim() != 2 or t.stride(1) != 1:
                raise ValueError(f"{name} must be a 2-D device tensor with contiguous rows")
            if t.shape[0] < self.n_pairs or t.shape[1] < self.t_pad:
                raise ValueError(f"{name} is smaller than the plan ({self.n_pairs}, {self.t_pad})")
        if X.dtype != Y.dtype or X.dtype not in (torch.float32, torch.float64):
            raise TypeError("X and Y must both be float32 or float64")

    def run(self, X, Y, out=None, *, graph=True):
        """Enqueue on torch's current stream; returns the float64 distances.

        With ``graph=True`` and no ``out``, the captured graph writes into a
        buffer the plan owns (so replays keep one capture) and the result is
        a copy of it, made on the same stream.
        """
        self._check_inputs(X, Y)
        dtype = _lib.F32 if X.dtype == torch.float32 else _lib.F64
        owned = out is None and graph
        if owned:
            if getattr(self, "_out", None) is None or self._out.device != X.device:
                self._out = torch.empty(self.n_pairs, dtype=torch.float64, device=X.device)
            out = self._out
        elif out is None:
            out = torch.empty(self.n_pairs, dtype=torch.float64, device=X.device)
        _lib.check(self._lib.wbx_batch_run(
            self._plan, dtype, X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0),
            out.data_ptr(), torch.cuda.current_stream().cuda_stream, 1 if graph else 0))
        return out.clone() if owned else out

    def close(self):
        """Releases the plan; waits for its last run to finish first."""
        if getattr(self, "_plan", None):
            self._lib.wbx_batch_free(self._plan)
            self._plan = None
        self._out = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dtw_batch(X, Y, r=1.0, *, len_x=None, len_y=None, precision="fp32", metric="dtw", g=0.05):
    """One-shot batched DTW; lengths default to the EOS positions of X and Y."""
    if len_x is None:
        len_x = eos_lengths(X).cpu().numpy()
    if len_y is None:
        len_y = eos_lengths(Y).cpu().numpy()
    plan = DtwBatch(len_x, len_y, X.shape[1], r, metric=metric, g=g, precision=precision)
    try:
        return plan.run(X, Y, graph=False)
    finally:
        torch.cuda.current_stream().synchronize()
        plan.close()
