"""The allocation-free pairwise entry point against wbx_dev_pairwise.

wbx_dev_pairwise_noalloc keeps its workspace in a fixed per-GPU device array
instead of allocating.  Both entry points are called through the C ABI on the
same device inputs; their outputs must be identical (same kernels, same
arithmetic), on two calls in a row with the second one smaller (the array
keeps the first call's contents), for each metric / precision / layout.
"""

import ctypes

import numpy as np
import pytest
import torch

from wildboar_amd import _lib

pytestmark = pytest.mark.gpu


def _rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


def _call(fn, m, X, Y, out):
    dtype = _lib.F32 if X.dtype == torch.float32 else _lib.F64
    st = torch.cuda.current_stream().cuda_stream
    if Y is None:
        return fn(m, dtype, X.data_ptr(), X.shape[0], X.stride(0), X.shape[1], None,
                  0, 0, 0, out.data_ptr(), st)
    return fn(m, dtype, X.data_ptr(), X.shape[0], X.stride(0), X.shape[1], Y.data_ptr(),
              Y.shape[0], Y.stride(0), Y.shape[1], out.data_ptr(), st)


CASES = [
    # kind, precision, r, input dtype, singleton
    (_lib.DTW, "fp32", 0.1, torch.float32, False),
    (_lib.DTW, "fp32", 0.1, torch.float64, False),   # inputs staged to float32
    (_lib.DTW, "fp32", 1.0, torch.float32, False),   # wide band: block kernel
    (_lib.DTW, "fp64", 0.1, torch.float64, False),
    (_lib.DTW, "fp64", 0.05, torch.float32, True),   # inputs staged to float64
    (_lib.EUCLIDEAN, "fp64", 1.0, torch.float64, False),
    (_lib.EUCLIDEAN, "fp32", 1.0, torch.float32, True),
]


@pytest.mark.parametrize("kind,prec,r,dt,singleton", CASES)
def test_noalloc_pairwise_equals_allocating_entry_point(kind, prec, r, dt, singleton):
    lib = _lib.load()
    _lib.check(lib.wbx_init(0))
    m = _lib.metric_struct(kind, _lib.PRECISIONS[prec], r, 0.05)
    for (nx, ny, t), seed in (((300, 200, 128), 1), ((40, 30, 96), 2)):  # second call smaller
        X = torch.from_numpy(_rw(seed, nx, t)).to("cuda", dt)
        Y = None if singleton else torch.from_numpy(_rw(seed + 10, ny, t)).to("cuda", dt)
        n_out = (nx, nx) if singleton else (nx, ny)
        want = torch.full(n_out, -7.0, dtype=torch.float64, device="cuda")
        got = torch.full(n_out, -5.0, dtype=torch.float64, device="cuda")
        _lib.check(_call(lib.wbx_dev_pairwise, m, X, Y, want))
        _lib.check(_call(lib.wbx_dev_pairwise_noalloc, m, X, Y, got))
        torch.cuda.synchronize()
        assert torch.isfinite(want).all()
        assert torch.equal(got, want), (nx, ny, t)


def test_noalloc_pairwise_limits():
    lib = _lib.load()
    _lib.check(lib.wbx_init(0))
    X = torch.zeros((5000, 64), dtype=torch.float64, device="cuda")
    out = torch.empty((5000, 5000), dtype=torch.float64, device="cuda")
    m = _lib.metric_struct(_lib.DTW, _lib.PRECISIONS["fp64"], 0.1, 0.05)
    with pytest.raises(ValueError, match="scratch"):  # 25e6 float64 costs > 64 MiB
        _lib.check(_call(lib.wbx_dev_pairwise_noalloc, m, X, X, out))
    small = torch.zeros((4, 16), dtype=torch.float64, device="cuda")
    o = torch.empty((4, 4), dtype=torch.float64, device="cuda")
    for kind in (_lib.WDTW, _lib.DDTW):
        mw = _lib.metric_struct(kind, _lib.PRECISIONS["fp64"], 0.1, 0.05)
        with pytest.raises(NotImplementedError):
            _lib.check(_call(lib.wbx_dev_pairwise_noalloc, mw, small, small, o))
    ma = _lib.metric_struct(_lib.DTW, _lib.PRECISIONS["fp32_acc64"], 0.1, 0.05)
    with pytest.raises(NotImplementedError):
        _lib.check(_call(lib.wbx_dev_pairwise_noalloc, ma, small, small, o))
