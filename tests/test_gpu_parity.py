"""GPU parity: libwbx.so against the float64 oracle on the same seeded inputs.

Bars
- FP64 mode: bit-identical to the oracle (assert_array_equal), indices exact.
- FP32 mode: distances within a max relative error of 1e-5 of the oracle
  (absolute 1e-6 near zero); argmin indices, heap order and returned
  distances exact (certified re-rank).
"""

import numpy as np
import pytest

import oracle
from wildboar_amd import _lib
from wildboar_amd import distance as wd

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-5


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


def assert_fp32_close(got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-1)
    assert np.isfinite(got).all()
    assert err.max() <= FP32_RTOL, float(err.max())


def setup_module(module):
    _lib.set_policy(0)


# ----------------------------------------------------------------- FP64 exact
@pytest.mark.parametrize("metric,params", [
    ("dtw", {"r": 0.1}), ("dtw", {"r": 0.0}), ("dtw", {"r": 1.0}), ("dtw", {"r": 0.37}),
    ("wdtw", {"r": 0.1, "g": 0.05}), ("wdtw", {"r": 1.0, "g": 0.5}), ("euclidean", {}),
])
def test_pairwise_fp64_bitwise_cfg1(metric, params):
    X = rw(0, 100, 128)
    Y = rw(1, 100, 128)
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params)
    want = oracle.pairwise(X, Y, metric=metric, r=params.get("r", 1.0), g=params.get("g", 0.05))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("lx,ly,r", [(1, 1, 0.5), (1, 7, 0.2), (9, 1, 0.9), (30, 47, 0.1),
                                      (47, 30, 0.1), (64, 200, 0.05), (200, 64, 1.0),
                                      (33, 33, 0.0)])
@pytest.mark.parametrize("metric", ["dtw", "wdtw"])
def test_pairwise_fp64_unequal_lengths(lx, ly, r, metric):
    X = np.random.default_rng(2).standard_normal((5, lx))
    Y = np.random.default_rng(3).standard_normal((6, ly))
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params={"r": r})
    want = oracle.pairwise(X, Y, metric=metric, r=r)
    np.testing.assert_array_equal(np.atleast_2d(got), want)


@pytest.mark.parametrize("metric", ["dtw", "wdtw", "euclidean"])
def test_singleton_and_paired_fp64(metric):
    X = rw(4, 37, 60)
    Y = rw(5, 37, 60)
    params = {} if metric == "euclidean" else {"r": 0.2}
    got = wd.pairwise_distance(X, metric=metric, metric_params=params)
    np.testing.assert_array_equal(got, oracle.pairwise(X, None, metric=metric, r=0.2))
    got = wd.paired_distance(X, Y, metric=metric, metric_params=params)
    np.testing.assert_array_equal(got, oracle.paired(X, Y, metric=metric, r=0.2))


def test_multivariate_dims_fp64():
    rng = np.random.default_rng(6)
    X = rng.standard_normal((7, 3, 40))
    Y = rng.standard_normal((5, 3, 40))
    for d in range(3):
        got = wd.pairwise_distance(X, Y, dim=d, metric="dtw", metric_params={"r": 0.3})
        np.testing.assert_array_equal(got, oracle.pairwise(X, Y, metric="dtw", r=0.3, dim=d))
    full = wd.pairwise_distance(X, Y, dim="full", metric="dtw", metric_params={"r": 0.3})
    assert full.shape == (3, 7, 5)
    mean = wd.pairwise_distance(X, Y, dim="mean", metric="dtw", metric_params={"r": 0.3})
    np.testing.assert_allclose(mean, full.mean(axis=0), rtol=0, atol=0)


@pytest.mark.parametrize("metric", ["dtw", "wdtw", "euclidean"])
@pytest.mark.parametrize("k", [1, 3, 7])
def test_argmin_fp64_heap_order(metric, k):
    X = rw(7, 40, 64)
    Y = rw(8, 90, 64)
    params = {} if metric == "euclidean" else {"r": 0.15}
    idx, dist = wd.argmin_distance(X, Y, k=k, metric=metric, metric_params=params,
                                   return_distance=True)
    oi, od = oracle.argmin(X, Y, k=k, metric=metric, r=0.15)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, od)


def test_argmin_lower_bound_and_docstring():
    X = np.array([[1, 2, 3, 4], [10, 1, 2, 3]], float)
    Y = np.array([[1, 2, 11, 2], [2, 4, 6, 7], [10, 11, 2, 3]], float)
    idx, dist = wd.argmin_distance(X, Y, k=2, return_distance=True)
    assert idx.tolist() == [[0, 1], [1, 2]]
    np.testing.assert_allclose(dist, [[8.24621125, 4.79583152], [10.24695077, 10.0]], atol=1e-8)
    Xr = rw(9, 20, 50)
    Yr = rw(10, 60, 50)
    lb = np.random.default_rng(11).uniform(0, 3, (20, 60))
    idx = wd.argmin_distance(Xr, Yr, k=4, metric="dtw", metric_params={"r": 0.1},
                             lower_bound=lb)
    oi, _ = oracle.argmin(Xr, Yr, k=4, metric="dtw", r=0.1, lower_bound=lb)
    np.testing.assert_array_equal(idx, oi)


@pytest.mark.parametrize("metric", ["dtw", "wdtw", "euclidean"])
def test_subsequence_fp64(metric):
    X = rw(12, 9, 100)
    rng = np.random.default_rng(13)
    shapelets = [X[rng.integers(9), s:s + L].copy() for s, L in [(3, 5), (40, 17), (0, 33), (70, 1)]]
    params = {} if metric == "euclidean" else {"r": 0.1}
    d, i = wd.pairwise_subsequence_distance(shapelets, X, metric=metric, metric_params=params,
                                            return_index=True)
    od, oi = oracle.subsequence(shapelets, X, metric=metric, r=0.1)
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(i, oi)


# ----------------------------------------------------------------- FP32
@pytest.mark.parametrize("n,r", [(128, 0.1), (512, 0.1), (512, 0.05), (256, 0.1), (128, 0.0),
                                 (100, 0.3), (64, 0.85), (512, 0.2)])
def test_pairwise_fp32_dtw(n, r):
    X = rw(20, 64, n)
    Y = rw(21, 70, n)
    got = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r}, precision="fp32")
    want = oracle.pairwise(X, Y, metric="dtw", r=r)
    assert_fp32_close(got, want)


def test_pairwise_fp32_every_exact_band_width():
    # Odd band widths 9 .. 111 run on kernels compiled for exactly that many
    # slots (no padding slot, no per-row reset); widths 2w - 1 at equal
    # lengths, 2w + 1 at lengths 60 / 62 (final cell off the band centre).
    X = rw(26, 10, 60)
    Y = rw(27, 12, 60)
    Y2 = rw(28, 12, 62)
    for w in range(5, 57):
        r = (w + 0.5) / 60.0
        got = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r}, precision="fp32")
        assert_fp32_close(got, oracle.pairwise(X, Y, metric="dtw", r=r))
        if 2 * w + 1 <= 111:
            got = wd.pairwise_distance(X, Y2, metric="dtw", metric_params={"r": r},
                                       precision="fp32")
            assert_fp32_close(got, oracle.pairwise(X, Y2, metric="dtw", r=r))


def test_pairwise_fp32_every_padded_band_width():
    # Lengths 60 / 61 give even band widths B = 2w, padded to a multiple of 8
    # with a reset slot: every padded class 8 .. 112 (the widest keeps its
    # group's loads at the group's start, the others load the next group's
    # during the current one), and the amerced variant on both sides of 80.
    X = rw(36, 10, 60)
    Y = rw(37, 12, 61)
    for w in range(4, 57):
        r = (w + 0.5) / 60.0
        got = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r}, precision="fp32")
        assert_fp32_close(got, oracle.pairwise(X, Y, metric="dtw", r=r))
    for w in (12, 30, 39, 45, 56):
        r = (w + 0.5) / 60.0
        got = wd.pairwise_distance(X, Y, metric="adtw", metric_params={"r": r, "p": 0.5},
                                   precision="fp32")
        assert_fp32_close(got, oracle.pairwise(X, Y, metric="adtw", r=r, p=0.5))


@pytest.mark.parametrize("t", [1, 2, 31, 33, 63, 65, 97])
def test_pairwise_euclidean_tile_edges(t):
    # Equal lengths take the 64 x 64-pair LDS tiles, which read two timesteps
    # at a time: odd lengths end on a single timestep, short ones inside the
    # first chunk, and 70 x 67 rows leave partial tiles on both edges.
    X = np.random.default_rng(48).standard_normal((70, t))
    Y = np.random.default_rng(49).standard_normal((67, t))
    want = oracle.pairwise(X, Y, metric="euclidean")
    np.testing.assert_array_equal(wd.pairwise_distance(X, Y, metric="euclidean"), want)
    got = wd.pairwise_distance(X, Y, metric="euclidean", precision="fp32")
    assert_fp32_close(got, want)


def test_pairwise_fp64_lane_classes_bitwise():
    # float64 lanes: DTW up to 40 slots at equal and unequal lengths, ADTW in
    # every class to 104 slots (the classes that load the next group during
    # the current one and those that do not); the reference's arithmetic, so
    # bit for bit.
    X = rw(46, 8, 60)
    for ly in (60, 61):
        Y = rw(47, 9, ly)
        for w in range(2, 21):
            r = (w + 0.5) / 60.0
            got = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r})
            np.testing.assert_array_equal(got, oracle.pairwise(X, Y, metric="dtw", r=r))
        for w in (6, 12, 16, 20, 26, 30, 36, 44, 52):
            r = (w + 0.5) / 60.0
            got = wd.pairwise_distance(X, Y, metric="adtw", metric_params={"r": r, "p": 0.5})
            np.testing.assert_array_equal(got, oracle.pairwise(X, Y, metric="adtw", r=r, p=0.5))


@pytest.mark.parametrize("lx,ly,r", [(100, 130, 0.1), (130, 100, 0.1), (37, 80, 0.2),
                                      (5, 3, 1.0), (1, 9, 0.5)])
def test_pairwise_fp32_unequal(lx, ly, r):
    X = np.random.default_rng(22).standard_normal((33, lx))
    Y = np.random.default_rng(23).standard_normal((65, ly))
    got = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": r}, precision="fp32")
    assert_fp32_close(np.atleast_2d(got), oracle.pairwise(X, Y, metric="dtw", r=r))


@pytest.mark.parametrize("lx,ly,r,g", [(128, 128, 0.1, 0.05), (100, 130, 0.1, 0.3),
                                        (130, 100, 0.2, 0.05), (37, 80, 0.2, 1.0),
                                        (5, 3, 1.0, 0.05), (1, 9, 0.5, 0.05), (256, 256, 0.2, 0.05),
                                        (64, 64, 0.0, 0.5)])
def test_pairwise_fp32_wdtw_lane_kernel(lx, ly, r, g):
    # FP32 WDTW on the lane kernel (per-slot weights, row-0 prefix weights
    # w[j-1]); also the singleton and paired layouts, and the block kernel
    X = np.random.default_rng(32).standard_normal((33, lx))
    Y = np.random.default_rng(33).standard_normal((65, ly))
    params = {"r": r, "g": g}
    got = wd.pairwise_distance(X, Y, metric="wdtw", metric_params=params, precision="fp32")
    want = oracle.pairwise(X, Y, metric="wdtw", r=r, g=g)
    assert_fp32_close(np.atleast_2d(got), want)
    try:
        _lib.set_policy(1)
        plain = wd.pairwise_distance(X, Y, metric="wdtw", metric_params=params, precision="fp32")
    finally:
        _lib.set_policy(0)
    assert_fp32_close(np.atleast_2d(plain), want)
    if lx == ly:
        got = wd.pairwise_distance(X, metric="wdtw", metric_params=params, precision="fp32")
        assert_fp32_close(got, oracle.pairwise(X, None, metric="wdtw", r=r, g=g))
        got = wd.paired_distance(X, Y[:33], metric="wdtw", metric_params=params, precision="fp32")
        assert_fp32_close(got, oracle.paired(X, Y[:33], metric="wdtw", r=r, g=g))


def test_pairwise_fp32_plain_kernel_matches():
    X = rw(24, 40, 200)
    Y = rw(25, 50, 200)
    want = oracle.pairwise(X, Y, metric="dtw", r=0.1)
    try:
        _lib.set_policy(1)
        plain = wd.pairwise_distance(X, Y, metric="dtw", metric_params={"r": 0.1}, precision="fp32")
    finally:
        _lib.set_policy(0)
    assert_fp32_close(plain, want)


@pytest.mark.parametrize("metric", ["dtw", "euclidean", "wdtw"])
@pytest.mark.parametrize("k", [1, 5])
def test_argmin_fp32_certified_exact(metric, k):
    X = rw(26, 150, 128)
    Y = rw(27, 300, 128)
    params = {} if metric == "euclidean" else {"r": 0.1}
    idx, dist = wd.argmin_distance(X, Y, k=k, metric=metric, metric_params=params,
                                   return_distance=True, precision="fp32")
    oi, od = oracle.argmin(X, Y, k=k, metric=metric, r=0.1)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, od)


def test_argmin_fp32_float32_inputs_and_ties():
    X = rw(28, 50, 96).astype(np.float32)
    Y = np.repeat(rw(29, 20, 96).astype(np.float32), 3, axis=0)  # exact duplicates -> ties
    idx = wd.argmin_distance(X, Y, k=1, metric="dtw", metric_params={"r": 0.1}, precision="fp32")
    oi, _ = oracle.argmin(X.astype(np.float64), Y.astype(np.float64), k=1, metric="dtw", r=0.1)
    np.testing.assert_array_equal(idx, oi)
    assert (idx % 3 == 0).all()


# ----------------------------------------------------------------- subsequence FP32
@pytest.mark.parametrize("metric,r", [("dtw", 0.1), ("dtw", 0.3), ("euclidean", 0.1)])
def test_subsequence_fp32_certified_exact(metric, r):
    X = rw(30, 23, 300)
    rng = np.random.default_rng(31)
    shapelets = []
    for L in (16, 17, 33, 64, 90, 128, 5, 1):
        i = int(rng.integers(23))
        s = int(rng.integers(0, 300 - L + 1))
        shapelets.append(X[i, s:s + L].copy())
    params = {} if metric == "euclidean" else {"r": r}
    d, i = wd.pairwise_subsequence_distance(shapelets, X, metric=metric, metric_params=params,
                                            return_index=True, precision="fp32")
    od, oi = oracle.subsequence(shapelets, X, metric=metric, r=r)
    np.testing.assert_array_equal(i, oi)
    np.testing.assert_array_equal(d, od)


def test_subsequence_fp32_ties_overflow_and_wide_band():
    # periodic series: many windows tie exactly -> candidate lists overflow
    t = np.arange(256)
    X = np.stack([np.sin(2 * np.pi * t / 16.0), np.cos(2 * np.pi * t / 32.0)]).astype(np.float32)
    shapelets = [X[0, 3:35].copy(), X[1, 10:42].copy(), X[0, 0:100].copy()]
    for metric, params in (("dtw", {"r": 0.2}), ("dtw", {"r": 1.0}), ("euclidean", {})):
        d, i = wd.pairwise_subsequence_distance(shapelets, X.astype(np.float64), metric=metric,
                                                metric_params=params, return_index=True,
                                                precision="fp32")
        od, oi = oracle.subsequence(shapelets, X.astype(np.float64), metric=metric,
                                    r=params.get("r", 1.0))
        np.testing.assert_array_equal(i, oi)
        np.testing.assert_array_equal(d, od)


# ----------------------------------------------------------------- FP32_ACC64
def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("metric,params", [("dtw", {"r": 0.1}), ("dtw", {"r": 1.0}),
                                           ("wdtw", {"r": 0.2, "g": 0.1}), ("euclidean", {})])
def test_fp32_acc64_is_fp64_on_rounded_inputs(metric, params):
    # float32 data, float64 recursion: bitwise the oracle on the rounded values
    X = rw(40, 30, 90)
    Y = rw(41, 25, 90)
    r, g = params.get("r", 1.0), params.get("g", 0.05)
    got = wd.pairwise_distance(X, Y, metric=metric, metric_params=params, precision="fp32_acc64")
    np.testing.assert_array_equal(got, oracle.pairwise(f32(X), f32(Y), metric=metric, r=r, g=g))
    got = wd.paired_distance(X, Y[:25].repeat(2, 0)[:30], metric=metric, metric_params=params,
                             precision="fp32_acc64")
    np.testing.assert_array_equal(
        got, oracle.paired(f32(X), f32(Y[:25].repeat(2, 0)[:30]), metric=metric, r=r, g=g))
    i, d = wd.argmin_distance(X, Y, k=3, metric=metric, metric_params=params,
                              return_distance=True, precision="fp32_acc64")
    oi, od = oracle.argmin(f32(X), f32(Y), k=3, metric=metric, r=r, g=g)
    np.testing.assert_array_equal(i, oi)
    np.testing.assert_array_equal(d, od)
    shp = [X[3, 5:25].copy(), Y[7, 40:47].copy()]
    d, i = wd.pairwise_subsequence_distance(shp, X, metric=metric, metric_params=params,
                                            return_index=True, precision="fp32_acc64")
    od, oi = oracle.subsequence([f32(s) for s in shp], f32(X), metric=metric, r=r, g=g)
    np.testing.assert_array_equal(i, oi)
    np.testing.assert_array_equal(d, od)


def test_device_subsequence_packed_matches_list_and_oracle():
    import torch
    from wildboar_amd import device

    X = rw(50, 12, 200)
    rng = np.random.default_rng(51)
    shp = [X[int(rng.integers(12)), s:s + L].copy() for s, L in [(3, 20), (50, 33), (100, 7), (9, 64)]]
    Xd = torch.from_numpy(X.astype(np.float32)).cuda()
    sd = [torch.from_numpy(s.astype(np.float32)) for s in shp]
    packed = device.pack_shapelets(sd, Xd.device)
    for metric in ("dtw", "euclidean"):
        d1, i1 = device.subsequence(sd, Xd, metric=metric, r=0.1)
        d2, i2 = device.subsequence(packed, Xd, metric=metric, r=0.1)
        assert torch.equal(d1, d2) and torch.equal(i1, i2)
        od, oi = oracle.subsequence([s.astype(np.float32).astype(np.float64) for s in shp],
                                    X.astype(np.float32).astype(np.float64), metric=metric, r=0.1)
        np.testing.assert_array_equal(i2.cpu().numpy(), oi)
        np.testing.assert_array_equal(d2.cpu().numpy(), od)
