"""GPU parity of the matrix-core euclidean subsequence scan (subseq_mfma.hip).

pairwise_subsequence_distance(metric="euclidean", precision="fp32") scans
shapelets up to 128 long with v_mfma_f32_32x32x2f32 dot products and a
certified candidate set, then evaluates the candidates with the reference's
float64 sum: distances and window indices must equal the float64 oracle's
bit for bit.  The cases cover every chunk length K = 8 .. 128, partial
chunks, exact ties (first window wins), lists that overflow into the exact
rescan, rows out of the float32 range, float64 inputs rounded for the scan,
a single window per series and shapelets past 128 (vector scan) in the same
call.
"""

import numpy as np
import pytest

import oracle
from wildboar_amd import _lib
from wildboar_amd import distance as wd

pytestmark = pytest.mark.gpu


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def setup_module(module):
    _lib.set_policy(0)


def scan(shapelets, X):
    return wd.pairwise_subsequence_distance(shapelets, X, metric="euclidean",
                                            return_index=True, precision="fp32")


def check(shapelets, X, exact_inputs=True):
    if exact_inputs:  # float32 data: the scan's inputs are exact
        shapelets = [np.asarray(s, np.float32) for s in shapelets]
        X = np.asarray(X, np.float32)
        d, i = scan(shapelets, X)
        od, oi = oracle.subsequence([f32(s) for s in shapelets], f32(X), metric="euclidean")
    else:
        d, i = scan(shapelets, X)
        od, oi = oracle.subsequence(shapelets, X, metric="euclidean")
    np.testing.assert_array_equal(i, oi)
    np.testing.assert_array_equal(d, od)


def cut(X, rng, lengths, noise=0.0):
    out = []
    for L in lengths:
        r = int(rng.integers(X.shape[0]))
        s = int(rng.integers(0, X.shape[1] - L + 1))
        out.append(X[r, s:s + L] + noise * rng.standard_normal(L))
    return out


def test_every_chunk_length_and_partial_chunk():
    X = f32(rw(60, 6, 300))
    rng = np.random.default_rng(61)
    # 32 shapelets in each length class (8 (k-1), 8 k], k = 1 .. 16, plus 7 more
    lengths = [int(rng.integers(8 * (k - 1) + 1, 8 * k + 1)) for k in range(1, 17) for _ in range(32)]
    lengths += [int(v) for v in rng.integers(1, 129, 7)]
    rng.shuffle(lengths)
    shp = cut(X, rng, lengths, noise=0.05)
    check(shp, X)


def test_float64_inputs_rounded_for_the_scan():
    X = rw(62, 8, 257)
    rng = np.random.default_rng(63)
    shp = cut(X, rng, [int(v) for v in rng.integers(1, 129, 70)], noise=0.01)
    check(shp, X, exact_inputs=False)


def test_exact_ties_first_window_and_overflow():
    base = np.array([0.5, -1.25, 2.0, 0.75, -0.5, 1.5, -2.0], np.float32)
    X = np.stack([np.tile(base, 40)[:270],                       # exact repeats every 7
                  np.full(270, 0.25, np.float32),                 # every window ties
                  np.tile(np.array([1.0, -1.0], np.float32), 135)])
    shp = [base[2:].copy(), np.tile(base, 3)[1:19].copy(), np.full(40, 0.25, np.float32),
           np.array([-1.0, 1.0, -1.0], np.float32), np.tile(base, 10)[:64].copy(),
           np.full(128, 0.25, np.float32)]
    check(shp, X)


def test_single_window_and_tiny_series():
    X = f32(rw(64, 5, 33))
    rng = np.random.default_rng(65)
    shp = cut(X, rng, [33, 32, 31, 1, 2, 17]) + [X[2].copy()]
    check(shp, X)
    Xs = f32(rw(66, 4, 3))
    check([Xs[0, :1].copy(), Xs[1, :3].copy(), Xs[3, 1:3].copy()], Xs)


def test_out_of_range_rows_and_long_shapelets_in_one_call():
    X = f32(rw(67, 7, 400)).astype(np.float32)
    X[3] *= np.float32(2.0 ** 45)           # past the scan's range: exact rescan
    rng = np.random.default_rng(68)
    shp = cut(X[[0, 1, 2, 4, 5, 6]], rng, [16, 40, 128, 129, 200, 5, 77])
    shp.append(X[3, 10:60].copy())
    check(shp, X)


def test_many_series_against_oracle():
    X = f32(rw(69, 120, 1024)).astype(np.float32)
    rng = np.random.default_rng(70)
    shp = cut(X, rng, [int(v) for v in rng.integers(16, 129, 40)])
    check(shp, X)
