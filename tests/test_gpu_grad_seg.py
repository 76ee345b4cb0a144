"""GPU checks of the band-checkpointed differentiable DTW / WDTW
(train_seg.hip: cost rows kept at anti-diagonal band checkpoints, each band
recomputed with its backtrack decisions in the backward).

It serves WDTW, bands wider than 112 slots, and ``checkpoint=True``.  The
bar is that of the decision-plane path:
- forward values within 1e-5 relative of the float64 oracle (and bitwise
  equal to the plane path's for DTW: the same float32 cell arithmetic);
- gradients against the float64 path-backtrack gradient (oracle.dtw_grad,
  restating _dtw_alignment's cost matrix with the forward's weights and
  dtw_mapping's backtrack).  Small integer inputs make every float32 cost an
  exact integer (for WDTW with g = 0 every weight is 1/2, still exact), so
  the GPU's path equals the float64 path cell for cell, ties included;
  tolerance 1e-5 of the gradient's max magnitude (float32 sums);
- WDTW with g = 0.05 on random walks: gradients on pairs whose float64 path
  has a relative margin above the float32 error bound of a path cost;
- bitwise determinism of the backward.
"""

import math

import numpy as np
import pytest
import torch

import oracle
from wildboar_amd.dtw import dtw, paired_dtw

pytestmark = pytest.mark.gpu


def ints(seed, n, t):
    return np.random.default_rng(seed).integers(-4, 5, (n, t)).astype(np.float32)


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return ((x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)).astype(np.float32)


def wvec(tx, ty, g):
    return oracle.wdtw_weights(max(tx, ty), g) if g is not None else None


def paired_oracle(X, Y, r, G, g=None):
    w = wvec(X.shape[1], Y.shape[1], g)
    gx = np.zeros(X.shape, np.float64)
    gy = np.zeros(Y.shape, np.float64)
    for i in range(X.shape[0]):
        _, ox, oy = oracle.dtw_grad(X[i].astype(np.float64), Y[i].astype(np.float64), r, w)
        gx[i], gy[i] = G[i] * ox, G[i] * oy
    return gx, gy


def grid_oracle(X, Y, r, G, g=None):
    w = wvec(X.shape[1], Y.shape[1], g)
    gx = np.zeros(X.shape, np.float64)
    gy = np.zeros(Y.shape, np.float64)
    for a in range(X.shape[0]):
        for b in range(Y.shape[0]):
            _, ox, oy = oracle.dtw_grad(X[a].astype(np.float64), Y[b].astype(np.float64), r, w)
            gx[a] += G[a, b] * ox
            gy[b] += G[a, b] * oy
    return gx, gy


def close(got, want):
    scale = max(np.abs(want).max(), 1e-3)
    assert np.max(np.abs(got - want)) <= 1e-5 * scale, float(np.max(np.abs(got - want)) / scale)


def run_paired(X, Y, r, G, **kw):
    Xd = torch.from_numpy(X).cuda().requires_grad_(True)
    Yd = torch.from_numpy(Y).cuda().requires_grad_(True)
    D = paired_dtw(Xd, Yd, r, **kw)
    (D * torch.from_numpy(G).cuda()).sum().backward()
    return D.detach().cpu().numpy(), Xd.grad.cpu().numpy(), Yd.grad.cpu().numpy()


def run_grid(X, Y, r, G, **kw):
    Xd = torch.from_numpy(X).cuda().requires_grad_(True)
    Yd = torch.from_numpy(Y).cuda().requires_grad_(True)
    D = dtw(Xd, Yd, r, **kw)
    (D * torch.from_numpy(G).cuda()).sum().backward()
    return D.detach().cpu().numpy(), Xd.grad.cpu().numpy(), Yd.grad.cpu().numpy()


@pytest.mark.parametrize("tx,ty,r", [(64, 64, 0.1), (100, 100, 0.3), (33, 33, 1.0),
                                     (37, 50, 0.2), (50, 37, 0.2)])
def test_checkpoint_dtw_paired_and_grid(tx, ty, r):
    X, Y = ints(30, 24, tx), ints(31, 24, ty)
    G = np.random.default_rng(32).uniform(0.5, 1.5, 24)
    D, gX, gY = run_paired(X, Y, r, G, checkpoint=True)
    D0, _, _ = run_paired(X, Y, r, G)  # decision planes
    assert np.array_equal(D, D0)
    wx, wy = paired_oracle(X, Y, r, G)
    close(gX, wx)
    close(gY, wy)
    Gg = np.random.default_rng(33).uniform(0.1, 1.0, (7, 5))
    Dg, gX, gY = run_grid(X[:7], Y[:5], r, Gg, checkpoint=True)
    want = oracle.pairwise(X[:7].astype(np.float64), Y[:5].astype(np.float64), metric="dtw", r=r)
    assert np.max(np.abs(Dg - want) / np.maximum(want, 1e-3)) <= 1e-5
    wx, wy = grid_oracle(X[:7], Y[:5], r, Gg)
    np.testing.assert_allclose(gX, wx, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(gY, wy, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("tx,ty,r", [(256, 256, 1.0), (128, 128, 0.5), (150, 120, 1.0),
                                     (300, 300, 0.25)])
def test_wide_band_dtw_gradient(tx, ty, r):
    # bands of 511, 127, 179 and 149 slots: past the decision-plane kernels
    X, Y = ints(34, 10, tx), ints(35, 10, ty)
    G = np.random.default_rng(36).uniform(0.5, 1.5, 10)
    D, gX, gY = run_paired(X, Y, r, G)
    want = np.array([oracle.metric_distance(X[i].astype(np.float64), Y[i].astype(np.float64),
                                            "dtw", r=r) for i in range(10)])
    assert np.max(np.abs(D - want) / np.maximum(want, 1e-3)) <= 1e-5
    wx, wy = paired_oracle(X, Y, r, G)
    close(gX, wx)
    close(gY, wy)
    Gg = np.random.default_rng(37).uniform(0.1, 1.0, (3, 4))
    _, gX, gY = run_grid(X[:3], Y[:4], r, Gg)
    wx, wy = grid_oracle(X[:3], Y[:4], r, Gg)
    np.testing.assert_allclose(gX, wx, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(gY, wy, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("tx,ty,r", [(64, 64, 0.1), (90, 70, 0.3), (48, 48, 1.0)])
def test_wdtw_gradient_g0_is_half_weighted_dtw(tx, ty, r):
    # g = 0: every weight is 1/2 exactly, integer inputs keep every cost exact
    X, Y = ints(38, 16, tx), ints(39, 16, ty)
    G = np.random.default_rng(40).uniform(0.5, 1.5, 16)
    D, gX, gY = run_paired(X, Y, r, G, metric="wdtw", g=0.0)
    want = np.array([oracle.metric_distance(X[i].astype(np.float64), Y[i].astype(np.float64),
                                            "wdtw", r=r, g=0.0) for i in range(16)])
    assert np.max(np.abs(D - want) / np.maximum(want, 1e-3)) <= 1e-5
    wx, wy = paired_oracle(X, Y, r, G, g=0.0)
    close(gX, wx)
    close(gY, wy)
    Gg = np.random.default_rng(41).uniform(0.1, 1.0, (5, 4))
    _, gX, gY = run_grid(X[:5], Y[:4], r, Gg, metric="wdtw", g=0.0)
    wx, wy = grid_oracle(X[:5], Y[:4], r, Gg, g=0.0)
    np.testing.assert_allclose(gX, wx, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(gY, wy, rtol=1e-5, atol=1e-5)


def weighted_margin(x, y, r, w):
    band = oracle.compute_r(min(len(x), len(y)), r)
    A = oracle.cost_matrix(x, y, band, w)
    pi, pj = oracle.mapping(A)
    best = math.inf
    for i, j in zip(pi[1:], pj[1:]):
        opts = sorted([A[i - 1, j - 1] if i > 0 and j > 0 else math.inf,
                       A[i - 1, j] if i > 0 else math.inf,
                       A[i, j - 1] if j > 0 else math.inf])
        if math.isfinite(opts[1]) and opts[0] > 0:
            best = min(best, (opts[1] - opts[0]) / opts[0])
    return best


@pytest.mark.parametrize("t,r,need", [(40, 0.1, 40), (64, 1.0, 20)])
def test_wdtw_gradient_random_walks_g005(t, r, need):
    # 64 pairs; `need` of them have a float64 path whose every decision has a
    # margin above the bound (47 and 26 with these seeds)
    X, Y = rw(42, 64, t), rw(43, 64, t)
    G = np.ones(64)
    D, gX, gY = run_paired(X, Y, r, G, metric="wdtw", g=0.05)
    w = oracle.wdtw_weights(t, 0.05)
    want = np.array([oracle.metric_distance(X[i].astype(np.float64), Y[i].astype(np.float64),
                                            "wdtw", r=r, g=0.05) for i in range(64)])
    assert np.max(np.abs(D - want) / want) <= 1e-5
    bound = 4 * (2 * t + 4) * 2.0 ** -24
    checked = 0
    for i in range(64):
        x64, y64 = X[i].astype(np.float64), Y[i].astype(np.float64)
        if weighted_margin(x64, y64, r, w) <= bound:
            continue
        _, ox, oy = oracle.dtw_grad(x64, y64, r, w)
        scale = max(np.abs(ox).max(), np.abs(oy).max())
        assert np.max(np.abs(gX[i] - ox)) <= 1e-4 * scale
        assert np.max(np.abs(gY[i] - oy)) <= 1e-4 * scale
        checked += 1
    assert checked >= need


def test_checkpoint_backward_is_deterministic():
    X = torch.from_numpy(rw(44, 40, 200)).cuda()
    Y = torch.from_numpy(rw(45, 6, 200)).cuda()
    grads = []
    for _ in range(2):
        Xd = X.clone().requires_grad_(True)
        Yd = Y.clone().requires_grad_(True)
        dtw(Xd, Yd, 0.7, metric="wdtw", g=0.1).sum().backward()
        grads.append((Xd.grad.clone(), Yd.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])


def test_wdtw_one_slot_band_is_refused():
    X = torch.zeros((2, 10), device="cuda")
    with pytest.raises(Exception, match="band of 2 slots"):
        paired_dtw(X, X, 0.0, metric="wdtw", g=0.05)
