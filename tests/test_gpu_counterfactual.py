"""GPU parity of the DTW counterfactual transforms: moves along the float64
optimal path equal a restatement of explain/counterfactual/_proto.py:586-673
over the oracle's alignment matrix and backtrack, bit for bit."""

import numpy as np
import pytest

import oracle
from wildboar_amd.counterfactual import (DynamicTimeWarpTransform,
                                         WeightedDynamicTimeWarpTransform, jeong_weight)

pytestmark = pytest.mark.gpu


def restated(o, target, r, weight, gamma=None, step=None):
    w = oracle.warp_size(o.shape[0], r, target.shape[0])
    oi, ti = oracle.mapping(oracle.dtw_alignment(o, target, w, weight))
    count = np.bincount(oi, minlength=o.shape[0])
    s = np.bincount(oi, weights=target[ti], minlength=o.shape[0])
    avg = np.zeros(o.shape[0])
    mask = count > 0
    avg[mask] = s[mask] / count[mask]
    if gamma is None:
        d = np.linalg.norm(avg - o)
        if d < 1e-10:
            return o.copy()
        gamma = min(1.0, step / d)
    return (1 - gamma) * o + gamma * avg


@pytest.mark.parametrize("r", [0.1, 0.5, 1.0])
def test_dtw_transform_moves(r):
    rng = np.random.default_rng(110)
    p = np.cumsum(rng.standard_normal(80))
    for lo in (80, 60, 95):
        o = np.cumsum(rng.standard_normal(lo))
        for step in (0.05, 1.0, 100.0):
            got = DynamicTimeWarpTransform(p, r=r).move(o, step)
            np.testing.assert_array_equal(got, restated(o, p, r, None, step=step))
        orig = o + rng.standard_normal(lo) * 0.1
        got = DynamicTimeWarpTransform(p, r=r).reverse_move(o, orig, 0.3)
        np.testing.assert_array_equal(got, restated(o, orig, r, None, gamma=0.3))
    # a sample already on the prototype stays
    np.testing.assert_array_equal(DynamicTimeWarpTransform(p, r=r).move(p, 1.0), p)


@pytest.mark.parametrize("g", [0.05, 0.5])
def test_weighted_dtw_transform_moves(g):
    rng = np.random.default_rng(111)
    p = np.cumsum(rng.standard_normal(64))
    o = np.cumsum(rng.standard_normal(64))
    tr = WeightedDynamicTimeWarpTransform(p, r=0.3, g=g)
    np.testing.assert_array_equal(tr.move(o, 0.7),
                                  restated(o, p, 0.3, jeong_weight(64, g), step=0.7))
    orig = o[::-1].copy()
    np.testing.assert_array_equal(tr.reverse_move(o, orig, 0.5),
                                  restated(o, orig, 0.3, jeong_weight(64, g), gamma=0.5))
