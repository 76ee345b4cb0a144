"""GPU parity of the DTW lower bounds (envelopes, LB_Keogh, LB_Kim) and of
argmin_distance fed with them, against the oracle, bit for bit."""

import math

import numpy as np
import pytest

import oracle
from wildboar_amd import distance as wd
from wildboar_amd.dtw import dtw_envelop, dtw_lb_keogh
from wildboar_amd.lb import DtwKeoghLowerBound, DtwKimLowerBound, envelope_warp

pytestmark = pytest.mark.gpu


def rw(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


@pytest.mark.parametrize("t,r", [(1, 1.0), (2, 1.0), (7, 0.3), (50, 0.05), (50, 1.0), (333, 0.1)])
def test_envelope(t, r):
    rng = np.random.default_rng(t)
    for x in (rng.standard_normal(t), rng.integers(0, 3, t).astype(float)):
        lo, up = dtw_envelop(x, r=r)
        olo, oup = oracle.dtw_envelop(x, envelope_warp(t, r))
        np.testing.assert_array_equal(lo, olo)
        np.testing.assert_array_equal(up, oup)


@pytest.mark.parametrize("kind", ["both", "left", "right"])
@pytest.mark.parametrize("r", [0.1, 1.0])
def test_keogh_transform(kind, r):
    X = rw(100, 30, 64)
    Q = rw(101, 12, 64)
    got = DtwKeoghLowerBound(r, kind=kind).fit(X).transform(Q)
    w = envelope_warp(64, r)
    envX = [oracle.dtw_envelop(x, w) for x in X]
    envQ = [oracle.dtw_envelop(q, w) for q in Q]
    want = np.empty((12, 30))
    for i in range(12):
        for j in range(30):
            left = oracle.lb_keogh(Q[i], *envX[j]) if kind != "right" else -np.inf
            right = oracle.lb_keogh(X[j], *envQ[i]) if kind != "left" else -np.inf
            want[i, j] = max(left, right)
    np.testing.assert_array_equal(got, want)
    d, cb = dtw_lb_keogh(Q[0], X[0], r=r)
    assert d == oracle.lb_keogh(Q[0], *envX[0])
    assert math.isclose(math.sqrt(cb.sum()), d, rel_tol=1e-12)


@pytest.mark.parametrize("t", [1, 2, 3, 40])
def test_kim_transform(t):
    X = rw(102, 25, 40)[:, :t]
    Q = rw(103, 9, 40)[:, :t]
    np.testing.assert_array_equal(DtwKimLowerBound().fit(X).transform(Q), oracle.lb_kim(X, Q))


def test_argmin_with_keogh_lower_bound():
    X = rw(104, 40, 80)
    Y = rw(105, 200, 80)
    lb = DtwKeoghLowerBound(0.1).fit(Y).transform(X)
    for k in (1, 5):
        gi, gd = wd.argmin_distance(X, Y, k=k, metric="dtw", metric_params={"r": 0.1},
                                    lower_bound=lb, return_distance=True)
        oi, od = oracle.argmin(X, Y, k=k, metric="dtw", r=0.1, lower_bound=lb)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)
        # a valid bound never changes the k nearest
        pi, pd = wd.argmin_distance(X, Y, k=k, metric="dtw", metric_params={"r": 0.1},
                                    return_distance=True, sorted=True)
        np.testing.assert_array_equal(np.sort(gd, axis=1), pd)
