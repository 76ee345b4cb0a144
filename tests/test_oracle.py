"""Pins the float64 oracle (oracle/oracle.c) before anything is compared to it.

Independent checks, each chosen so a plausible slip in the restatement fails:
- a pure-Python masked full-matrix DTW written from the band definition
  (all lengths 1..8 x 1..8, a grid of r, with and without weights) -- catches
  band-edge, index and row-0 mistakes;
- r = 0 is euclidean bit for bit (diagonal-only band, same summation order);
- r = 1 is the textbook unconstrained DTW;
- D(x, x) = 0, symmetry of DTW, monotone non-increasing in r;
- WDTW with g = 0 has every weight exactly 0.5, so WDTW^2 == 0.5 * DTW^2;
- the WDTW weight vector by its closed forms: w[n/2] = 1/2 exactly, point
  symmetry w[i] + w[n-i] = 1, strictly increasing for g > 0, agreement with an
  independent logistic (scipy expit), and the size rule n = max(widths)
  through a brute force with expit weights;
- the argmin_distance docstring known answer (distance/_distance.py:1380-1387);
- argmin equals a brute-force "k smallest, first index on ties" selection;
- subsequence scans against brute force over windows, euclidean == dtw(r=0),
  pairwise == pairwise_subsequence.T at equal lengths
  (tests/wildboar/distance/test_distance.py:949-958, 1440-1477);
- the path-backtrack gradient against central finite differences and the
  r = 0 closed form dD/dx_i = (x_i - y_i) / D.
"""

import math

import numpy as np
import pytest

import oracle

INF = math.inf


def band_w(lx, ly, r):
    return max(math.floor(min(lx, ly) * r), 1)


def brute_dtw(x, y, r, weights=None):
    """Full-matrix DTW with an explicit band mask, from the definition.

    Cell (i, j) is in the band iff
      -(max(0, ly-lx) + w - 1) <= i - j <= max(0, lx-ly) + w - 1.
    Costs outside the band or the matrix are +inf; (0, 0) starts the path;
    row-0 cells weigh with w[j-1] (distance/_elastic.pyx:898-903), every
    other cell with w[|i-j|].
    """
    lx, ly = len(x), len(y)
    w = band_w(lx, ly, r)
    hi = max(0, lx - ly) + w - 1      # i - j <= hi
    lo = -(max(0, ly - lx) + w - 1)   # i - j >= lo
    C = [[INF] * ly for _ in range(lx)]
    for i in range(lx):
        for j in range(ly):
            if not (lo <= i - j <= hi):
                continue
            v = float(x[i]) - float(y[j])
            if weights is None:
                wt = 1.0
            elif i == 0 and j > 0:
                wt = float(weights[j - 1])
            else:
                wt = float(weights[abs(i - j)])
            if i == 0 and j == 0:
                C[i][j] = v * v * wt
                continue
            up = C[i - 1][j] if i > 0 else INF
            left = C[i][j - 1] if j > 0 else INF
            diag = C[i - 1][j - 1] if (i > 0 and j > 0) else INF
            m = left if left < up else up
            m = diag if diag < m else m
            C[i][j] = m + v * v * wt
    return C[lx - 1][ly - 1]


def textbook_dtw(x, y):
    n, m = len(x), len(y)
    D = np.full((n + 1, m + 1), np.inf)
    D[0, 0] = 0.0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            D[i, j] = (x[i - 1] - y[j - 1]) ** 2 + min(D[i - 1, j], D[i, j - 1], D[i - 1, j - 1])
    return D[n, m]


def rw(rng, n, t):
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    x = (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)
    return x


R_GRID = [0.0, 0.05, 0.1, 0.25, 0.5, 0.75, 1.0]


def test_brute_force_all_small_shapes():
    rng = np.random.default_rng(1)
    for lx in range(1, 9):
        for ly in range(1, 9):
            x = rng.standard_normal(lx)
            y = rng.standard_normal(ly)
            for r in R_GRID:
                band = oracle.compute_r(min(lx, ly), r)
                got = oracle.dtw_distance(x, y, band)
                want = brute_dtw(x, y, r)
                assert got == want, (lx, ly, r, got, want)


def test_brute_force_weighted_with_row0_quirk():
    rng = np.random.default_rng(2)
    for lx in range(1, 8):
        for ly in range(1, 8):
            x = rng.standard_normal(lx)
            y = rng.standard_normal(ly)
            wts = rng.uniform(0.1, 2.0, max(lx, ly))
            for r in (0.0, 0.3, 1.0):
                band = oracle.compute_r(min(lx, ly), r)
                got = oracle.dtw_distance(x, y, band, weights=wts)
                assert got == brute_dtw(x, y, r, wts), (lx, ly, r)


def _ulps(a, b):
    """Distance in units in the last place of the larger magnitude."""
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


@pytest.mark.parametrize("n,g", [(2, 0.05), (8, 0.5), (128, 0.05), (512, 0.05), (1024, 0.01),
                                 (300, 1.3)])
def test_wdtw_weights_closed_forms_even_n(n, g):
    # The logistic weight 1 / (1 + exp(-g (i - n/2))) (distance/_elastic.pyx:3334-3341)
    # is pinned by properties the formula fixes, not by retyping it:
    #  - centre: i = n/2 gives exp(0) = 1, so w = 1/2 exactly;
    #  - point symmetry about the centre: w[i] + w[n - i] = 1 (logistic);
    #  - strictly increasing in i for g > 0 (a sign slip reverses it);
    #  - an independent logistic (scipy's expit) agrees to a few ulp.
    # A centring on n instead of n/2, or i - 1 for i, fails the first two.
    from scipy.special import expit

    w = oracle.wdtw_weights(n, g)
    assert w.shape == (n,)
    assert w[n // 2] == 0.5
    i = np.arange(1, n)
    s = w[i] + w[n - i]
    assert np.all(np.abs(s - 1.0) <= 2 * np.spacing(1.0)), float(np.max(np.abs(s - 1.0)))
    d = np.diff(w)
    unsat = (w[:-1] < 1.0) & (w[1:] < 1.0)  # exp(-g (i - n/2)) below 2^-53: w rounds to 1
    assert np.all(d[unsat] > 0.0)
    assert np.all(d >= 0.0)
    assert np.max(_ulps(w, expit(g * (np.arange(n) - n / 2.0)))) <= 4


@pytest.mark.parametrize("n", [1, 7, 33, 513])
def test_wdtw_weights_closed_forms_odd_n(n):
    # odd n: the centre n/2 falls between two samples, so w[(n-1)/2] and
    # w[(n+1)/2] are symmetric about 1/2: w[i] + w[n - i] = 1 still holds
    # for i in 1..n-1, and g = 0 makes every weight exactly 1/2.
    from scipy.special import expit

    g = 0.2
    w = oracle.wdtw_weights(n, g)
    if n > 1:
        i = np.arange(1, n)
        assert np.all(np.abs(w[i] + w[n - i] - 1.0) <= 2 * np.spacing(1.0))
        assert w[(n - 1) // 2] < 0.5 < w[(n + 1) // 2]
    assert np.max(_ulps(w, expit(g * (np.arange(n) - n / 2.0)))) <= 4
    assert np.all(oracle.wdtw_weights(n, 0.0) == 0.5)


def test_r0_is_euclidean_bitwise():
    rng = np.random.default_rng(3)
    for t in (1, 2, 17, 128):
        x, y = rng.standard_normal((2, t))
        d_dtw = oracle.metric_distance(x, y, "dtw", r=0.0)
        d_euc = oracle.metric_distance(x, y, "euclidean")
        assert d_dtw == d_euc


def test_r1_is_textbook_dtw():
    rng = np.random.default_rng(4)
    for lx, ly in [(5, 5), (9, 4), (3, 11), (32, 32)]:
        x = rng.standard_normal(lx)
        y = rng.standard_normal(ly)
        got = oracle.metric_distance(x, y, "dtw", r=1.0) ** 2
        assert got == pytest.approx(textbook_dtw(x, y), rel=1e-12)


def test_invariants_identity_symmetry_monotone():
    rng = np.random.default_rng(5)
    X = rw(rng, 6, 40)
    for a in range(6):
        assert oracle.metric_distance(X[a], X[a], "dtw", r=0.3) == 0.0
        for b in range(6):
            assert oracle.metric_distance(X[a], X[b], "dtw", r=0.2) == oracle.metric_distance(
                X[b], X[a], "dtw", r=0.2)
            prev = INF
            for r in R_GRID:
                d = oracle.metric_distance(X[a], X[b], "dtw", r=r)
                assert d <= prev
                prev = d


def test_wdtw_g0_is_half_dtw_and_asymmetric():
    rng = np.random.default_rng(6)
    X = rw(rng, 4, 30)
    for a in range(4):
        for b in range(4):
            dw = oracle.metric_distance(X[a], X[b], "wdtw", r=0.4, g=0.0)
            dd = oracle.metric_distance(X[a], X[b], "dtw", r=0.4)
            assert dw * dw == pytest.approx(0.5 * dd * dd, rel=1e-15)
    asym = [abs(oracle.metric_distance(X[a], X[b], "wdtw", r=0.4, g=0.3)
                - oracle.metric_distance(X[b], X[a], "wdtw", r=0.4, g=0.3))
            for a in range(4) for b in range(4)]
    assert max(asym) > 0.0


def test_wdtw_pairwise_weights_sized_by_wider_array():
    # WeightedDtwMetric.reset sizes the weights by max(X.shape[2], Y.shape[2])
    # (distance/_elastic.pyx:3336); the brute force uses expit weights of that
    # size, so the vector is pinned end to end (value, centre and size).
    from scipy.special import expit

    rng = np.random.default_rng(61)
    X = rw(rng, 3, 30)
    Y = rw(rng, 4, 41)
    g, r = 0.15, 0.3
    D = oracle.pairwise(X, Y, metric="wdtw", r=r, g=g)
    wts = expit(g * (np.arange(41) - 41 / 2.0))
    for a in range(3):
        for b in range(4):
            want = math.sqrt(brute_dtw(X[a], Y[b], r, wts))
            assert D[a, b] == pytest.approx(want, rel=1e-13)
    # the smaller width would centre the weights elsewhere: values differ
    wts30 = expit(g * (np.arange(41) - 30 / 2.0))
    assert abs(D[0, 0] - math.sqrt(brute_dtw(X[0], Y[0], r, wts30))) > 1e-6 * D[0, 0]


def test_argmin_docstring_known_answer():
    X = np.array([[1, 2, 3, 4], [10, 1, 2, 3]], float)
    Y = np.array([[1, 2, 11, 2], [2, 4, 6, 7], [10, 11, 2, 3]], float)
    idx, dist = oracle.argmin(X, Y, k=2, metric="euclidean")
    assert idx.tolist() == [[0, 1], [1, 2]]
    np.testing.assert_allclose(dist, [[8.24621125, 4.79583152], [10.24695077, 10.0]], atol=1e-8)


@pytest.mark.parametrize("metric", ["dtw", "euclidean", "wdtw"])
def test_argmin_is_k_smallest(metric):
    rng = np.random.default_rng(7)
    X = rw(rng, 7, 24)
    Y = rw(rng, 19, 24)
    D = oracle.pairwise(X, Y, metric=metric, r=0.2, g=0.1)
    for k in (1, 3, 5):
        idx, dist = oracle.argmin(X, Y, k=k, metric=metric, r=0.2, g=0.1)
        for i in range(7):
            order = np.argsort(D[i], kind="stable")[:k]
            assert sorted(idx[i].tolist()) == sorted(order.tolist())
            np.testing.assert_array_equal(np.sort(dist[i]), np.sort(D[i][order]))
            ri, rd = oracle.argmin_replay(D[i], k)
            assert ri.tolist() == idx[i].tolist() and rd.tolist() == dist[i].tolist()


def test_argmin_ties_first_index():
    X = np.zeros((1, 4))
    Y = np.zeros((5, 4))
    idx, _ = oracle.argmin(X, Y, k=1, metric="dtw", r=0.5)
    assert idx.tolist() == [[0]]


def test_singleton_and_paired_orientation():
    rng = np.random.default_rng(8)
    X = rw(rng, 5, 20)
    Y = rw(rng, 5, 20)
    S = oracle.pairwise(X, None, metric="wdtw", r=0.5, g=0.3)
    for i in range(5):
        assert S[i, i] == 0.0
        for j in range(i + 1, 5):
            d = oracle.metric_distance(X[i], X[j], "wdtw", r=0.5, g=0.3)
            assert S[i, j] == d and S[j, i] == d
    P = oracle.paired(X, Y, metric="wdtw", r=0.5, g=0.3)
    for i in range(5):
        assert P[i] == oracle.metric_distance(Y[i], X[i], "wdtw", r=0.5, g=0.3)


@pytest.mark.parametrize("metric", ["dtw", "euclidean", "wdtw"])
def test_subsequence_brute_force(metric):
    rng = np.random.default_rng(9)
    X = rw(rng, 4, 30)
    shapelets = [rng.standard_normal(L) for L in (1, 3, 7, 12)]
    dist, idx = oracle.subsequence(shapelets, X, metric=metric, r=0.25, g=0.2)
    wts = oracle.wdtw_weights(30, 0.2) if metric == "wdtw" else None
    for i in range(4):
        for j, s in enumerate(shapelets):
            L = len(s)
            best, bi = INF, 0
            for w0 in range(30 - L + 1):
                win = X[i, w0:w0 + L]
                if metric == "euclidean":
                    c = 0.0
                    for q in range(L):
                        v = win[q] - s[q]
                        c += v * v
                else:
                    c = brute_dtw(s, win, 0.25, wts)
                if c < best:
                    best, bi = c, w0
            assert idx[i, j] == bi
            assert dist[i, j] == math.sqrt(best)


def test_subsequence_match_known_answer():
    # tests/wildboar/distance/test_distance.py:1047-1060 (inline data): the
    # euclidean windows of x within threshold 2 of y are [0, 1, 6] and [1, 6].
    x = np.array([1, 2, 3, 4], float)
    ys = [np.array([1, 2, 3, 3, 3, 3, 1, 2, 3, 4], float),
          np.array([10, 2, 3, 3, 3, 3, 1, 2, 3, 4], float)]
    for y, want in zip(ys, ([0, 1, 6], [1, 6])):
        prof = [oracle.metric_distance(x, y[s:s + 4], "euclidean") for s in range(7)]
        assert [s for s, d in enumerate(prof) if d <= 2] == want
        dist, idx = oracle.subsequence([x], y[None, :], metric="euclidean")
        assert idx[0, 0] == 6 and dist[0, 0] == 0.0


def test_subsequence_euclidean_is_dtw_r0_and_pairwise_transposed():
    rng = np.random.default_rng(10)
    X = rw(rng, 5, 25)
    S = rw(rng, 3, 9)
    de, ie = oracle.subsequence(list(S), X, metric="euclidean")
    dd, id_ = oracle.subsequence(list(S), X, metric="dtw", r=0.0)
    np.testing.assert_array_equal(de, dd)
    np.testing.assert_array_equal(ie, id_)
    Z = rw(rng, 3, 25)
    dsub, _ = oracle.subsequence(list(Z), X, metric="dtw", r=0.3)
    np.testing.assert_array_equal(dsub, oracle.pairwise(Z, X, metric="dtw", r=0.3).T)


def test_gradient_finite_differences_and_closed_form():
    rng = np.random.default_rng(11)
    x = rw(rng, 1, 24)[0]
    y = rw(rng, 1, 24)[0]
    D, gx, gy = oracle.dtw_grad(x, y, 0.2)
    assert D == oracle.metric_distance(x, y, "dtw", r=0.2)
    eps = 1e-6
    for i in range(24):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        fd = (oracle.metric_distance(xp, y, "dtw", r=0.2)
              - oracle.metric_distance(xm, y, "dtw", r=0.2)) / (2 * eps)
        assert fd == pytest.approx(gx[i], rel=1e-5, abs=1e-7)
        yp, ym = y.copy(), y.copy()
        yp[i] += eps
        ym[i] -= eps
        fd = (oracle.metric_distance(x, yp, "dtw", r=0.2)
              - oracle.metric_distance(x, ym, "dtw", r=0.2)) / (2 * eps)
        assert fd == pytest.approx(gy[i], rel=1e-5, abs=1e-7)
    D0, g0x, g0y = oracle.dtw_grad(x, y, 0.0)
    np.testing.assert_allclose(g0x, (x - y) / D0, rtol=1e-14)
    np.testing.assert_allclose(g0y, -(x - y) / D0, rtol=1e-14)


def test_mapping_tie_rule_prefers_diag_then_up():
    A = np.array([[0.0, 1.0], [1.0, 0.0]])
    pi, pj = oracle.mapping(A)
    assert list(zip(pi, pj)) == [(0, 0), (1, 1)]
    A = np.array([[0.0, 0.0, 5.0], [0.0, 0.0, 0.0]])
    pi, pj = oracle.mapping(A)
    # from (1,2): diag (0,1)=0 ties up (0,2)=5? no: diag 0 < up 5, left (1,1)=0 ties diag -> diag
    assert list(zip(pi, pj))[-2] == (0, 1)


def _window_costs(s, x, metric, r, wts):
    """Squared cost of every window of x against s, from the definitions."""
    L = len(s)
    out = []
    for w0 in range(len(x) - L + 1):
        win = x[w0:w0 + L]
        if metric == "euclidean":
            c = 0.0
            for q in range(L):
                v = win[q] - s[q]
                c += v * v
        else:
            c = brute_dtw(s, win, r, wts)
        out.append(c)
    return out


def test_subsequence_matches_known_answer():
    # tests/wildboar/distance/test_distance.py:1047-1060: windows [0, 1, 6]
    # and [1, 6]; the distances follow by hand: |(0,0,0,-1)| = 1,
    # |(1,1,0,-1)| = sqrt(3), the exact copy 0, |(9,0,0,-1)| > 2.
    s = np.array([1, 2, 3, 4], float)
    x0 = np.array([1, 2, 3, 3, 3, 3, 1, 2, 3, 4], float)
    x1 = np.array([10, 2, 3, 3, 3, 3, 1, 2, 3, 4], float)
    idx, dist = oracle.subsequence_matches(s, x0, 2.0, "euclidean")
    assert idx.tolist() == [0, 1, 6]
    assert dist.tolist() == [1.0, math.sqrt(3.0), 0.0]
    idx, dist = oracle.subsequence_matches(s, x1, 2.0, "euclidean")
    assert idx.tolist() == [1, 6]
    # the bound is compared squared: a negative threshold acts as its magnitude
    assert oracle.subsequence_matches(s, x0, -2.0, "euclidean")[0].tolist() == [0, 1, 6]
    assert oracle.subsequence_matches(s, x0, float("nan"), "euclidean")[0].size == 0


@pytest.mark.parametrize("metric", ["dtw", "euclidean", "wdtw"])
def test_subsequence_matches_brute_force(metric):
    rng = np.random.default_rng(21)
    x = rw(rng, 1, 40)[0]
    for L in (1, 4, 9, 40):
        s = rng.standard_normal(L)
        wts = oracle.wdtw_weights(40, 0.1) if metric == "wdtw" else None
        costs = _window_costs(s, x, metric, 0.3, wts)
        for thr in (INF, 0.0, math.sqrt(np.median(costs)), math.sqrt(min(costs))):
            idx, dist = oracle.subsequence_matches(s, x, thr, metric, r=0.3, g=0.1)
            want = [w for w, c in enumerate(costs) if c <= thr * thr]
            assert idx.tolist() == want, (L, thr)
            assert dist.tolist() == [math.sqrt(costs[w]) for w in want]
        # threshold inf: every window; its minimum is the subsequence distance
        idx, dist = oracle.subsequence_matches(s, x, INF, metric, r=0.3, g=0.1)
        sd, si = oracle.subsequence([s], x[None, :], metric=metric, r=0.3, g=0.1)
        assert dist.min() == sd[0, 0] and int(np.argmin(dist)) == si[0, 0]


@pytest.mark.parametrize("metric", ["dtw", "euclidean", "wdtw"])
def test_argmin_subsequence_is_heap_of_window_distances(metric):
    rng = np.random.default_rng(22)
    X = rw(rng, 5, 33)
    shapelets = [rng.standard_normal(L) for L in (1, 5, 8, 12, 33)]
    wts = oracle.wdtw_weights(33, 0.2) if metric == "wdtw" else None
    for k in (1, 2, 5):
        idx, dist = oracle.argmin_subsequence(shapelets, X, k=k, metric=metric, r=0.25, g=0.2)
        for i, s in enumerate(shapelets):
            d = np.sqrt(_window_costs(s, X[i], metric, 0.25, wts))
            kk = min(k, d.size)
            np.testing.assert_array_equal(np.sort(dist[i, :kk]), np.sort(d)[:kk])
            ri, rd = oracle.argmin_replay(d, kk)
            assert idx[i, :kk].tolist() == ri.tolist() and dist[i, :kk].tolist() == rd.tolist()


def test_argmin_subsequence_euclidean_is_dtw_r0():
    rng = np.random.default_rng(23)
    X = rw(rng, 4, 30)
    shapelets = [rng.standard_normal(L) for L in (3, 6, 10, 2)]
    a = oracle.argmin_subsequence(shapelets, X, k=4, metric="euclidean")
    b = oracle.argmin_subsequence(shapelets, X, k=4, metric="dtw", r=0.0)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


# ------------------------------------------------------ derivative / amercing
def brute_adtw(x, y, r, p):
    """Full-matrix ADTW from the definition: the dtw_distance band, row 0 a
    plain prefix sum, later cells min(up + p, left + p, diag) + (x_i - y_j)^2."""
    lx, ly = len(x), len(y)
    w = band_w(lx, ly, r)
    hi = max(0, lx - ly) + w - 1
    lo = -(max(0, ly - lx) + w - 1)
    C = [[INF] * ly for _ in range(lx)]
    for i in range(lx):
        for j in range(ly):
            if not (lo <= i - j <= hi):
                continue
            v = float(x[i]) - float(y[j])
            if i == 0:
                C[i][j] = v * v if j == 0 else C[0][j - 1] + v * v
                continue
            up = C[i - 1][j] + p
            left = (C[i][j - 1] if j > 0 else INF) + p
            diag = C[i - 1][j - 1] if j > 0 else INF
            m = left if left < up else up
            m = diag if diag < m else m
            C[i][j] = m + v * v
    return C[lx - 1][ly - 1]


def test_average_slope_known_answers():
    # by hand: ((2-1) + (4-1)/2)/2 = 1.25, ((4-2) + (7-2)/2)/2 = 2.25, ...
    assert oracle.average_slope([1, 2, 4, 7, 11]).tolist() == [1.25, 2.25, 3.25]
    assert oracle.average_slope(3.0 + 0.5 * np.arange(9)).tolist() == [0.5] * 7
    i = np.arange(1, 11)
    assert oracle.average_slope(np.arange(12.0) ** 2).tolist() == (2.0 * i - 0.5).tolist()
    assert oracle.average_slope([1.0, 2.0]).size == 0


def test_adtw_brute_force_and_closed_forms():
    rng = np.random.default_rng(31)
    for lx in range(1, 8):
        for ly in range(1, 8):
            x = rng.standard_normal(lx)
            y = rng.standard_normal(ly)
            for r in (0.0, 0.4, 1.0):
                band = oracle.compute_r(min(lx, ly), r)
                for p in (0.0, 0.25, 1.5, INF):
                    assert oracle.adtw_distance(x, y, band, p) == brute_adtw(x, y, r, p), (lx, ly, r, p)
                # penalty 0 is DTW, bit for bit
                assert oracle.adtw_distance(x, y, band, 0.0) == oracle.dtw_distance(x, y, band)
    # penalty +inf at equal lengths: only the diagonal, the squared euclidean sum
    for t in (1, 5, 33):
        x = rng.standard_normal(t)
        y = rng.standard_normal(t)
        c = 0.0
        for a, b in zip(x, y):
            v = a - b
            c += v * v
        for r in (0.0, 0.3, 1.0):
            assert oracle.adtw_distance(x, y, oracle.compute_r(t, r), INF) == c
    X = rw(rng, 4, 20)
    D = oracle.pairwise(X, X[::-1].copy(), metric="adtw", r=0.5, p=0.7)
    assert D[1, 2] == math.sqrt(brute_adtw(X[1], X[::-1][2], 0.5, 0.7))


def test_ddtw_is_dtw_on_slopes_with_the_band_quirk():
    rng = np.random.default_rng(32)
    X = rw(rng, 5, 12)
    Y = rw(rng, 6, 12)
    r = 0.25  # band 3 by the series' lengths (pairwise), 2 by the slopes' (argmin)
    assert oracle.compute_r(12, r) == 3 and oracle.compute_r(10, r) == 2
    D = oracle.pairwise(X, Y, metric="ddtw", r=r)
    for i in range(5):
        for j in range(6):
            want = oracle.dtw_distance(oracle.average_slope(X[i]), oracle.average_slope(Y[j]), 3)
            assert D[i, j] == math.sqrt(want)
    idx, dist = oracle.argmin(X, Y, k=3, metric="ddtw", r=r)
    for i in range(5):
        row = np.array([math.sqrt(oracle.dtw_distance(oracle.average_slope(X[i]),
                                                      oracle.average_slope(Y[j]), 2))
                        for j in range(6)])
        ri, rd = oracle.argmin_replay(row, 3)
        assert idx[i].tolist() == ri.tolist() and dist[i].tolist() == rd.tolist()
    # WDDTW: WDTW on the slopes, weights sized by the slope length
    W = oracle.pairwise(X, Y, metric="wddtw", r=r, g=0.3)
    wts = oracle.wdtw_weights(10, 0.3)
    want = oracle.dtw_distance(oracle.average_slope(X[2]), oracle.average_slope(Y[4]), 3,
                               weights=wts)
    assert W[2, 4] == math.sqrt(want)
    # series shorter than 3 give 0; argmin never accepts them (zero outputs)
    assert (oracle.pairwise(X[:, :2], Y[:, :2], metric="ddtw") == 0).all()
    idx, dist = oracle.argmin(X[:, :2], Y[:, :2], k=2, metric="ddtw")
    assert (idx == 0).all() and (dist == 0).all()
    # straight lines: constant slopes b and d, the diagonal path of t - 2 cells
    t = 40
    i = np.arange(t, dtype=float)
    d = oracle.metric_distance(1.0 + 0.75 * i, -2.0 + 0.25 * i, "ddtw", r=0.1)
    assert d == pytest.approx(0.5 * math.sqrt(t - 2), rel=1e-15)


def test_mm_dtw_average_fixed_points():
    rng = np.random.default_rng(33)
    s = rw(rng, 1, 30)[0]
    # four identical series are a fixed point: every path is the diagonal and
    # 4 s / 4 == s
    mean, cost = oracle.mm_dtw_average(np.stack([s] * 4), s, r=0.2)
    np.testing.assert_array_equal(mean, s)
    assert cost == 0.0
    # r = 0 leaves only the diagonal: one step from any start averages the
    # series themselves
    X = np.stack([s, s, s, s])
    mean, _ = oracle.mm_dtw_average(X, s + 1e-3, r=0.0, max_epoch=1)
    np.testing.assert_array_equal(mean, s)
    # the mapping of identical series is the diagonal; a one-step shift puts
    # two cells on the first and last rows
    pi, pj = oracle.mapping(oracle.cost_matrix(s, s, oracle.warp_size(30, 0.2)))
    assert pi.tolist() == list(range(30)) and pj.tolist() == list(range(30))
    pi, pj = oracle.mapping(oracle.cost_matrix(s, np.r_[s[0], s[:-1]], 30))
    assert (pi[0], pj[0], pi[-1], pj[-1]) == (0, 0, 29, 29) and len(pi) == 31


# ------------------------------------------------------------- lower bounds
def test_envelope_deque_is_the_windowed_min_max():
    rng = np.random.default_rng(34)
    for trial in range(600):
        n = int(rng.integers(1, 40))
        r = int(rng.integers(0, n))
        x = rng.integers(0, 4, n).astype(float) if trial % 2 else rng.standard_normal(n)
        lo, up = oracle.dtw_envelop(x, r)
        assert lo.tolist() == [x[max(0, k - r):k + r + 1].min() for k in range(n)], (n, r)
        assert up.tolist() == [x[max(0, k - r):k + r + 1].max() for k in range(n)], (n, r)


def test_lb_keogh_known_answer_and_bound():
    assert oracle.lb_keogh([0.0, 3.0, 0.0], [1.0] * 3, [2.0] * 3) == math.sqrt(3.0)
    assert oracle.lb_keogh([1.5, 1.0, 2.0], [1.0] * 3, [2.0] * 3) == 0.0
    rng = np.random.default_rng(35)
    X = rw(rng, 20, 50)
    for r in (0.05, 0.2, 1.0):
        warp = max(math.floor(50 * r), 1)
        warp = warp - 1 if warp == 50 else warp
        for i in range(0, 20, 2):
            lo, up = oracle.dtw_envelop(X[i + 1], warp)
            lb = oracle.lb_keogh(X[i], lo, up)
            assert lb <= oracle.metric_distance(X[i], X[i + 1], "dtw", r=r) + 1e-12
    # LB_Kim (squared) bounds the squared unconstrained DTW cost
    K = oracle.lb_kim(X[:10], X[10:])
    for i in range(10):
        for j in range(10):
            assert K[i, j] <= oracle.metric_distance(X[j], X[10 + i], "dtw", r=1.0) ** 2 + 1e-12


def test_alignment_matrix():
    # unweighted: the alignment is the distance cost matrix; weighted: a
    # brute-force matrix with w[|i - j|] in every row (row 0 included)
    rng = np.random.default_rng(36)
    for lx, ly in ((1, 1), (5, 9), (9, 5), (12, 12)):
        x = rng.standard_normal(lx)
        y = rng.standard_normal(ly)
        for band in (1, 2, 4, max(lx, ly)):
            np.testing.assert_array_equal(oracle.dtw_alignment(x, y, band),
                                          oracle.cost_matrix(x, y, band))
            w = rng.uniform(0.2, 1.5, max(lx, ly))
            A = oracle.dtw_alignment(x, y, band, w)
            hi = max(0, lx - ly) + band - 1
            lo = -(max(0, ly - lx) + band - 1)
            B = np.full((lx, ly), np.inf)
            for i in range(lx):
                for j in range(ly):
                    if not lo <= i - j <= hi:
                        continue
                    c = (x[i] - y[j]) * (x[i] - y[j]) * w[abs(i - j)]
                    if i or j:
                        m = min(B[i - 1, j] if i else np.inf, B[i, j - 1] if j else np.inf,
                                B[i - 1, j - 1] if i and j else np.inf)
                        c = m + c
                    B[i, j] = c
            np.testing.assert_array_equal(A, B)


def test_derivative_and_amercing_subsequence():
    rng = np.random.default_rng(37)
    X = rw(rng, 4, 50)
    pats = [X[1, 7:19].copy(), rng.standard_normal(6), X[2, 0:2].copy(), X[3, 10:50].copy()]
    dd, di = oracle.subsequence(pats, X, metric="ddtw", r=0.25)
    for i in range(4):
        for j, s in enumerate(pats):
            L = len(s)
            if L < 3:
                assert dd[i, j] == 0.0
                continue
            # DTW of the slopes, band of the pattern's own length (the window's
            # slope is the series' slope over the same span)
            band = oracle.compute_r(L, 0.25)
            xs = oracle.average_slope(X[i])
            ss = oracle.average_slope(s)
            costs = [oracle.dtw_distance(ss, xs[w:w + L - 2], band) for w in range(50 - L + 1)]
            assert dd[i, j] == math.sqrt(min(costs)) and di[i, j] == int(np.argmin(costs))
    assert dd[1, 0] == 0.0 and di[1, 0] == 7
    # adtw: penalty 0 is dtw, penalty inf at equal lengths the euclidean scan
    a0 = oracle.subsequence(pats, X, metric="adtw", r=0.3, p=0.0)
    dt = oracle.subsequence(pats, X, metric="dtw", r=0.3)
    np.testing.assert_array_equal(a0[0], dt[0])
    np.testing.assert_array_equal(a0[1], dt[1])
    ai = oracle.subsequence(pats, X, metric="adtw", r=0.3, p=math.inf)
    eu = oracle.subsequence(pats, X, metric="euclidean")
    np.testing.assert_array_equal(ai[0], eu[0])
    np.testing.assert_array_equal(ai[1], eu[1])


def test_ssg_average_on_the_diagonal_is_the_convex_step():
    # r = 0 gives a warp size of 1, the diagonal path at equal lengths: each
    # step is mean <- mean - lr * (mean - w * x_o) (distance/dtw.py:618-633),
    # series in RandomState's shuffled order, the epoch-0 rate compounded by
    # decay**i
    X = np.cumsum(np.random.default_rng(15).standard_normal((6, 20)), axis=1)
    mean0 = X[2].copy()
    got, cost = oracle.ssg_dtw_average(X, mean0, r=0.0, max_epoch=4, max_stable=10,
                                       learning_rate=0.3, decay=0.8, random_state=7)
    rs = np.random.RandomState(7)
    mean, lr, order = mean0.copy(), 0.3, np.arange(6)
    best, best_cost = None, np.inf
    for epoch in range(4):
        rs.shuffle(order)
        for i, o in enumerate(order):
            mean = mean - lr * (mean - X[o] * 1.0)
            if epoch == 0:
                lr = 0.8**i * lr
        c = np.mean(oracle.pairwise(mean[None], X, metric="dtw", r=0.0)[0])
        if c < best_cost:
            best, best_cost = mean.copy(), c
    np.testing.assert_array_equal(got, best)
    assert cost == best_cost
