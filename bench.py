"""Benchmark: wildboar's DTW distance path on one MI355X.

Headline workload (one step): pairwise_distance(X, Y, metric="dtw", r=0.1)
with X and Y 2000 x 512 float32 z-normalised Gaussian random walks (seeds 10,
11), inputs resident in HBM, FP32 cells, float64 2000 x 2000 output -- the
full-matrix configuration the metric is quoted on.  The step runs through the
C ABI (wbx_dev_pairwise) on torch's current stream.

Printed JSON line (stdout):
  value         algorithmic band cells per second over the timed steps
  roofline      the DTW kernel's cells/s (HIP events on its launch stream)
                against the VALU min-plus peak of 3 instructions per cell
  cpu_baseline  the float64 C restatement of the reference, all host cores,
                on a bounded slice of the headline workload
  accuracy      FP32 max relative error against the float64 oracle, 1-NN
                labels and indices of all 2000 rows against the oracle's
                argmin, the plain one-pair-per-block kernel on a slice of the
                same workload
  pairwise_legs the other legs of the same pairwise workload: dtw r = 0.05
                and r = 1.0, wdtw g = 0.05, euclidean (float32), dtw r = 0.1
                and r = 1.0 in float64; rates and oracle checks on a few rows
  batched       50,000 EOS-padded pairs of uneven length (64..2048) with
                per-pair r in [0.05, 0.2]: the persistent strip launch in a
                hipGraph replay against the plain per-pair kernel, its own
                roofline, and the float64 strips
  train         differentiable DTW, 4096 x 64 pairs of length 256, r = 0.1:
                forward, forward + backward, and 200 barycenter steps
  subsequence   pairwise_subsequence_distance, 4096 shapelets (lengths
                16..128) over 10000 random-walk series of length 1024, dtw
                r = 0.1 and euclidean, with exact indices
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DTW GCUPS (cells/s) at r=0.1, batched vs plain on 1 GPU; fwd+bwd time; max rel err"
# MI355X published FP32 vector peak 157.3 TFLOP/s = 256 CUs x 128 FMA lanes x 2 flop
# x 2.4 GHz, i.e. 78.6e12 VALU lane-instructions/s; a DTW cell needs 3 of them
# (v_sub_f32, v_min3_f32, v_fmac_f32).
PEAK_CELLS = 256 * 128 * 2.4e9 / 3.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_walks(seed, n, t):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, t)), axis=1)
    return (x - x.mean(axis=1, keepdims=True)) / x.std(axis=1, keepdims=True)


def band_cells(lx, ly, r):
    """Exact cell count of the reference's dtw_distance loops (_elastic.pyx:869-940)."""
    w = max(math.floor(min(lx, ly) * r), 1)
    min_len = max(0, lx - ly)
    max_len = max(0, ly - lx) + w
    cells = min(ly, max_len)
    for i in range(1, lx):
        cells += min(ly, i + max_len) - max(0, i - min_len - w + 1)
    return cells


def timed(torch, fn, reps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


# --------------------------------------------------------------------- batched
def uneven_pairs(n, t_pad, seed_len=30, seed_data=31):
    """EOS-padded random-walk pairs: lengths U[64, t_pad], r U[0.05, 0.2],
    z-normalised over each series' own length, plus Laplace(0, 1) spikes with
    probability 0.05 per timestep."""
    from wildboar_amd.variable_len import EOS

    rng = np.random.default_rng(seed_len)
    lx = rng.integers(64, t_pad + 1, n).astype(np.int32)
    ly = rng.integers(64, t_pad + 1, n).astype(np.int32)
    r = rng.uniform(0.05, 0.2, n)
    drng = np.random.default_rng(seed_data)
    cols = np.arange(t_pad)

    def side(lens):
        out = np.empty((n, t_pad), np.float32)
        for c0 in range(0, n, 4096):
            L = lens[c0:c0 + 4096].astype(np.float64)[:, None]
            m = cols[None, :] < L
            wlk = np.cumsum(drng.standard_normal((m.shape[0], t_pad), dtype=np.float32), axis=1,
                            dtype=np.float64)
            mu = (wlk * m).sum(1, keepdims=True) / L
            sd = np.sqrt((((wlk - mu) * m) ** 2).sum(1, keepdims=True) / L)
            z = (wlk - mu) / sd
            spike = drng.random(z.shape, dtype=np.float32) < 0.05
            z[spike] += drng.laplace(0.0, 1.0, int(spike.sum()))
            out[c0:c0 + 4096] = np.where(m, z, EOS).astype(np.float32)
        return out

    return side(lx), side(ly), lx, ly, r


def run_batched(torch, dev, n_pairs, t_pad, reps):
    import oracle
    from wildboar_amd.batch import DtwBatch

    t0 = time.perf_counter()
    X, Y, lx, ly, r = uneven_pairs(n_pairs, t_pad)
    log(f"batched: data ready in {time.perf_counter() - t0:.1f} s")
    Xd = torch.from_numpy(X).to(dev)
    Yd = torch.from_numpy(Y).to(dev)
    side = torch.cuda.Stream()
    res = {}
    tuned = None
    tuned_dt = None
    for plain in (False, True, "padded"):
        plan = DtwBatch(lx, ly, t_pad, r, precision="fp32", plain=plain)
        out = torch.empty(n_pairs, dtype=torch.float64, device=dev)

        def replay():
            with torch.cuda.stream(side):
                plan.run(Xd, Yd, out=out, graph=True)

        replay()
        replay()
        dt = timed(torch, replay, 1 if plain is not False else reps)
        res[plain] = (dt, out.cpu().numpy(), plan.cells, plan.n_strip, plan.n_plain)
        if plain is False:
            # the autotuner's per-bucket launch shapes, timed the same way
            with torch.cuda.stream(side):
                tuned_ms = plan.autotune(Xd, Yd, reps=5)
            tuned_kernel_ms = tuned_ms
            tuned = [{"R": bk["R"], "pairs": bk["pairs"], "waves": bk["waves"],
                      "stream": bk["stream"], "ms": round(bk["ms"], 4)} for bk in plan.buckets()]
            replay()
            replay()
            tuned_dt = timed(torch, replay, reps)
            assert np.array_equal(out.cpu().numpy(), res[False][1])
            log(f"batched: autotuned {tuned_dt * 1e3:.3f} ms (strip launch alone {tuned_ms:.3f} ms) "
                f"vs default launch {dt * 1e3:.3f} ms")
        plan.close()
    # the autotuned launch shape is the plan's own choice: its replay time is
    # the batched result
    dt, got, cells, n_strip, n_plainp = res[False]
    default_dt = dt
    dt = tuned_dt
    pdt = res[True][0]
    ppdt = res["padded"][0]
    # the float64 strips (FP64 mode: the reference's arithmetic, bit-exact)
    plan64 = DtwBatch(lx, ly, t_pad, r, precision="fp64")
    X64d, Y64d = Xd.double(), Yd.double()
    out64 = torch.empty(n_pairs, dtype=torch.float64, device=dev)

    def replay64():
        with torch.cuda.stream(side):
            plan64.run(X64d, Y64d, out=out64, graph=True)

    replay64()
    replay64()
    dt64 = timed(torch, replay64, 3)
    got64 = out64.cpu().numpy()
    plan64.close()
    assert np.array_equal(res["padded"][1], res[True][1])
    pick = np.random.default_rng(0).choice(n_pairs, 200, replace=False)
    want = np.array([oracle.metric_distance(X[p, :lx[p]].astype(np.float64),
                                            Y[p, :ly[p]].astype(np.float64), "dtw", r=r[p])
                     for p in pick])
    err = float(np.max(np.abs(got[pick] - want) / want))
    mism64 = int((got64[pick] != want).sum())
    kern_s = tuned_kernel_ms / 1e3
    out = {
        "workload": f"{n_pairs} pairs, lengths U[64,{t_pad}] padded to {t_pad} with EOS, "
                    "r U[0.05,0.2] per pair, random walks + 5% Laplace spikes, FP32",
        "cells": int(cells),
        "ms": dt * 1e3,
        "default_launch_ms": default_dt * 1e3,
        "value": cells / dt,
        "unit": "cells/s",
        "frac_of_peak": cells / dt / PEAK_CELLS,
        "plain_ms": pdt * 1e3,
        "plain_value": cells / pdt,
        "speedup_vs_plain": pdt / dt,
        # the plain kernel walking every pair over the padded length (the
        # survey's K1 definition), same results
        "plain_padded_ms": ppdt * 1e3,
        "speedup_vs_plain_padded": ppdt / dt,
        "strip_pairs": int(n_strip),
        "plain_pairs": int(n_plainp),
        "max_rel_err_200_pairs": err,
        "autotuned_ms": tuned_dt * 1e3,
        "autotuned_buckets": tuned,
        "graph": True,
        # the strip launch alone (HIP events around it on its stream, by the
        # autotuner): the batched path's own roofline
        "roofline": {
            "bound": "valu",
            "achieved": cells / kern_s * 6 / 1e12,
            "peak": 157.3,
            "unit": "TFLOP/s",
            "frac": cells / kern_s * 6 / 157.3e12,
            "traffic": None,
            "kernel": "k_strip_batch<float>",
            "kernel_ms": kern_s * 1e3,
            "achieved_cells_per_s": cells / kern_s,
            "peak_cells_per_s": PEAK_CELLS,
        },
        "fp64": {"ms": dt64 * 1e3, "value": cells / dt64, "unit": "cells/s",
                 "mismatch_200_pairs": mism64,
                 "note": "FP64 mode (the reference's float64 arithmetic) on the float32 data "
                         "widened, float64 strips; bit-exact against the oracle"},
    }
    log(f"batched: {dt * 1e3:.3f} ms, {cells / dt / 1e12:.2f} Tcells/s (strip launch "
        f"{kern_s * 1e3:.3f} ms, {cells / kern_s / 1e12:.2f} Tcells/s); plain {pdt * 1e3:.2f} ms "
        f"(x{pdt / dt:.1f}); plain over the padded length {ppdt * 1e3:.2f} ms (x{ppdt / dt:.1f}); "
        f"max rel err {err:.2e}; fp64 {dt64 * 1e3:.3f} ms ({cells / dt64 / 1e12:.2f} Tcells/s, "
        f"{mism64} mismatches)")
    return out


# ------------------------------------------------------- pairwise metric legs
def run_pairwise_legs(torch, dev, X64, Y64, reps):
    """The other legs of the pairwise 2000 x 2000 x 512 workload on the same
    inputs: dtw r = 0.05 and r = 1.0, wdtw g = 0.05 (r = 0.1) and euclidean in
    float32, and dtw r = 0.1 / r = 1.0 in float64 (the reference's
    arithmetic).  Each leg: time per call, algorithmic cells/s, the kernel
    family that ran, and a few rows against the float64 oracle (relative
    error for float32, mismatches for float64)."""
    import oracle
    from wildboar_amd import _lib, device

    n, t = X64.shape
    X32 = torch.from_numpy(X64.astype(np.float32)).to(dev)
    Y32 = torch.from_numpy(Y64.astype(np.float32)).to(dev)
    Xd = torch.from_numpy(X64).to(dev)
    Yd = torch.from_numpy(Y64).to(dev)
    Xr = X64.astype(np.float32).astype(np.float64)
    Yr = Y64.astype(np.float32).astype(np.float64)
    out = torch.empty((n, n), dtype=torch.float64, device=dev)
    legs = [("fp32_dtw_r0.05", "dtw", 0.05, 0.05, "fp32"), ("fp32_dtw_r1.0", "dtw", 1.0, 0.05, "fp32"),
            ("fp32_wdtw_r0.1_g0.05", "wdtw", 0.1, 0.05, "fp32"),
            ("fp32_euclidean", "euclidean", 1.0, 0.05, "fp32"),
            ("fp64_dtw_r0.1", "dtw", 0.1, 0.05, "fp64"), ("fp64_dtw_r1.0", "dtw", 1.0, 0.05, "fp64")]
    res = {"workload": f"pairwise_distance of {n}x{t} against {n}x{t} (seeds 10 / 11)"}
    kinds = {1: "euclid", 2: "lane / strip", 3: "plain block"}
    for name, metric, r, g, prec in legs:
        Xa, Ya = (X32, Y32) if prec == "fp32" else (Xd, Yd)

        def call():
            device.pairwise(Xa, Ya, metric=metric, r=r, g=g, precision=prec, out=out)

        call()
        torch.cuda.synchronize()
        _lib.prof_enable(True)
        _lib.prof_take()
        k = 1 if (r == 1.0 and metric != "euclidean") else reps
        dt = timed(torch, call, k)
        kern_ms, launches, kind = _lib.prof_take()
        _lib.prof_enable(False)
        if metric == "euclidean":
            work, unit = float(n) * n * t, "MAC/s"
        else:
            work, unit = float(n) * n * band_cells(t, t, r), "cells/s"
        rows = 2 if r == 1.0 else 4
        got = out[:rows].cpu().numpy()
        ox, oy = (Xr, Yr) if prec == "fp32" else (X64, Y64)
        want = oracle.pairwise(ox[:rows], oy, metric=metric, r=r, g=g)
        leg = {"ms": dt * 1e3, "kernel_ms": kern_ms / max(launches, 1), "work": work,
               "value": work / dt, "unit": unit, "kernel": kinds.get(kind, kind)}
        # the leg against the VALU peak of its own cell arithmetic
        peak, basis = LEG_PEAKS[(metric, prec)]
        leg["frac_of_peak"] = work / dt / peak
        leg["peak"] = peak
        leg["peak_basis"] = basis
        if prec == "fp32":
            leg["max_rel_err_rows"] = float(np.max(np.abs(got - want) / np.abs(want)))
        else:
            leg["mismatch_rows"] = int((got != want).sum())
        leg["rows_checked"] = rows
        res[name] = leg
        log(f"pairwise {name}: {dt * 1e3:.2f} ms, {work / dt / 1e12:.3f} T{unit} "
            f"({leg['kernel']}); checked {rows} rows")
    return res


# VALU peak of each pairwise leg's cell arithmetic: 78.6e12 FP32 and 39.3e12
# FP64 lane-instructions per second (the published 157.3 / 78.6 TFLOP/s
# vector peaks, an FMA counted as 2 flops), divided by the instructions one
# algorithmic cell needs.
LEG_PEAKS = {
    ("dtw", "fp32"): (78.6e12 / 3, "3 FP32 VALU per cell (sub, min3, fma)"),
    ("wdtw", "fp32"): (78.6e12 / 4, "4 FP32 VALU per cell (sub, mul, min3, fma)"),
    ("euclidean", "fp32"): (78.6e12 / 2, "2 FP32 VALU per multiply-add (sub, fma)"),
    ("dtw", "fp64"): (39.3e12 / 5, "5 FP64 VALU per cell (sub, mul, min, min, add)"),
}


# ----------------------------------------------------------------------- train
def run_train(torch, dev, reps, dba_steps):
    from wildboar_amd import device
    from wildboar_amd.dtw import dtw, dtw_barycenter, dtw_state_bytes

    X = torch.from_numpy(random_walks(40, 4096, 256).astype(np.float32)).to(dev)
    P = torch.from_numpy(random_walks(41, 64, 256).astype(np.float32)).to(dev)
    cells = 4096 * 64 * band_cells(256, 256, 0.1)
    out = torch.empty((4096, 64), dtype=torch.float64, device=dev)
    Xg = X.clone().requires_grad_(True)
    Pg = P.clone().requires_grad_(True)

    def inf():
        device.pairwise(X, P, metric="dtw", r=0.1, precision="fp32", out=out)

    def fwd():
        return dtw(Xg, Pg, 0.1)

    def fwd_bwd():
        dtw(Xg, Pg, 0.1).sum().backward()

    def fwd_ck():
        return dtw(Xg, Pg, 0.1, checkpoint=True)

    def fwd_bwd_ck():
        dtw(Xg, Pg, 0.1, checkpoint=True).sum().backward()

    for _ in range(3):
        inf()
        fwd_bwd()
        fwd_bwd_ck()
    t_inf = timed(torch, inf, reps)
    t_fwd = timed(torch, fwd, reps)
    t_fb = timed(torch, fwd_bwd, reps)
    t_fwd_ck = timed(torch, fwd_ck, reps)
    t_fb_ck = timed(torch, fwd_bwd_ck, reps)
    state_planes, _ = dtw_state_bytes(4096, 64, 256, 256, 0.1, checkpoint=False)
    state_ck, _ = dtw_state_bytes(4096, 64, 256, 256, 0.1, checkpoint=True)
    D = torch.from_numpy(random_walks(42, 512, 256).astype(np.float32)).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = dtw_barycenter(D, 0.1, steps=dba_steps, lr=1e-2, return_history=True)
    torch.cuda.synchronize()
    t_dba = time.perf_counter() - t0
    res = {
        "workload": "dtw(X, P, r=0.1) autograd, X 4096x256, P 64x256 float32, grads w.r.t. both",
        "cells": int(cells),
        "inference_fwd_ms": t_inf * 1e3,
        "fwd_ms": t_fwd * 1e3,
        "fwd_bwd_ms": t_fb * 1e3,
        "fwd_bwd_over_fwd": t_fb / t_fwd,
        "fwd_bwd_over_inference_fwd": t_fb / t_inf,
        "state_bytes": state_planes,
        "checkpointed": {
            "note": "dtw(..., checkpoint=True): cost rows kept every 32 anti-diagonal steps, "
                    "each band recomputed with its decisions in the backward",
            "fwd_ms": t_fwd_ck * 1e3,
            "fwd_bwd_ms": t_fb_ck * 1e3,
            "fwd_bwd_over_fwd": t_fb_ck / t_fwd_ck,
            "fwd_bwd_over_inference_fwd": t_fb_ck / t_inf,
            "state_bytes": state_ck,
        },
        "dba": {"series": "512x256 (seed 42)", "steps": dba_steps, "seconds": t_dba,
                "loss_first": float(hist[0]), "loss_last": float(hist[-1])},
    }
    log(f"train: inference {t_inf * 1e3:.3f} ms, fwd {t_fwd * 1e3:.3f} ms, fwd+bwd "
        f"{t_fb * 1e3:.3f} ms (x{t_fb / t_fwd:.2f}); checkpointed fwd {t_fwd_ck * 1e3:.3f} ms, "
        f"fwd+bwd {t_fb_ck * 1e3:.3f} ms; DBA {dba_steps} steps {t_dba:.2f} s")
    return res


# ----------------------------------------------------------------- subsequence
def run_subsequence(torch, dev, n_series):
    import oracle
    from wildboar_amd import device

    Xall = random_walks(20, 10000, 1024)
    rng = np.random.default_rng(21)
    L = rng.integers(16, 129, 4096)
    src = rng.integers(0, 10000, 4096)
    start = np.array([rng.integers(0, 1024 - l + 1) for l in L])
    shp64 = [Xall[s, a:a + l].copy() for s, a, l in zip(src, start, L)]
    X = torch.from_numpy(Xall[:n_series].astype(np.float32)).to(dev)
    # shapelets resident on the device, packed once (as X is)
    shp = device.pack_shapelets([torch.from_numpy(s.astype(np.float32)) for s in shp64], dev)
    res = {"workload": f"4096 shapelets, L U[16,128], from 10000 random walks of length 1024; "
                       f"scan over the first {n_series} series, float32"}
    for metric in ("dtw", "euclidean"):
        device.subsequence(shp, X[:8], metric=metric, r=0.1, precision="fp32")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dist, idx = device.subsequence(shp, X, metric=metric, r=0.1, precision="fp32")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if metric == "dtw":
            per = np.array([(1025 - l) * band_cells(int(l), int(l), 0.1) for l in L])
            unit = "cells/s"
        else:
            per = np.array([(1025 - l) * int(l) for l in L])
            unit = "MAC/s"
        work = float(per.sum()) * n_series
        # indices and distances of a block of tasks against the float64 oracle
        xs = Xall[:8].astype(np.float32).astype(np.float64)
        od, oi = oracle.subsequence([s.astype(np.float32).astype(np.float64) for s in shp64[:64]],
                                    xs, metric=metric, r=0.1)
        gi = idx[:8, :64].cpu().numpy()
        gd = dist[:8, :64].cpu().numpy()
        res[metric] = {"seconds": dt, "work": work, "value": work / dt, "unit": unit,
                       "index_mismatch_8x64": int((gi != oi).sum()),
                       "dist_mismatch_8x64": int((gd != od).sum())}
        log(f"subsequence {metric}: {dt:.3f} s, {work / dt / 1e12:.2f} T{unit}; "
            f"index mismatches {res[metric]['index_mismatch_8x64']}")
    return res


# -------------------------------------------------------------- float64 paths
def run_fp64(torch, dev):
    """The exact float64 paths (the drop-in default precision): window
    profiles, k best windows and pairwise DTW, each by the register-band
    float64 kernels and, on the same inputs, by the block kernels (plain
    policy); a slice of every result against the oracle, bit for bit."""
    import oracle
    from wildboar_amd import _lib, device

    Xh = random_walks(40, 2000, 1024)
    X = torch.from_numpy(Xh).to(dev)
    rng = np.random.default_rng(41)
    res = {"workload": "float64 random walks, 2000 series x 1024; dtw r = 0.1 unless noted"}

    def both(name, fn, cells, unit="cells/s"):
        fn()
        t_lane = timed(torch, fn, 3)
        _lib.set_policy(1)
        try:
            fn()
            t_plain = timed(torch, fn, 1)
        finally:
            _lib.set_policy(0)
        res[name] = {"seconds": t_lane, "work": cells, "value": cells / t_lane, "unit": unit,
                     "plain_seconds": t_plain, "speedup_vs_block": t_plain / t_lane}
        log(f"fp64 {name}: {t_lane * 1e3:.2f} ms, {cells / t_lane / 1e12:.3f} T{unit}; "
            f"block kernels {t_plain * 1e3:.1f} ms ({t_plain / t_lane:.1f}x)")

    # 1. profile of one pattern (L = 128) over every series: subsequence_match
    pat = Xh[7, 300:428].copy()
    pd = [torch.from_numpy(pat)]
    nwin = 1024 - 128 + 1
    both("profile_dtw", lambda: device.subsequence_profile(pd, X, metric="dtw", r=0.1),
         2000.0 * nwin * band_cells(128, 128, 0.1))
    prof = device.subsequence_profile(pd, X[:4], metric="dtw", r=0.1).cpu().numpy()
    band = oracle.compute_r(128, 0.1)
    want = np.array([[oracle.dtw_distance(pat, Xh[q, s:s + 128], band) for s in range(nwin)]
                     for q in range(4)])
    res["profile_dtw"]["mismatch_4_rows"] = int((prof != want).sum())
    both("profile_euclidean", lambda: device.subsequence_profile(pd, X, metric="euclidean"),
         2000.0 * nwin * 128, unit="MAC/s")
    res["profile_euclidean"]["note"] = (
        "one kernel under both policies (the block policy switches only the elastic "
        "kernels): speedup_vs_block compares two timings of the same launch")
    # 2. k = 5 best windows, pattern i (L U[16, 128]) against series i
    L = rng.integers(16, 129, 2000)
    pats = [Xh[int(rng.integers(2000)), a:a + l].copy()
            for a, l in zip(rng.integers(0, 1024 - 128, 2000), L)]
    pds = device.pack_shapelets([torch.from_numpy(p) for p in pats], dev)
    cells = float(sum((1025 - l) * band_cells(int(l), int(l), 0.1) for l in L))
    both("argmin_subsequence_k5",
         lambda: device.argmin_subsequence(pds, X, k=5, metric="dtw", r=0.1), cells)
    gi, gd = device.argmin_subsequence(pds, X, k=5, metric="dtw", r=0.1)
    oi, od = oracle.argmin_subsequence(pats[:8], Xh[:8], k=5, metric="dtw", r=0.1)
    res["argmin_subsequence_k5"]["mismatch_8_rows"] = int(
        (gi[:8].cpu().numpy() != oi).sum() + (gd[:8].cpu().numpy() != od).sum())
    # 3. pairwise DTW 1000 x 1000, t = 256, r = 0.05 (B = 23)
    A, Bm = X[:1000, :256].contiguous(), X[1000:, :256].contiguous()
    both("pairwise_dtw_t256_r0.05",
         lambda: device.pairwise(A, Bm, metric="dtw", r=0.05, precision="fp64"),
         1e6 * band_cells(256, 256, 0.05))
    got = device.pairwise(A[:4], Bm, metric="dtw", r=0.05, precision="fp64").cpu().numpy()
    want = oracle.pairwise(Xh[:4, :256], Xh[1000:, :256], metric="dtw", r=0.05)
    res["pairwise_dtw_t256_r0.05"]["mismatch_4_rows"] = int((got != want).sum())
    # 4. the same at r = 0.1 (B = 49, the widest register class)
    both("pairwise_dtw_t256_r0.1",
         lambda: device.pairwise(A, Bm, metric="dtw", r=0.1, precision="fp64"),
         1e6 * band_cells(256, 256, 0.1))
    got = device.pairwise(A[:4], Bm, metric="dtw", r=0.1, precision="fp64").cpu().numpy()
    want = oracle.pairwise(Xh[:4, :256], Xh[1000:, :256], metric="dtw", r=0.1)
    res["pairwise_dtw_t256_r0.1"]["mismatch_4_rows"] = int((got != want).sum())
    # the default float64 precision on a 512-long band of 101 slots (the widest
    # register class, 104), 500 x 500
    A5, B5 = X[:500, :512].contiguous(), X[1000:1500, :512].contiguous()
    both("pairwise_dtw_t512_r0.1",
         lambda: device.pairwise(A5, B5, metric="dtw", r=0.1, precision="fp64"),
         2.5e5 * band_cells(512, 512, 0.1))
    got = device.pairwise(A5[:2], B5, metric="dtw", r=0.1, precision="fp64").cpu().numpy()
    want = oracle.pairwise(Xh[:2, :512], Xh[1000:1500, :512], metric="dtw", r=0.1)
    res["pairwise_dtw_t512_r0.1"]["mismatch_2_rows"] = int((got != want).sum())
    # 5. derivative and amercing DTW, pairwise 1000 x 1000, t = 256, r = 0.1
    both("pairwise_ddtw_t256_r0.1",
         lambda: device.pairwise(A, Bm, metric="ddtw", r=0.1, precision="fp64"),
         1e6 * square_band_w_cells(254, band_w_of(256, 0.1)))
    got = device.pairwise(A[:4], Bm, metric="ddtw", r=0.1, precision="fp64").cpu().numpy()
    want = oracle.pairwise(Xh[:4, :256], Xh[1000:, :256], metric="ddtw", r=0.1)
    res["pairwise_ddtw_t256_r0.1"]["mismatch_4_rows"] = int((got != want).sum())
    both("pairwise_wdtw_t256_r0.1_g0.05",
         lambda: device.pairwise(A, Bm, metric="wdtw", r=0.1, g=0.05, precision="fp64"),
         1e6 * band_cells(256, 256, 0.1))
    got = device.pairwise(A[:4], Bm, metric="wdtw", r=0.1, g=0.05, precision="fp64").cpu().numpy()
    want = oracle.pairwise(Xh[:4, :256], Xh[1000:, :256], metric="wdtw", r=0.1, g=0.05)
    res["pairwise_wdtw_t256_r0.1_g0.05"]["mismatch_4_rows"] = int((got != want).sum())
    both("pairwise_adtw_t256_r0.1_p0.5",
         lambda: device.pairwise(A, Bm, metric="adtw", r=0.1, p=0.5, precision="fp64"),
         1e6 * band_cells(256, 256, 0.1))
    got = device.pairwise(A[:4], Bm, metric="adtw", r=0.1, p=0.5, precision="fp64").cpu().numpy()
    want = oracle.pairwise(Xh[:4, :256], Xh[1000:, :256], metric="adtw", r=0.1, p=0.5)
    res["pairwise_adtw_t256_r0.1_p0.5"]["mismatch_4_rows"] = int((got != want).sum())
    # 6. z-normalised scan (scale=True, the UCR-suite cascade reproduced):
    # 64 patterns, L U[16, 128], against 1000 series; the block baseline is
    # the scratch-row kernel
    Ls = rng.integers(16, 129, 64)
    spats = [Xh[int(rng.integers(2000)), a:a + l] * 1.7 + 0.3
             for a, l in zip(rng.integers(0, 1024 - 128, 64), Ls)]
    sp = [torch.from_numpy(p.copy()) for p in spats]
    Xs = X[:1000]
    cells = 1000.0 * sum((1025 - int(l)) * square_band_cells(int(l), 0.1) for l in Ls)
    both("scaled_dtw_subsequence_64x1000",
         lambda: device.scaled_subsequence(sp, Xs, metric="dtw", r=0.1), cells)
    gd, gi = device.scaled_subsequence(sp, Xs[:3], metric="dtw", r=0.1)
    od, oi = oracle.scaled_subsequence(spats, Xh[:3], "dtw", r=0.1)
    res["scaled_dtw_subsequence_64x1000"]["mismatch_3_rows"] = int(
        (gd.cpu().numpy() != od).sum() + (gi.cpu().numpy() != oi).sum())
    return res


def band_w_of(n, r):
    """_compute_r (distance/_elastic.pyx:1924-1925): max(floor(n * r), 1)."""
    return max(int(np.floor(n * r)), 1)


def square_band_w_cells(n, w):
    """Cells of dtw_distance's band on n x n with width w: |i - j| < w."""
    return sum(min(n, i + w) - max(0, i - w + 1) for i in range(n))


def square_band_cells(L, r):
    """Band cells of the scaled scan's DTW: |i - j| <= w on an L x L matrix,
    w = _compute_warp_width(L, r) (distance/_elastic.pyx:1917-1921)."""
    w = L - 1 if r == 1 else int(np.floor(L * r))
    return sum(min(L - 1, i + w) - max(0, i - w) + 1 for i in range(L))


# ------------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--t", type=int, default=512)
    ap.add_argument("--r", type=float, default=0.1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the other pairwise metric legs")
    ap.add_argument("--no-train", action="store_true")
    ap.add_argument("--no-subsequence", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--batched-pairs", type=int, default=50000)
    ap.add_argument("--subsequence-series", type=int, default=10000)
    ap.add_argument("--dba-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB three times (scales the PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(every box has one GPU; the multi-GPU design is described in DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        src = torch.ones(64 * 1024 * 1024, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied 3 x 268435456 bytes")
        return

    from wildboar_amd import _lib, device

    lib = _lib.load()
    _lib.check(lib.wbx_init(0))
    dev = torch.device("cuda:0")
    n, t, r = args.n, args.t, args.r
    X64 = random_walks(10, n, t)
    Y64 = random_walks(11, n, t)
    X = torch.from_numpy(X64.astype(np.float32)).to(dev)
    Y = torch.from_numpy(Y64.astype(np.float32)).to(dev)
    out = torch.empty((n, n), dtype=torch.float64, device=dev)
    cells_pair = band_cells(t, t, r)
    cells = n * n * cells_pair

    def step():
        device.pairwise(X, Y, metric="dtw", r=r, precision="fp32", out=out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.prof_enable(True)
    _lib.prof_take()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kern_ms, launches, kind = _lib.prof_take()
    _lib.prof_enable(False)
    ms_step = elapsed / args.steps * 1e3
    value = cells / (elapsed / args.steps)
    kern_avg_s = kern_ms / max(launches, 1) / 1e3
    achieved = cells / kern_avg_s
    log(f"step {ms_step:.3f} ms, kernel {kern_avg_s * 1e3:.3f} ms (kind {kind}), "
        f"{value / 1e12:.3f} Tcells/s whole step, {achieved / 1e12:.3f} Tcells/s kernel")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_lane_dtw_pairwise_2000x512_r0.1.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    probe = _lib.valu_peak()
    log(f"VALU probe: {probe / 1e12:.2f} T lane-instr/s = {probe / 3 / 1e12:.2f} Tcells/s at 3/cell")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: z-normalised Gaussian random walks, numpy default_rng seeds 10 and 11",
        "config": {
            "workload": f"pairwise_distance(X, Y, metric='dtw', metric_params={{'r': {r}}}), "
                        f"X and Y {n}x{t} float32 resident in HBM, float64 {n}x{n} output",
            "cells_per_pair": cells_pair,
            "cells_per_step": cells,
            "kernel": {1: "euclid", 2: "lane-per-pair register band", 3: "plain"}.get(kind, kind),
        },
        "roofline": {
            # DTW cells are VALU min-plus work (no MFMA shape, < 1% of HBM): the
            # bound is the FP32 vector issue rate.  The published 157.3 TFLOP/s
            # counts one FMA lane-instruction as 2 flops; a cell is 3 lane-
            # instructions (v_sub, v_min3, v_fma), so it is counted as 6.
            "bound": "valu",
            "achieved": achieved * 6 / 1e12,
            "peak": 157.3,
            "unit": "TFLOP/s",
            "frac": achieved * 6 / 157.3e12,
            "traffic": traffic,
            "basis": "FMA-equivalent FP32 VALU work: 3 lane-instructions per algorithmic band "
                     "cell, 2 flops each, against the published FP32 vector peak",
            "achieved_cells_per_s": achieved,
            "peak_cells_per_s": PEAK_CELLS,
            "measured_valu_peak_cells_per_s": probe / 3,
            "frac_of_measured": achieved / (probe / 3),
            "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE "
                            "x 1.0 (copy-kernel calibration); algorithmic bytes 24.6 MB",
        },
    }

    if not args.no_extra:
        import oracle

        # FP32 accuracy and 1-NN labels against the float64 oracle on a slice.
        rows = 20
        Dg = out[:rows].cpu().numpy()
        Xs = X.cpu().numpy().astype(np.float64)
        Ys = Y.cpu().numpy().astype(np.float64)
        Do = oracle.pairwise(Xs[:rows], Ys, metric="dtw", r=r)
        acc = {"max_rel_err": float((np.abs(Dg - Do) / np.abs(Do)).max())}
        labels = np.random.default_rng(12).integers(0, 10, n)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        idx, _ = device.argmin(X, Y, 1, metric="dtw", r=r, precision="fp32")
        torch.cuda.synchronize()
        acc["argmin_1nn_ms"] = (time.perf_counter() - ta) * 1e3
        # 1-NN labels of every row against the oracle's early-abandoning
        # argmin (the reference's heap loop) on all host cores
        tl = time.perf_counter()
        oi, _ = oracle.argmin(Xs, Ys, k=1, metric="dtw", r=r, n_threads=os.cpu_count() or 1)
        tl = time.perf_counter() - tl
        gi = idx[:, 0].cpu().numpy()
        acc["label_mismatch"] = int((labels[gi] != labels[oi[:, 0]]).sum())
        acc["index_mismatch"] = int((gi != oi[:, 0]).sum())
        acc["label_rows_checked"] = n
        acc["oracle_argmin_seconds"] = tl

        # argmin with the library's LB_Keogh: only the pairs the reference's
        # loop evaluates (plus seeds) are computed, float64, exact
        from wildboar_amd.lb import DtwKeoghLowerBound

        Xe = torch.from_numpy(Xs).to(dev)
        Ye = torch.from_numpy(Ys).to(dev)
        lbk = torch.from_numpy(DtwKeoghLowerBound(r).fit(Ys).transform(Xs)).to(dev)
        device.argmin(Xe, Ye, 1, metric="dtw", r=r, precision="fp64", lower_bound=lbk)
        _lib.prof_pruned()
        t_lb = timed(torch, lambda: device.argmin(Xe, Ye, 1, metric="dtw", r=r, precision="fp64",
                                                  lower_bound=lbk), 3)
        evaluated = _lib.prof_pruned() / 3
        idx_lb, _ = device.argmin(Xe, Ye, 1, metric="dtw", r=r, precision="fp64", lower_bound=lbk)
        t_full = timed(torch, lambda: device.argmin(Xe, Ye, 1, metric="dtw", r=r, precision="fp64"),
                       3)
        acc["argmin_lb_keogh_fp64"] = {
            "ms": t_lb * 1e3, "no_bound_ms": t_full * 1e3, "speedup": t_full / t_lb,
            "pairs_evaluated_frac": evaluated / (n * n),
            "index_mismatch": int((idx_lb[:, 0].cpu().numpy() != oi[:, 0]).sum()),
            "note": "argmin_distance(k=1, lower_bound=DtwKeoghLowerBound(r).fit(Y).transform(X)) "
                    "in float64 against the same call without a bound"}
        log(f"argmin with LB_Keogh: {t_lb * 1e3:.2f} ms ({evaluated / (n * n):.3f} of the pairs) "
            f"vs {t_full * 1e3:.2f} ms without a bound")

        # Plain one-pair-per-block kernel on a slice of the same workload.
        prow = 64
        _lib.set_policy(1)
        try:
            outp = torch.empty((prow, n), dtype=torch.float64, device=dev)
            device.pairwise(X[:prow], Y, metric="dtw", r=r, precision="fp32", out=outp)
            tp = timed(torch, lambda: device.pairwise(X[:prow], Y, metric="dtw", r=r,
                                                      precision="fp32", out=outp), 1)
        finally:
            _lib.set_policy(0)
        acc["plain_cells_per_s"] = prow * n * cells_pair / tp
        acc["speedup_vs_plain"] = value / acc["plain_cells_per_s"]
        result["accuracy"] = acc
        log(f"plain {acc['plain_cells_per_s'] / 1e12:.4f} Tcells/s; max rel err "
            f"{acc['max_rel_err']:.2e}; label mismatches {acc['label_mismatch']}")

        if not args.no_legs:
            result["pairwise_legs"] = run_pairwise_legs(torch, dev, X64, Y64, 3)
        if not args.no_batched:
            result["batched"] = run_batched(torch, dev, args.batched_pairs, 2048, 5)
        if not args.no_train:
            result["train"] = run_train(torch, dev, 10, args.dba_steps)
        if not args.no_subsequence:
            result["subsequence"] = run_subsequence(torch, dev, args.subsequence_series)
        if not args.no_fp64:
            result["fp64"] = run_fp64(torch, dev)

    if not args.no_cpu:
        import oracle

        cores = os.cpu_count() or 1
        Xs = X.cpu().numpy().astype(np.float64)
        Ys = Y.cpu().numpy().astype(np.float64)
        rows = max(1, cores)
        tc = time.perf_counter()
        oracle.pairwise(Xs[:rows], Ys, metric="dtw", r=r, n_threads=cores)
        tc = time.perf_counter() - tc
        rows2 = int(min(n, max(rows, rows * args.cpu_seconds / max(tc, 1e-3))))
        if rows2 > rows:
            tc = time.perf_counter()
            oracle.pairwise(Xs[:rows2], Ys, metric="dtw", r=r, n_threads=cores)
            tc = time.perf_counter() - tc
            rows = rows2
        result["cpu_baseline"] = {
            "value": rows * n * cells_pair / tc,
            "unit": "cells/s",
            "cores": cores,
            "kind": "port",
            "sample": f"float64 oracle (C restatement of _elastic.pyx dtw_distance, -O2, "
                      f"{cores} threads over row blocks) on X[:{rows}] x Y ({rows}x{n} pairs, "
                      f"{tc:.1f} s)",
        }
        log(f"cpu {result['cpu_baseline']['value'] / 1e9:.3f} Gcells/s on {cores} threads")

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
