"""Benchmark: wildboar's pairwise DTW on one MI355X.

Workload (one step): pairwise_distance(X, Y, metric="dtw", r=0.1) with X and Y
2000 x 512 float32 z-normalised Gaussian random walks (seeds 10, 11), inputs
resident in HBM, FP32 cells, float64 2000 x 2000 output -- the full-matrix
configuration the metric is quoted on.  The whole step runs through the C ABI
(wbx_dev_pairwise) on torch's current stream.

Printed JSON line (stdout):
  value     algorithmic band cells per second over the timed steps
  roofline  the DTW kernel's cells/s (HIP events on its launch stream) against
            the VALU min-plus peak of 3 instructions per cell
  cpu_baseline  the float64 C restatement of the reference, all host cores,
            on a bounded slice of the same workload
  plain_*   the plain one-pair-per-block kernel on a slice, and the speedup
  max_rel_err / label_mismatch  FP32 results against the float64 oracle
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--t", type=int, default=512)
    ap.add_argument("--r", type=float, default=0.1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
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
            "bound": "valu",
            "achieved": achieved / 1e9,
            "peak": PEAK_CELLS / 1e9,
            "unit": "Gcells/s",
            "frac": achieved / PEAK_CELLS,
            "traffic": traffic,
            "peak_basis": "published FP32 vector rate (157.3 TFLOP/s = 78.6e12 VALU lane-instr/s)"
                          " / 3 VALU instructions per cell (sub, min3, fma)",
            "measured_valu_peak": probe / 3 / 1e9,
            "frac_of_measured": achieved / (probe / 3),
            "note": "DTW is VALU min-plus bound; HBM traffic is < 1% of the 8 TB/s roofline "
                    "and MFMA does not apply",
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
        rel = np.abs(Dg - Do) / np.abs(Do)
        result["max_rel_err"] = float(rel.max())
        labels = np.random.default_rng(12).integers(0, 10, n)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        idx, _ = device.argmin(X, Y, 1, metric="dtw", r=r, precision="fp32")
        torch.cuda.synchronize()
        result["argmin_1nn_ms"] = (time.perf_counter() - ta) * 1e3
        oi, _ = oracle.argmin(Xs[:rows], Ys, k=1, metric="dtw", r=r)
        got = labels[idx[:rows, 0].cpu().numpy()]
        result["label_mismatch"] = int((got != labels[oi[:, 0]]).sum())
        result["label_rows_checked"] = rows

        # Plain one-pair-per-block kernel on a slice of the same workload.
        prow = 64
        _lib.set_policy(1)
        try:
            outp = torch.empty((prow, n), dtype=torch.float64, device=dev)
            device.pairwise(X[:prow], Y, metric="dtw", r=r, precision="fp32", out=outp)
            torch.cuda.synchronize()
            tp = time.perf_counter()
            device.pairwise(X[:prow], Y, metric="dtw", r=r, precision="fp32", out=outp)
            torch.cuda.synchronize()
            tp = time.perf_counter() - tp
        finally:
            _lib.set_policy(0)
        plain = prow * n * cells_pair / tp
        result["plain_cells_per_s"] = plain
        result["speedup_vs_plain"] = value / plain
        log(f"plain {plain / 1e12:.4f} Tcells/s; max rel err {result['max_rel_err']:.2e}; "
            f"label mismatches {result['label_mismatch']}")

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
