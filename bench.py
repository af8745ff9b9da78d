"""Benchmark of the embedding path: pca(50) + neighbors(15) on one MI355X.

A step is one pass of the path over the resident layer: the streamed block-Krylov
PCA (q product passes, FP64 dense steps on the host, embedding left on the device),
then the kNN graph of the embedding, downloaded to the host. The layer is the
synthetic 1,000,000 cells x 2,000 genes log1p CSR (20 cell types, ~8% density),
uploaded before the timed region.

    python bench.py [--steps K] [--warmup W] [--workload 1m] [--search exact|learned]
                    [--precision bf16|fp32] [--q 4] [--no-cpu-baseline]

Prints ONE JSON line (the driver's contract). --gpus N with N > 1 exits: this path
runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
PEAKS_TFLOPS = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(layer, embedding, n_genes, q, threads):
    """The oracle (kyto's CPU path restated) on a bounded sample of the workload:
    PCA on the first `pca_cells` cells (its cost is linear in cells), and the filtered
    exact kNN for `knn_queries` query rows against all rows (linear in queries)."""
    import oracle

    n = len(layer.indptr) - 1
    pca_cells = min(n, 50_000)
    ip = layer.indptr[: pca_cells + 1]
    nnz = int(ip[-1])
    t0 = time.perf_counter()
    oracle.pca(ip, layer.indices[:nnz], layer.values[:nnz], n_genes, 50, seed=0, q=q)
    pca_s = time.perf_counter() - t0
    knn_queries = min(n, 4096)
    t0 = time.perf_counter()
    oracle.knn_exact_range(embedding, 14, 0, knn_queries, threads=threads)
    knn_s = time.perf_counter() - t0
    pca_rate = pca_cells / pca_s
    knn_rate = knn_queries / knn_s
    value = 1.0 / (1.0 / pca_rate + 1.0 / knn_rate)
    return {
        "value": value,
        "unit": "cells/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle pca(50, q={q}) fold on 1 thread over the first {pca_cells} cells "
                   f"({pca_s:.1f} s, {pca_rate:.0f} cells/s) + oracle filtered exact kNN(15) of "
                   f"{knn_queries} queries against all {n} rows on {threads} threads "
                   f"({knn_s:.1f} s, {knn_rate:.0f} queries/s); value = 1/(1/pca + 1/knn)"),
    }


def copy_calibration(nbytes=256 << 20):
    """One plain device-to-device copy of `nbytes` through the HIP runtime (its copy
    kernel reads and writes exactly nbytes): profiles divide the HBM counters of the
    path's kernels by what the counters read for this copy."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: {nbytes} bytes device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="1m")
    ap.add_argument("--search", default="learned", choices=["exact", "learned"])
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16"])
    ap.add_argument("--n-lists", type=int, default=4096)
    ap.add_argument("--kmeans-iters", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--train-queries", type=int, default=200_000)
    ap.add_argument("--confidence", type=float, default=0.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one device-to-device copy of 256 MiB (the HBM counters' reference kernel)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the recall check and the plain FP32 kNN stage timed beside the learned path")
    ap.add_argument("--q", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this path runs on one GPU: every run here uses one device; "
                                   "see DESIGN.md for how it would shard"}))
        sys.exit(2)

    import kyto_amd
    from kyto_amd import synthetic as S

    w = S.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    layer = S.workload(args.workload)
    log(f"workload {w.name}: {w.n_cells} x {w.n_genes}, nnz {layer.nnz} (density {layer.density:.4f}), "
        f"generated in {time.perf_counter() - t0:.1f} s")

    dev = kyto_amd.Device(0)
    csr = dev.upload_csr(layer.indptr, layer.indices, layer.values, w.n_genes)
    log("layer resident on", dev.info())
    n = w.n_cells
    k = 15

    def step():
        res = dev.pca(csr, 50, seed=0, power_iterations=args.q, download_embedding=False)
        emb = res.device_embedding
        if args.search == "exact":
            g = dev.neighbors(emb, k, precision=args.precision)
        else:
            t0 = time.perf_counter()
            index = dev.ivf_build(emb, n_lists=args.n_lists, kmeans_iters=args.kmeans_iters, seed=0, hidden=256,
                                  train_queries=args.train_queries, batch=4096, epochs=args.epochs, lr=1e-3,
                                  label_k=k - 1)
            build_ms = (time.perf_counter() - t0) * 1e3
            g = index.search(emb, k, n_probe=8, use_router=True, confidence=args.confidence)
            g.info["build_ms"] = build_ms
            g.info.update({f"build_{key}": v for key, v in index.info.items()})
            g.info["search_ms"] = g.info["route_ms"] + g.info["search_ms"]
            g.info["knn_total_ms"] = build_ms + g.info["total_ms"]
            g.info["total_ms"] = g.info["knn_total_ms"]
            g.index = index
        return res, g, emb

    def release(g, emb):
        index = getattr(g, "index", None)
        if index is not None:
            index.free()
        emb.free()

    for i in range(args.warmup):
        res, g, emb = step()
        log(f"warmup {i}: pca {res.timing['total_ms']:.1f} ms, knn {g.info['total_ms']:.1f} ms")
        release(g, emb)
    times, pcas, knns = [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        res, g, emb = step()
        dt = time.perf_counter() - t0
        times.append(dt)
        pcas.append(res.timing)
        knns.append(g.info)
        log(f"step {i}: {dt * 1e3:.1f} ms (pca {res.timing['total_ms']:.1f} ms, knn {g.info['total_ms']:.1f} ms, "
            f"search {g.info['search_ms']:.1f}, fallback rows {g.info['fallback']}) {g.info}")
        if i < args.steps - 1:
            release(g, emb)
    ms = 1e3 * float(np.mean(times))
    value = n / (ms / 1e3)
    d = 50
    peak = PEAKS_TFLOPS[args.precision]
    if args.search == "exact":
        # Dominant kernel: the full candidate scan, 2 n^2 d flops a launch (unpadded d).
        scan = [x.get("scan_kernel_ms", 0.0) or x["search_ms"] for x in knns]
        scan_ms = float(np.mean(scan))
        achieved = 2.0 * n * n * d / (scan_ms / 1e3) / 1e12
        roofline = {"bound": "mfma" if args.precision == "bf16" else "valu",
                    "kernel": "search_bf16<false>" if args.precision == "bf16" else "scan_points (FP32)",
                    "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": None, "launch_ms": scan_ms, "algorithmic_per_launch": f"2 n^2 d = {2.0 * n * n * d:.3e} flop"}
    else:
        roofline = None

    def spmm_roofline(kind):
        ms = float(np.mean([p[f"{kind}_ms"] for p in pcas]))
        launches = pcas[-1][f"{kind}_launches"]
        nbytes = pcas[-1][f"{kind}_bytes"]
        gbs = nbytes / (ms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "launches": launches, "launch_ms": ms / max(launches, 1), "bytes_per_launch": nbytes / max(launches, 1)}

    spmm = {"column_product": spmm_roofline("column_product"), "row_project": spmm_roofline("row_product")}
    embedding = emb.download()

    def mean_fields(rows):
        keys = [key for key, v in rows[-1].items() if isinstance(v, (int, float))]
        return {key: float(np.mean([r[key] for r in rows])) for key in keys}

    out = {
        "metric": "cells/s pca(50)+neighbors(15) @1M×2k",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": f"{w.n_cells} cells x {w.n_genes} genes log1p CSR, {w.n_types} cell types, "
                               f"density {layer.density:.4f}, nnz {layer.nnz}",
                   "pca": f"n_components=50, block 60, q={args.q}, seed 0",
                   "neighbors": f"n_neighbors=15, euclidean, search={args.search}, candidates {args.precision}"},
        "roofline": roofline,
        "roofline_spmm": spmm,
        "breakdown": {"pca": mean_fields(pcas), "pca_pass_ms": pcas[-1]["pass_ms"], "knn": mean_fields(knns)},
    }
    if args.search == "learned" and not args.no_compare:
        # Outside the timed region: the exact graph (kyto's) for recall, and the plain
        # FP32 exact kNN stage on the same embedding for the speed-up of the kNN stage.
        exact = dev.neighbors(emb, k, precision="bf16")
        got = g.rows()
        ref = exact.rows()
        def recall_of(rows):
            return sum(np.intersect1d(a, b, assume_unique=True).size for a, b in zip(rows, ref)) / ref.size

        recall = recall_of(got)
        centroid = g.index.search(emb, k, n_probe=8, use_router=False)
        recall_centroid = recall_of(centroid.rows())
        plain = dev.neighbors(emb, k, precision="fp32")
        learned_knn_ms = float(np.mean([x["knn_total_ms"] for x in knns]))
        out["learned"] = {"recall_at_15": recall, "recall_at_15_nearest_centroid_probing": recall_centroid,
                          "recall_definition": "share of the exact graph's 14 stored neighbours (self excluded) found",
                          "exact_graph_equal_rows": int(sum(np.array_equal(a, b) for a, b in zip(got, ref))),
                          "plain_fp32_knn_ms": plain.info["total_ms"], "learned_knn_ms": learned_knn_ms,
                          "knn_speedup_vs_plain_fp32": plain.info["total_ms"] / learned_knn_ms}
        log(f"recall@15 {recall:.5f}; plain FP32 kNN {plain.info['total_ms']:.1f} ms vs learned "
            f"{learned_knn_ms:.1f} ms ({plain.info['total_ms'] / learned_knn_ms:.1f}x)")
    release(g, emb)
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(layer, embedding, w.n_genes, args.q, threads)
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        copy_calibration()


if __name__ == "__main__":
    main()
