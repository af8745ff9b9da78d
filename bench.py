"""Benchmark of the embedding path: pca(50) + neighbors(15) on one MI355X.

A step is one pass of the path over the resident layer: the streamed block-Krylov
PCA (q product passes, FP64 dense steps on the host, embedding left on the device),
then the kNN graph of the embedding with BF16 MFMA candidates, exact re-rank and
certificate, downloaded to the host. The layer is the synthetic 1,000,000 cells x
2,000 genes log1p CSR (20 cell types, ~8% density), uploaded before the timed region:
`value` is the step with the layer resident in HBM. The "end_to_end" field times the
same step from the CSR in pageable host memory (upload, row checks and column copy
inside the clock), the PCIe-inclusive rate.

The metric compares the learned + BF16 search with the plain FP32 GPU path at
recall@15 >= 0.99. On this workload the learned router (an IVF probe of 8 of 4,096
lists) cannot reach that recall (a row's 14 neighbours sit in ~13 distinct lists), so
its confidence gate sends every row to the BF16 exact search: the timed step is that
search (recall 1, the exact graph). Outside the timed region the run also measures
the learned router's own recall and time, the nearest-centroid probe it is compared
with, and the plain FP32 kNN stage on the same embedding.

    python bench.py [--steps K] [--warmup W] [--precision bf16|fp32] [--q 4]
                    [--workload 1m|4m|100k] [--compare] [--no-compare] [--no-cpu-baseline] [--copy-calibration]
                    [--no-store] [--from-store]

Outside the timed region (1m, 100k) the resident layer is also written as a kyto store
under /tmp and read back onto the device through the store reader (the "store" field).

--workload 4m is the streamed form: 4,000,000 x 3,000 counts (40 types, ~7%) read in
32,768-cell blocks through the pinned double-buffered reader, the layer made on the
device, kNN with k = 30; its timed step is the same PCA + exact BF16 kNN.

Prints ONE JSON line. --gpus N with N > 1 exits: this path runs on one GPU.
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

METRIC = "cells/s pca(50)+neighbors(15) @1M\u00d72k: learned+BF16 vs FP32 at recall@15\u22650.99"
HBM_PEAK_GBS = 8000.0
PEAKS_TFLOPS = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "search_bf16_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(layer, embedding, n_genes, q, threads, device_pca, repeats=3):
    """The oracle (kyto's CPU path restated) on bounded samples of the workload, on
    `threads` host threads, median of `repeats` warm runs each: the PCA (kyto's parallel
    fold: leaves of 4,096 cells folded in parallel and combined in its fixed tree) over
    the first `pca_cells` cells (its cost is linear in cells), and the filtered exact
    kNN (blocked f32 products, f64 re-rank) of `knn_queries` queries against all rows
    (linear in queries). Also, as the checker: the device's PCA against the FP64 truth
    (the eigenpairs of the dense centred covariance) beside the oracle's own error at
    the same q, over the whole layer."""
    import oracle

    oracle.lib().ko_set_threads(int(threads))
    n = len(layer.indptr) - 1
    pca_cells = min(n, 250_000)
    ip = layer.indptr[: pca_cells + 1]
    nnz = int(ip[-1])
    pca_s = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        oracle.pca(ip, layer.indices[:nnz], layer.values[:nnz], n_genes, 50, seed=0, q=q)
        pca_s.append(time.perf_counter() - t0)
    knn_queries = min(n, 4096)
    knn_s = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        oracle.knn_exact_range(embedding, 14, 0, knn_queries)
        knn_s.append(time.perf_counter() - t0)
    pca_t, knn_t = float(np.median(pca_s)), float(np.median(knn_s))
    pca_rate = pca_cells / pca_t
    knn_rate = knn_queries / knn_t
    value = 1.0 / (1.0 / pca_rate + 1.0 / knn_rate)
    baseline = {
        "value": value,
        "unit": "cells/s",
        "cores": int(threads),
        "kind": "port",
        "host": {"cpu_model": _cpu_model(), "nproc": os.cpu_count()},
        "sample": (f"oracle pca(50, q={q}) over the first {pca_cells} cells, kyto's parallel fold on {threads} "
                   f"threads (median of {repeats}: {pca_t:.2f} s, {pca_rate:.0f} cells/s) + oracle filtered exact "
                   f"kNN(15) of {knn_queries} queries against all {n} rows on {threads} threads (median of "
                   f"{repeats}: {knn_t:.2f} s, {knn_rate:.0f} queries/s); value = 1/(1/pca + 1/knn)"),
    }
    t0 = time.perf_counter()
    tv, tl, _ = oracle.pca_truth_sparse(layer.indptr, layer.indices, layer.values, n_genes, 50)
    dev_var = np.abs(device_pca.variance - tv) / tv
    angles = oracle.subspace_angles(device_pca.loadings, tl)
    fp64 = {
        "variance_rel_err_max": float(dev_var.max()),
        "variance_rel_err_first": float(dev_var[0]),
        "subspace_angle_max_rad": float(angles.max()),
        "leading_component_angle_rad": float(oracle.subspace_angles(device_pca.loadings[:, :1], tl[:, :1]).max()),
        "truth": "eigh of the dense centred covariance (X^T X - n mu mu^T) / (n - 1) in f64 from the CSR",
        "note": f"q = {q} block-Krylov on a nearly flat spectrum (variances {tv[1]:.2f} .. {tv[-1]:.2f} after "
                f"{tv[0]:.2f}); the device equals the oracle's PCA to 1e-4 (tests/test_gpu_scale.py)",
        "ms": (time.perf_counter() - t0) * 1e3,
    }
    return baseline, fp64


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


def recall_of(rows, ref):
    """Share of the reference graph's stored neighbours (self excluded) found."""
    return sum(np.intersect1d(a, b, assume_unique=True).size for a, b in zip(rows, ref)) / ref.size


def learned_leg(dev, emb, k, ref_rows, args):
    """The learned router (k-means lists, MLP trained on exact-kNN list labels) probing
    8 lists, against probing the 8 nearest centroids; recall against the exact graph."""
    t0 = time.perf_counter()
    index = dev.ivf_build(emb, n_lists=args.n_lists, kmeans_iters=args.kmeans_iters, seed=0, hidden=256,
                          train_queries=args.train_queries, batch=4096, epochs=args.epochs, lr=1e-3, label_k=k - 1)
    build_ms = (time.perf_counter() - t0) * 1e3
    try:
        g = index.search(emb, k, n_probe=8, use_router=True, confidence=0.0)
        c = index.search(emb, k, n_probe=8, use_router=False)
        gated = index.search(emb, k, n_probe=8, use_router=True, confidence=args.confidence)
    finally:
        index.free()
    return {
        "n_lists": args.n_lists, "n_probe": 8, "train_queries": args.train_queries, "epochs": args.epochs,
        "build_ms": build_ms, "kmeans_ms": index.info["kmeans_ms"], "labels_ms": index.info["labels_ms"],
        "train_ms": index.info["train_ms"], "final_loss": index.info["final_loss"],
        "label_mass_in_top_8": index.info["label_recall_top"],
        "router_search_ms": g.info["total_ms"], f"router_recall_at_{k}": recall_of(g.rows(), ref_rows),
        "centroid_search_ms": c.info["total_ms"], f"centroid_recall_at_{k}": recall_of(c.rows(), ref_rows),
        "confidence_gate": args.confidence, "gated_rows_to_exact": int(gated.info["fallback"]),
        f"gated_recall_at_{k}": recall_of(gated.rows(), ref_rows), "gated_search_ms": gated.info["total_ms"],
    }


def store_leg(dev, layer, csr, n_genes):
    """Outside the timed region: the resident layer written as a kyto store (tiles of
    4,096 cells x 1,024 genes, inner chunks of 65,536 values, shards of 2^20, zstd 3)
    under /tmp, then read back tile by tile into the pinned buffers and onto the
    device (ky_csr_stream_layer), its values compared with the resident layer's. The
    files are in the page cache when read: this times decoding and transfer, not the
    disk."""
    import shutil
    import tempfile
    from kyto_amd import store as ks
    tmp = tempfile.mkdtemp(prefix="kyto-store-", dir="/tmp")
    try:
        if shutil.disk_usage(tmp).free < 4 * layer.values.nbytes:
            return {"skipped": "not enough free space under /tmp"}
        path = os.path.join(tmp, "layer.kyto")
        t0 = time.perf_counter()
        ks.create_store(path, n_cells=len(layer.indptr) - 1, n_genes=n_genes, cell_block=4096, gene_block=1024,
                        layers={"log1p": ks.LayerData(layer.indptr, layer.indices, layer.values)})
        write_ms = (time.perf_counter() - t0) * 1e3
        disk = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(os.path.join(path, "layers")) for f in fs)
        infos, equal = [], False
        for _ in range(2):  # the second read is reported (the first warms the reader's buffers)
            with ks.StoreLayer(path, "log1p") as opened:
                t0 = time.perf_counter()
                made = dev.stream_layer(opened)
                wall = (time.perf_counter() - t0) * 1e3
            if infos:
                equal = bool(np.array_equal(made.values().view(np.uint32), csr.values().view(np.uint32)))
            infos.append(dict(made.stream_info, wall_ms=wall))
            made.free()
        info = infos[-1]
        decoded = float(info["bytes"])
        return {"write_ms": write_ms, "bytes_on_disk": disk, "decoded_bytes": decoded, "blocks": int(info["blocks"]),
                "read_ms": info["read_ms"], "stream_ms": info["total_ms"], "wall_ms": info["wall_ms"],
                "decode_GBps": decoded / (info["read_ms"] / 1e3) / 1e9,
                "stream_GBps": decoded / (info["total_ms"] / 1e3) / 1e9, "equal_to_resident_layer": equal,
                "tiling": "4096 cells x 1024 genes, K 65536, S 1048576, zstd level 3"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="1m")
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16"])
    ap.add_argument("--q", type=int, default=4)
    ap.add_argument("--n-lists", type=int, default=4096)
    ap.add_argument("--kmeans-iters", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--train-queries", type=int, default=200_000)
    ap.add_argument("--confidence", type=float, default=0.99,
                    help="router confidence (softmax mass of the probed lists) below which a row is searched exactly")
    ap.add_argument("--compare", action="store_true",
                    help="run the comparison legs at 4m too (plain FP32 kNN stage, learned leg: minutes)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the learned leg, the recall check and the plain FP32 kNN stage")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--from-store", action="store_true",
                    help="4m: write the counts as a kyto store first and read the layer from it")
    ap.add_argument("--no-store", action="store_true", help="skip the store leg (write the layer as a kyto store, "
                                                            "read it back onto the device)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one device-to-device copy of 256 MiB (the HBM counters' reference kernel)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this path runs on one GPU: every run here uses one device; "
                                   "see DESIGN.md for how it would shard"}))
        sys.exit(2)

    import kyto_amd
    from kyto_amd import synthetic as S

    w = S.WORKLOADS[args.workload]
    dev = kyto_amd.Device(0)
    n = w.n_cells
    ingest = None
    if args.workload == "4m" and args.from_store:
        # The counts written as a kyto store under /tmp first (outside any timing but
        # its own), then read tile by tile into the pinned buffers and made into the
        # log1p layer on the device.
        import shutil
        import tempfile
        from kyto_amd import store as ks
        c = S.counts(w.n_cells, w.n_genes, w.n_types, w.density, w.seed)
        tmp = tempfile.mkdtemp(prefix="kyto-store-", dir="/tmp")
        try:
            path = os.path.join(tmp, "counts.kyto")
            t0 = time.perf_counter()
            ks.create_store(path, n_cells=w.n_cells, n_genes=w.n_genes, cell_block=4096, gene_block=1024,
                            layers={"counts": ks.LayerData(c.indptr, c.indices, c.values.astype(np.uint32))})
            write_ms = (time.perf_counter() - t0) * 1e3
            del c
            disk = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(os.path.join(path, "layers"))
                       for f in fs)
            with ks.StoreLayer(path, "counts") as opened:
                t0 = time.perf_counter()
                csr = dev.stream_layer(opened, block_rows=32_768, target_sum=1e4, log1p=True)
                wall = (time.perf_counter() - t0) * 1e3
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        ingest = dict(csr.stream_info, wall_ms=wall, store_write_ms=write_ms, store_bytes_on_disk=disk,
                      source="kyto store (tiles 4,096 x 1,024, zstd 3)")
        layer = None
        nnz = csr.nnz
        log(f"workload {w.name}: {w.n_cells} x {w.n_genes}, nnz {nnz}, store written in {write_ms:.0f} ms "
            f"({disk / 1e9:.2f} GB on disk), read onto the device in {ingest['blocks']} blocks, "
            f"{ingest['total_ms']:.0f} ms (reader {ingest['read_ms']:.0f} ms)")
    elif args.workload == "4m":
        # Read in 32,768-cell row blocks through the pinned double-buffered reader; the
        # counts become the log1p layer on the device; the host holds two blocks.
        reader = S.block_reader(w.n_cells, w.n_genes, w.n_types, w.density, w.seed)
        t0 = time.perf_counter()
        csr = dev.stream_counts(w.n_cells, w.n_genes, reader, block_rows=32_768, target_sum=1e4, log1p=True)
        ingest = dict(csr.stream_info)
        ingest["wall_ms"] = (time.perf_counter() - t0) * 1e3
        layer = None
        nnz = csr.nnz
        log(f"workload {w.name}: {w.n_cells} x {w.n_genes}, nnz {nnz} (density {nnz / (n * w.n_genes):.4f}), "
            f"streamed in {ingest['blocks']} blocks, {ingest['total_ms']:.0f} ms (reader {ingest['read_ms']:.0f} ms)")
    else:
        t0 = time.perf_counter()
        layer = S.workload(args.workload)
        nnz = layer.nnz
        log(f"workload {w.name}: {w.n_cells} x {w.n_genes}, nnz {layer.nnz} (density {layer.density:.4f}), "
            f"generated in {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        csr = dev.upload_csr(layer.indptr, layer.indices, layer.values, w.n_genes)
        ingest = {"upload_ms": (time.perf_counter() - t0) * 1e3,
                  "bytes": int(layer.indptr.nbytes + layer.indices.nbytes + layer.values.nbytes)}
        log(f"layer uploaded from pageable host memory in {ingest['upload_ms']:.0f} ms "
            f"({ingest['bytes'] / 1e9:.2f} GB, column copy and checks included)")
    log("layer resident on", dev.info())
    k = 30 if args.workload == "4m" else 15
    if args.workload == "4m":
        args.n_lists = 8192
        args.no_compare = not args.compare

    def step():
        res = dev.pca(csr, 50, seed=0, power_iterations=args.q, download_embedding=False)
        g = dev.neighbors(res.device_embedding, k, precision=args.precision)
        return res, g

    for i in range(args.warmup):
        res, g = step()
        log(f"warmup {i}: pca {res.timing['total_ms']:.1f} ms, knn {g.info['total_ms']:.1f} ms")
        res.device_embedding.free()
    times, pcas, knns = [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        res, g = step()
        dt = time.perf_counter() - t0
        times.append(dt)
        pcas.append(res.timing)
        knns.append(g.info)
        log(f"step {i}: {dt * 1e3:.1f} ms (pca {res.timing['total_ms']:.1f} ms, knn {g.info['total_ms']:.1f} ms, "
            f"scan kernel {g.info['scan_kernel_ms']:.1f} ms, rows searched again {g.info['fallback']})")
        if i < args.steps - 1:
            res.device_embedding.free()
    ms = 1e3 * float(np.mean(times))
    value = n / (ms / 1e3)
    end_to_end = None
    if layer is not None:
        # The metric as BASELINE.md words it: the clock starts with the log1p CSR in
        # (pageable) host memory and stops with the graph on the host, so each step
        # uploads the layer (checks and the column copy included) before the same
        # pca + neighbors. One warm-up, then the same number of timed steps.
        res.device_embedding.free()
        e2e, uploads = [], []
        for i in range(args.steps + 1):
            t0 = time.perf_counter()
            csr.free()
            t1 = time.perf_counter()
            csr = dev.upload_csr(layer.indptr, layer.indices, layer.values, w.n_genes)
            up = time.perf_counter() - t1
            res, g = step()
            dt = time.perf_counter() - t1
            if i > 0:
                e2e.append(dt)
                uploads.append(up)
            if i < args.steps:
                res.device_embedding.free()
            log(f"end-to-end step {i}: {dt * 1e3:.1f} ms (upload {up * 1e3:.1f} ms)")
        e2e_ms = 1e3 * float(np.mean(e2e))
        end_to_end = {"value": n / (e2e_ms / 1e3), "unit": "cells/s", "ms_per_step": e2e_ms,
                      "upload_ms": 1e3 * float(np.mean(uploads)), "steps": args.steps,
                      "clock": "log1p CSR in pageable host memory -> kNN graph on the host (upload, checks and "
                               "column copy, pca, neighbors, graph download)"}
    emb = res.device_embedding
    d = 50
    peak = PEAKS_TFLOPS[args.precision]

    # Dominant kernel: the full candidate scan, 2 n^2 d flops a launch (unpadded d).
    scan_ms = float(np.mean([x.get("scan_kernel_ms", 0.0) or x["search_ms"] for x in knns]))
    achieved = 2.0 * n * n * d / (scan_ms / 1e3) / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("kernel") == "search_bf16" and args.precision == "bf16" and t.get("workload_name") == args.workload:
            traffic = t["bytes_per_launch"]
    roofline = {"bound": "mfma" if args.precision == "bf16" else "valu",
                "kernel": "search_bf16" if args.precision == "bf16" else "scan_points (FP32)",
                "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": traffic, "launch_ms": scan_ms,
                "algorithmic_per_launch": f"2 n^2 d = {2.0 * n * n * d:.3e} flop (n = {n}, d = {d})"}
    if args.precision == "bf16" and knns[-1].get("rounds", 1) == 1 and n >= 1 << 17:
        # The one-round scan is symmetric: each pair of rows is keyed once (for both rows)
        # by MFMAs over the K = 64 augmented components, n^2 / 2 pairs, so the matrix
        # cores execute 64 n^2 flops, not 2 n^2 d. Their own rate, stated beside the
        # algorithmic one above.
        executed = 2.0 * 64 * (n * n / 2.0)
        roofline["mfma_executed_per_launch"] = (f"2 K n^2 / 2 = {executed:.3e} flop (K = 64 augmented components, "
                                                f"each pair of rows keyed once)")
        roofline["mfma_executed_tflops"] = executed / (scan_ms / 1e3) / 1e12
        roofline["frac_mfma_executed"] = roofline["mfma_executed_tflops"] / peak

    def spmm_roofline(kind):
        kms = float(np.mean([p[f"{kind}_ms"] for p in pcas]))
        launches = pcas[-1][f"{kind}_launches"]
        nbytes = pcas[-1][f"{kind}_bytes"]
        gbs = nbytes / (kms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "launches": launches, "launch_ms": kms / max(launches, 1), "bytes_per_launch": nbytes / max(launches, 1)}

    def mean_fields(rows):
        keys = [key for key, v in rows[-1].items() if isinstance(v, (int, float))]
        return {key: float(np.mean([r[key] for r in rows])) for key in keys}

    out = {
        "metric": METRIC if args.workload == "1m" else f"cells/s pca(50)+neighbors({k}) @{args.workload}",
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
                               f"density {nnz / (n * w.n_genes):.4f}, nnz {nnz}",
                   "pca": f"n_components=50, block 60, q={args.q}, seed 0, FP32 SpMM, FP64 QR/Rayleigh-Ritz",
                   "neighbors": f"n_neighbors={k}, euclidean, {args.precision} candidates, exact f64 re-rank "
                                f"and certificate, exact search of uncertified rows",
                   "timed": "pca + neighbors with the layer resident in HBM (end_to_end: the same with the "
                            "upload from host memory); the learned router is not in the timed step: at "
                            "recall@15 >= 0.99 its confidence gate sends every row to this exact search "
                            "(field 'learned')"},
        "roofline": roofline,
        "roofline_spmm": {"column_product": spmm_roofline("column_product"),
                          "row_project": spmm_roofline("row_product")},
        "breakdown": {"pca": mean_fields(pcas), "pca_pass_ms": pcas[-1]["pass_ms"], "knn": mean_fields(knns)},
    }
    if end_to_end is not None:
        out["end_to_end"] = end_to_end
    if ingest is not None:
        out["ingest"] = ingest
    if args.workload == "4m":
        out["config"]["ingest"] = ("counts read from a kyto store" if args.from_store else "counts read") + \
            " in 32,768-cell blocks through pinned double buffers, normalize_total(1e4) + log1p on the device; " \
            "host holds two blocks"
    if not args.no_compare:
        # Outside the timed region: the plain FP32 exact path (kyto's exact graph, the
        # kNN stage the metric compares with), the BF16 graph's equality with it, and
        # the learned router's own recall.
        plain = dev.neighbors(emb, k, precision="fp32")
        exact = plain.rows()
        same = bool(np.array_equal(g.rows(), exact) and np.array_equal(g.distances, plain.distances))
        knn_ms = float(np.mean([x["total_ms"] for x in knns]))
        out[f"recall_at_{k}"] = recall_of(g.rows(), exact)
        out["recall_definition"] = f"share of the exact graph's {k - 1} stored neighbours (self excluded) found"
        out["graph_equal_to_plain_fp32"] = same
        out["plain_fp32_knn_ms"] = plain.info["total_ms"]
        out["knn_ms"] = knn_ms
        out["knn_speedup_vs_plain_fp32"] = plain.info["total_ms"] / knn_ms
        out["learned"] = learned_leg(dev, emb, k, exact, args)
        log(f"recall@{k} {out[f'recall_at_{k}']:.5f} (graph equal to FP32: {same}); kNN {knn_ms:.1f} ms vs plain FP32 "
            f"{plain.info['total_ms']:.1f} ms ({out['knn_speedup_vs_plain_fp32']:.1f}x); learned {out['learned']}")
    # Outside the timed region: the connectivities graph the neighbors step declares
    # (UMAP's fuzzy weights on the last step's graph), the second of two calls timed.
    dev.connectivities(g)
    conn = dev.connectivities(g)
    out["connectivities"] = {"edges": int(conn.info["edges"]), "ms": conn.info["total_ms"],
                             "upload_ms": conn.info["upload_ms"], "scales_ms": conn.info["scales_ms"],
                             "union_ms": conn.info["union_ms"], "zero_rho_rows": int(conn.info["zero_rho_rows"])}
    log(f"connectivities: {conn.info['edges']} edges in {conn.info['total_ms']:.1f} ms "
        f"(upload {conn.info['upload_ms']:.1f}, scales {conn.info['scales_ms']:.1f}, union {conn.info['union_ms']:.1f})")
    if layer is not None and not args.no_store:
        out["store"] = store_leg(dev, layer, csr, w.n_genes)
        log(f"store: {out['store']}")
    if not args.no_cpu_baseline:
        import oracle
        t0 = time.perf_counter()
        ref = oracle.connectivities(g.rows(), g.distances.reshape(-1, k - 1), k)
        out["connectivities"]["cpu_oracle_ms"] = (time.perf_counter() - t0) * 1e3
        out["connectivities"]["cpu_oracle_cores"] = 1
        out["connectivities"]["equal_to_oracle"] = bool(
            np.array_equal(ref.indices, conn.indices) and
            np.array_equal(ref.weights.view(np.uint32), conn.weights.view(np.uint32)))
    embedding = emb.download()
    emb.free()
    if not args.no_cpu_baseline and layer is not None:
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or min(os.cpu_count() or 1, 16)
        out["cpu_baseline"], out["fp64_errors"] = cpu_baseline(layer, embedding, w.n_genes, args.q, threads, res)
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        copy_calibration()


if __name__ == "__main__":
    main()
