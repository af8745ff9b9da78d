"""Device handles over libkyto_gpu.so: a context, resident CSR layers and dense
matrices, and the pca / neighbors / IVF entry points as numpy-facing calls.

Mirrors kyto-gpu's DeviceKernels (crates/kyto-gpu/src/launch.rs:42-93): one handle
per device, results returned as owned arrays, failures raised as KytoError with
kyto's codes (KY-E215 arguments/data, KY-E501 no kernel, KY-E504 no device,
KY-E505 device failure).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_uint32, c_uint64, c_void_p
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


def device_count() -> int:
    n = ctypes.c_int(0)
    N.check(N.gpu_lib().ky_device_count(byref(n)))
    return int(n.value)


@dataclass
class PcaOutput:
    embedding: np.ndarray | None   # cells x k f32 (None when left on the device only)
    loadings: np.ndarray           # genes x k f32
    variance: np.ndarray           # k f64
    variance_ratio: np.ndarray     # k f64
    means: np.ndarray              # genes f64
    passes: int
    timing: dict = field(default_factory=dict)
    device_embedding: "DenseMatrix | None" = None
    made_values: "np.ndarray | None" = None  # the layer's values when the device made it (kept for a sink)


@dataclass
class Graph:
    """kyto's NeighborGraph (neighbors.rs:52-175): CSR, k = n_neighbors - 1 a row."""
    n_neighbors: int
    indptr: np.ndarray     # u64, n + 1
    indices: np.ndarray    # u32, n * k, increasing within a row
    distances: np.ndarray  # f32
    info: dict = field(default_factory=dict)

    @property
    def k(self) -> int:
        return self.n_neighbors - 1

    def rows(self) -> np.ndarray:
        return self.indices.reshape(-1, self.k)


@dataclass
class ConnectivityGraph:
    """kyto's ConnectivityGraph (connectivities.rs:33-72): symmetric CSR, columns
    increasing in each row, float32 weights in (0, 1], no stored zeros."""
    indptr: np.ndarray   # u64, n + 1
    indices: np.ndarray  # u32
    weights: np.ndarray  # f32
    sigma: np.ndarray    # f32, each row's sigma
    rho: np.ndarray      # f32, each row's rho
    info: dict = field(default_factory=dict)

    @property
    def n_cells(self) -> int:
        return len(self.indptr) - 1


class Device:
    """One GPU context (ky_open)."""

    def __init__(self, ordinal: int = 0):
        self._lib = N.gpu_lib()
        h = c_void_p()
        N.check(self._lib.ky_open(int(ordinal), byref(h)))
        self._h = h
        self.ordinal = ordinal

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            N.check(self._lib.ky_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ run control --

    def run_begin(self, total_rows: int) -> None:
        """A new run on this context: the cancel flag cleared, no rows done."""
        N.check(self._lib.ky_run_begin(self._h, int(total_rows)))

    def cancel(self) -> None:
        """Stop the run at its next pass, block or round (KY-E503); any thread."""
        N.check(self._lib.ky_cancel(self._h))

    def progress(self) -> tuple[int, int]:
        """(cell rows computed, cell rows of the run); any thread."""
        done, total = c_uint64(0), c_uint64(0)
        N.check(self._lib.ky_progress(self._h, byref(done), byref(total)))
        return int(done.value), int(total.value)

    def trim(self) -> None:
        """Release the context's pass buffers and this thread's cached device blocks."""
        N.check(self._lib.ky_trim(self._h))

    def info(self) -> dict:
        name = ctypes.create_string_buffer(128)
        total = c_uint64(0)
        N.check(self._lib.ky_device_info(self._h, name, 128, byref(total)))
        return {"name": name.value.decode(), "memory_bytes": int(total.value)}

    # ------------------------------------------------------------ residency --

    def upload_csr(self, indptr, indices, values, n_genes: int) -> "CsrLayer":
        return CsrLayer(self, indptr, indices, values, n_genes)

    def stream_counts(self, n_rows: int, n_genes: int, read_block, block_rows: int = 32_768, nnz_hint: int = 0,
                      target_sum=1e4, log1p=True) -> "CsrLayer":
        """A counts layer read block by block through pinned double buffers
        (ky_csr_stream), made into the layer on the device as upload_counts does.
        read_block(first, n_rows, indptr, indices, values, capacity) fills the
        library's pinned buffers (ctypes pointers) and returns the block's nnz."""
        def trampoline(user, first, count, indptr, indices, values, capacity):
            try:
                return int(read_block(int(first), int(count), indptr, indices, values, int(capacity)))
            except Exception:  # noqa: BLE001 - a failure is reported to the library as -1
                return -1

        cb = N.BlockReader(trampoline)
        h = c_void_p()
        info = N.StreamInfo()
        N.check(self._lib.ky_csr_stream(self._h, int(n_rows), int(n_genes), int(block_rows), cb, None, int(nnz_hint),
                                        float(target_sum or 0.0), 1 if log1p else 0, byref(h), byref(info)))
        layer = CsrLayer._adopt(self, h, int(n_rows), int(n_genes), int(info.nnz))
        layer.stream_info = _info(info)
        return layer

    def stream_layer(self, layer, block_rows: int = 0, target_sum=None, log1p=False) -> "CsrLayer":
        """A kyto store's layer (kyto_amd.store.StoreLayer) read cell block by cell block
        through the pinned double buffers (ky_csr_stream_layer: the tiles decode
        straight into them), with normalize_total(target_sum) / log1p made on the
        device as upload_counts does. block_rows 0: eight cell blocks."""
        h = c_void_p()
        info = N.StreamInfo()
        N.check(self._lib.ky_csr_stream_layer(self._h, layer.handle, int(block_rows), float(target_sum or 0.0),
                                              1 if log1p else 0, byref(h), byref(info)))
        out = CsrLayer._adopt(self, h, layer.n_cells, layer.n_genes, int(info.nnz))
        out.stream_info = _info(info)
        return out

    def upload_counts(self, indptr, indices, counts, n_genes: int, target_sum=1e4, log1p=True) -> "CsrLayer":
        """A counts CSR made into the layer on the device: kyto's normalize_total
        (target_sum; None skips it) then log1p (ky_csr_upload_counts)."""
        return CsrLayer(self, indptr, indices, counts, n_genes, target_sum=target_sum, log1p=log1p)

    def upload_dense(self, x) -> "DenseMatrix":
        return DenseMatrix.upload(self, x)

    # -------------------------------------------------------------------- pca --

    def pca(self, layer: "CsrLayer", n_components: int = 50, seed: int = 0, power_iterations: int = 9,
            download_embedding: bool = True, keep_on_device: bool = True) -> PcaOutput:
        genes = layer.n_genes
        k = int(n_components)
        emb = np.empty((layer.n_rows, k), np.float32) if download_embedding else None
        loadings = np.empty((genes, k), np.float32)
        variance = np.empty(k)
        ratio = np.empty(k)
        means = np.empty(genes)
        passes = c_uint32(0)
        timing = N.PcaTiming()
        dh = c_void_p()
        N.check(self._lib.ky_pca(self._h, layer.handle, k, int(seed), int(power_iterations),
                                 N.ptr(emb, N.f32p) if emb is not None else None,
                                 byref(dh) if keep_on_device else None,
                                 N.ptr(loadings, N.f32p), N.ptr(variance, N.f64p), N.ptr(ratio, N.f64p),
                                 N.ptr(means, N.f64p), byref(passes), byref(timing)))
        dev = DenseMatrix._adopt(self, dh, layer.n_rows, k) if keep_on_device else None
        t = {"pass_ms": [timing.pass_ms[i] for i in range(min(16, timing.n_passes))],
             "dense_ms": timing.dense_ms, "total_ms": timing.total_ms,
             "row_product_ms": timing.spmm_ms[0], "row_product_launches": timing.spmm_launches[0],
             "column_product_ms": timing.spmm_ms[1], "column_product_launches": timing.spmm_launches[1],
             "row_product_bytes": timing.spmm_bytes[0], "column_product_bytes": timing.spmm_bytes[1]}
        return PcaOutput(emb, loadings, variance, ratio, means, int(passes.value), t, dev)

    def pca_sketch(self, layer, seed, l, cell_offset=0):
        g = layer.n_genes
        sketch = np.empty((g, l))
        sums = np.empty(g)
        osum = np.empty(l)
        N.check(self._lib.ky_pca_sketch(self._h, layer.handle, int(seed), int(cell_offset), int(l),
                                        N.ptr(sketch, N.f64p), N.ptr(sums, N.f64p), N.ptr(osum, N.f64p)))
        return sketch, sums, osum

    def pca_product(self, layer, q, block_means):
        q = np.ascontiguousarray(q, np.float64)
        bm = np.ascontiguousarray(block_means, np.float64)
        g, w = q.shape
        prod = np.empty((g, w))
        sums = np.empty(w)
        N.check(self._lib.ky_pca_product(self._h, layer.handle, N.ptr(q, N.f64p), w, N.ptr(bm, N.f64p),
                                         N.ptr(prod, N.f64p), N.ptr(sums, N.f64p)))
        return prod, sums

    def pca_project(self, layer, v, mean_projection, means):
        v = np.ascontiguousarray(v, np.float64)
        mp = np.ascontiguousarray(mean_projection, np.float64)
        mu = np.ascontiguousarray(means, np.float64)
        g, k = v.shape
        emb = np.empty((layer.n_rows, k), np.float32)
        sq = np.empty(g)
        st = np.empty(g, np.uint64)
        N.check(self._lib.ky_pca_project(self._h, layer.handle, N.ptr(v, N.f64p), k, N.ptr(mp, N.f64p),
                                         N.ptr(mu, N.f64p), N.ptr(emb, N.f32p), None, N.ptr(sq, N.f64p),
                                         N.ptr(st, N.u64p)))
        return emb, sq, st

    # -------------------------------------------------------------- neighbours --

    def neighbors(self, points: "DenseMatrix", n_neighbors: int = 15, n_components: int | None = None,
                  metric: str = "euclidean", precision: str = "fp32", round_rows: int = 0) -> Graph:
        """Exact neighbours (ky_neighbors_rounds): the candidate search in rounds of
        round_rows queries (0: one round)."""
        n = points.rows
        k = int(n_neighbors) - 1
        if metric not in N.METRIC:
            raise N.KytoError(N.KY_E215, f"neighbors: metric {metric!r} is not 'euclidean' or 'cosine'")
        if precision not in N.PRECISION:
            raise N.KytoError(N.KY_E215, f"neighbors: precision {precision!r} is not 'fp32' or 'bf16'")
        idx = np.empty(max(n * k, 1), np.uint32)
        dist = np.empty(max(n * k, 1), np.float32)
        info = N.KnnInfo()
        N.check(self._lib.ky_neighbors_rounds(self._h, points.handle, int(n_components or 0), int(n_neighbors),
                                              N.METRIC[metric], N.PRECISION[precision], int(round_rows),
                                              N.ptr(idx, N.u32p), N.ptr(dist, N.f32p), byref(info)))
        indptr = np.arange(0, n * k + 1, k, dtype=np.uint64)
        return Graph(int(n_neighbors), indptr, idx[: n * k], dist[: n * k], _info(info))

    def neighbors_bytes(self, rows: int, cols: int, n_neighbors: int = 15, n_components: int | None = None,
                        metric: str = "euclidean", precision: str = "bf16", round_rows: int = 0) -> int:
        """Device scratch of a neighbours search (ky_neighbors_bytes)."""
        return neighbors_bytes(rows, cols, n_neighbors, n_components, metric, precision, round_rows, self._h)

    def neighbors_plan(self, rows: int, cols: int, limit: int, n_neighbors: int = 15, n_components: int | None = None,
                       metric: str = "euclidean", precision: str = "bf16") -> tuple[int, int]:
        """(round_rows, bytes) of the largest round that fits `limit` bytes; KytoError
        KY-E401 when none does (ky_neighbors_plan)."""
        return neighbors_plan(rows, cols, limit, n_neighbors, n_components, metric, precision, self._h)

    def connectivities(self, graph: Graph) -> ConnectivityGraph:
        """scanpy's connectivities of a neighbour graph (ky_connectivities)."""
        n = len(graph.indptr) - 1
        k = graph.n_neighbors - 1
        idx = np.ascontiguousarray(graph.indices, np.uint32)
        dist = np.ascontiguousarray(graph.distances, np.float32)
        if len(idx) != n * k or len(dist) != n * k:
            raise N.KytoError(N.KY_E215, f"connectivities: {len(idx)} indices and {len(dist)} distances for "
                                         f"{n} rows of {k}")
        sigma = np.empty(max(n, 1), np.float32)
        rho = np.empty(max(n, 1), np.float32)
        h = c_void_p()
        info = N.ConnInfo()
        N.check(self._lib.ky_connectivities(self._h, n, int(graph.n_neighbors), N.ptr(idx, N.u32p),
                                            N.ptr(dist, N.f32p), byref(h), N.ptr(sigma, N.f32p),
                                            N.ptr(rho, N.f32p), byref(info)))
        try:
            edges = int(info.edges)
            indptr = np.empty(n + 1, np.uint64)
            cols = np.empty(max(edges, 1), np.uint32)
            w = np.empty(max(edges, 1), np.float32)
            N.check(self._lib.ky_graph_download(h, N.ptr(indptr, N.u64p), N.ptr(cols, N.u32p), N.ptr(w, N.f32p)))
        finally:
            N.check(self._lib.ky_graph_free(h))
        return ConnectivityGraph(indptr, cols[:edges], w[:edges], sigma[:n], rho[:n], _info(info))

    def nearest_keys(self, queries, points: "DenseMatrix", m: int, key: str = "squared_euclidean", skip=None,
                     query_norms=None, point_norms=None):
        """DeviceKernels::nearest_keys (ky_nearest_keys): each query row's m smallest f32
        keys against every point, the point skip[i] left out (UINT32_MAX: none), ranked
        by (key, point). Returns (keys f32, rows u32), n_queries x m."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        sk = np.full(nq, 0xFFFFFFFF, np.uint32) if skip is None else np.ascontiguousarray(skip, np.uint32)
        keys = np.empty((max(nq, 1), m), np.float32)
        rows = np.empty((max(nq, 1), m), np.uint32)
        qn = None if query_norms is None else np.ascontiguousarray(query_norms, np.float32)
        pn = None if point_norms is None else np.ascontiguousarray(point_norms, np.float32)
        N.check(self._lib.ky_nearest_keys(self._h, N.ptr(q, N.f32p), nq, points.handle, N.KEY[key],
                                          None if qn is None else N.ptr(qn, N.f32p),
                                          None if pn is None else N.ptr(pn, N.f32p), N.ptr(sk, N.u32p), int(m),
                                          N.ptr(keys, N.f32p), N.ptr(rows, N.u32p)))
        return keys[:nq], rows[:nq]

    # ---------------------------------------------------------------------- IVF --

    def ivf_build(self, points: "DenseMatrix", n_lists=4096, kmeans_iters=10, seed=0, hidden=256,
                  train_queries=200_000, batch=4096, epochs=4, lr=1e-3, label_k=14) -> "IvfIndex":
        p = N.IvfParams(n_lists, kmeans_iters, seed, hidden, train_queries, batch, epochs, lr, label_k)
        h = c_void_p()
        info = N.IvfInfo()
        N.check(self._lib.ky_ivf_build(self._h, points.handle, byref(p), byref(h), byref(info)))
        return IvfIndex(self, h, {f: getattr(info, f) for f, _ in N.IvfInfo._fields_})


def set_threads(threads: int) -> None:
    """Host threads of the library's worker pool (0: every hardware thread, at most 32)."""
    N.check(N.gpu_lib().ky_set_threads(int(threads)))


def pca_bytes(n_rows, n_cols, nnz, n_components, power_iterations) -> int:
    """Device memory of a resident layer and one pca() over it (ky_pca_bytes)."""
    out = c_uint64(0)
    N.check(N.gpu_lib().ky_pca_bytes(int(n_rows), int(n_cols), int(nnz), int(n_components), int(power_iterations),
                                     byref(out)))
    return int(out.value)


def connectivities_bytes(n_cells, n_neighbors) -> int:
    """Device memory of ky_connectivities (ky_connectivities_bytes)."""
    out = c_uint64(0)
    N.check(N.gpu_lib().ky_connectivities_bytes(int(n_cells), int(n_neighbors), byref(out)))
    return int(out.value)


def neighbors_bytes(rows, cols, n_neighbors=15, n_components=None, metric="euclidean", precision="bf16",
                    round_rows=0, handle=None) -> int:
    """Device scratch of a neighbours search (ky_neighbors_bytes; no device needed)."""
    out = c_uint64(0)
    N.check(N.gpu_lib().ky_neighbors_bytes(handle, int(rows), int(cols), int(n_components or 0), int(n_neighbors),
                                           N.METRIC[metric], N.PRECISION[precision], int(round_rows), byref(out)))
    return int(out.value)


def neighbors_plan(rows, cols, limit, n_neighbors=15, n_components=None, metric="euclidean", precision="bf16",
                   handle=None) -> tuple[int, int]:
    """(round_rows, bytes) of the largest round of a neighbours search whose scratch fits
    `limit` bytes; KytoError KY-E401 when none does (ky_neighbors_plan)."""
    rr, b = c_uint64(0), c_uint64(0)
    N.check(N.gpu_lib().ky_neighbors_plan(handle, int(rows), int(cols), int(n_components or 0), int(n_neighbors),
                                          N.METRIC[metric], N.PRECISION[precision], int(limit), byref(rr), byref(b)))
    return int(rr.value), int(b.value)


def _info(info) -> dict:
    return {f: getattr(info, f) for f, _ in type(info)._fields_}


class CsrLayer:
    """A CSR layer resident on the device (ky_csr_upload)."""

    def __init__(self, dev: Device, indptr, indices, values, n_genes: int, target_sum=None, log1p=False):
        """values as given (ky_csr_upload); or counts made into the layer on the device
        (ky_csr_upload_counts): normalize_total(target_sum) when target_sum is set,
        then log1p when `log1p`."""
        self.dev = dev
        indptr = np.ascontiguousarray(indptr, np.uint64)
        indices = np.ascontiguousarray(indices, np.uint32)
        values = np.ascontiguousarray(values, np.float32)
        self.n_rows = len(indptr) - 1
        self.n_genes = int(n_genes)
        self.nnz = int(indptr[-1]) if len(indptr) else 0
        h = c_void_p()
        if target_sum is None and not log1p:
            N.check(dev._lib.ky_csr_upload(dev.handle, N.ptr(indptr, N.u64p), N.ptr(indices, N.u32p),
                                           N.ptr(values, N.f32p), self.n_rows, self.n_genes, byref(h)))
        else:
            N.check(dev._lib.ky_csr_upload_counts(dev.handle, N.ptr(indptr, N.u64p), N.ptr(indices, N.u32p),
                                                  N.ptr(values, N.f32p), self.n_rows, self.n_genes,
                                                  float(target_sum or 0.0), 1 if log1p else 0, byref(h)))
        self._h = h

    @classmethod
    def _adopt(cls, dev, h, n_rows, n_genes, nnz) -> "CsrLayer":
        self = cls.__new__(cls)
        self.dev, self._h, self.n_rows, self.n_genes, self.nnz = dev, h, n_rows, n_genes, nnz
        return self

    @property
    def handle(self):
        return self._h

    def values(self) -> np.ndarray:
        """The resident values (nnz floats, row order): ky_csr_values."""
        out = np.empty(max(self.nnz, 1), np.float32)
        N.check(self.dev._lib.ky_csr_values(self._h, N.ptr(out, N.f32p)))
        return out[: self.nnz]

    def free(self):
        if self._h:
            N.check(self.dev._lib.ky_csr_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DenseMatrix:
    """A row-major f32 matrix resident on the device."""

    def __init__(self):
        self.dev = None
        self._h = None
        self.rows = 0
        self.cols = 0

    @classmethod
    def upload(cls, dev: Device, x) -> "DenseMatrix":
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2:
            raise N.KytoError(N.KY_E215, "a dense matrix is 2-D")
        h = c_void_p()
        N.check(dev._lib.ky_dense_upload(dev.handle, N.ptr(x, N.f32p), x.shape[0], x.shape[1], byref(h)))
        return cls._adopt(dev, h, x.shape[0], x.shape[1])

    @classmethod
    def _adopt(cls, dev, h, rows, cols) -> "DenseMatrix":
        m = cls()
        m.dev, m._h, m.rows, m.cols = dev, h, int(rows), int(cols)
        return m

    @property
    def handle(self):
        return self._h

    def download(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), np.float32)
        N.check(self.dev._lib.ky_dense_download(self._h, N.ptr(out, N.f32p)))
        return out

    def free(self):
        if self._h:
            N.check(self.dev._lib.ky_dense_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class IvfIndex:
    """A learned IVF index on the device (ky_ivf_build)."""

    def __init__(self, dev: Device, h, info: dict):
        self.dev, self._h, self.info = dev, h, info

    def search(self, points: DenseMatrix, n_neighbors=15, n_probe=8, use_router=True, confidence=0.0) -> Graph:
        n = points.rows
        k = int(n_neighbors) - 1
        idx = np.empty(n * k, np.uint32)
        dist = np.empty(n * k, np.float32)
        info = N.IvfSearchInfo()
        N.check(self.dev._lib.ky_ivf_search(self.dev.handle, self._h, points.handle, int(n_neighbors), int(n_probe),
                                            1 if use_router else 0, float(confidence), N.ptr(idx, N.u32p),
                                            N.ptr(dist, N.f32p), byref(info)))
        indptr = np.arange(0, n * k + 1, k, dtype=np.uint64)
        return Graph(int(n_neighbors), indptr, idx, dist, _info(info))

    def layout(self, n_rows: int, d: int, n_lists: int) -> dict:
        """The k-means layout and the router's alpha (ky_ivf_layout)."""
        c = np.empty((n_lists, d), np.float32)
        assign = np.empty(n_rows, np.uint32)
        ptr = np.empty(n_lists + 1, np.uint32)
        rows = np.empty(n_rows, np.uint32)
        alpha = ctypes.c_float(0)
        N.check(self.dev._lib.ky_ivf_layout(self._h, N.ptr(c, N.f32p), N.ptr(assign, N.u32p), N.ptr(ptr, N.u32p),
                                            N.ptr(rows, N.u32p), byref(alpha)))
        return {"centroids": c, "assign": assign, "list_ptr": ptr, "list_rows": rows, "alpha": alpha.value}

    def free(self):
        if self._h:
            N.check(self.dev._lib.ky_ivf_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# Host-only FP64 dense steps (no device needed).

def dense_extend_basis(basis: np.ndarray, block: np.ndarray) -> np.ndarray:
    lib = N.gpu_lib()
    basis = np.ascontiguousarray(basis, np.float64)
    block = np.ascontiguousarray(block, np.float64)
    rows, bw = block.shape
    bc = basis.shape[1] if basis.size else 0
    out = np.zeros((rows, bw))
    kept = c_uint32(0)
    N.check(lib.ky_dense_extend_basis(N.ptr(basis, N.f64p) if bc else None, rows, bc, N.ptr(block, N.f64p), bw,
                                      N.ptr(out, N.f64p), byref(kept)))
    return out[:, : kept.value].copy()


def dense_rayleigh_ritz(basis: np.ndarray, products: np.ndarray, k: int):
    lib = N.gpu_lib()
    basis = np.ascontiguousarray(basis, np.float64)
    products = np.ascontiguousarray(products, np.float64)
    rows, cols = basis.shape
    ev = np.zeros(k)
    lo = np.zeros((rows, k))
    N.check(lib.ky_dense_rayleigh_ritz(N.ptr(basis, N.f64p) if cols else None,
                                       N.ptr(products, N.f64p) if cols else None, rows, cols, k,
                                       N.ptr(ev, N.f64p), N.ptr(lo, N.f64p)))
    return ev, lo
