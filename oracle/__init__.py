"""ORACLE — test infrastructure only.

A CPU restatement of kyto's CPU path for ``pca()`` and ``neighbors()``: the heavy
loops in C (``kyto_oracle.c``, built by ``Makefile`` into ``_build/``), the driver and
the small FP64 dense steps in numpy (LAPACK's SVD and symmetric eigensolver stand in
for ``faer``'s, which kyto pins only through tolerances).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker. The product
(``kyto_amd``) never imports it.

Every function cites the reference file:line it follows (paths relative to the
reference checkout).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# KYTO_AMD_SANITIZED=1 (tests/sanitize.sh): the AddressSanitizer + UBSan build.
_SANITIZED = os.environ.get("KYTO_AMD_SANITIZED") == "1"
_LIB_PATH = os.path.join(_HERE, "_build", "libkyto_oracle_san.so" if _SANITIZED else "libkyto_oracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)

# kyto-plan constants: physical_plan.rs:73 (q) and estimate.rs:153 (oversamples).
PCA_POWER_ITERATIONS = 9
PCA_OVERSAMPLES = 10
# pca_dense.rs:30
DROP_TOLERANCE = 1e-10


def build() -> str:
    """Compile the C restatement (``make -C oracle``)."""
    subprocess.run(["make", "-s", "-C", _HERE] + (["sanitize"] if _SANITIZED else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ko_mix.restype = ctypes.c_uint64
        L.ko_mix.argtypes = [ctypes.c_uint64]
        L.ko_omega_row.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _f64p, ctypes.c_size_t]
        L.ko_log1p.restype = ctypes.c_float
        L.ko_log1p.argtypes = [ctypes.c_float]
        L.ko_normalize_log1p.argtypes = [_u64p, _f32p, ctypes.c_size_t, ctypes.c_double, _f32p, _f32p]
        L.ko_pca_sketch.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, _u64p, ctypes.c_uint64,
                                    ctypes.c_size_t, _u64p, _f64p, _f64p, _f64p]
        L.ko_pca_product.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t,
                                     _f64p, _f64p, _f64p]
        L.ko_pca_rows.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t,
                                  _f64p, _f64p, _f64p, _f64p, _u64p]
        L.ko_pca_sketch_tree.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, _u64p, ctypes.c_uint64,
                                         ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _u64p, _f64p, _f64p,
                                         _f64p]
        L.ko_pca_product_tree.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t, _f64p,
                                          ctypes.c_size_t, ctypes.c_size_t, _f64p, _f64p]
        L.ko_pca_rows_tree.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, _f64p, ctypes.c_size_t, _f64p, _f64p,
                                       ctypes.c_size_t, ctypes.c_size_t, _f64p, _f64p, _u64p]
        L.ko_gram.argtypes = [_u64p, _u32p, _f32p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _f64p, _f64p]
        L.ko_knn_exact_list.restype = ctypes.c_int
        L.ko_knn_exact_list.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t,
                                        _u32p, ctypes.c_size_t, ctypes.c_size_t, _f64p, _u32p, _u64p]
        L.ko_squared_distance.restype = ctypes.c_double
        L.ko_squared_distance.argtypes = [_f32p, _f32p, ctypes.c_size_t]
        L.ko_dot.restype = ctypes.c_double
        L.ko_dot.argtypes = [_f32p, _f32p, ctypes.c_size_t]
        L.ko_nearest_offer_many.restype = ctypes.c_size_t
        L.ko_nearest_offer_many.argtypes = [ctypes.c_size_t, _f64p, _u32p, ctypes.c_size_t, _f64p, _u32p]
        L.ko_norms.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, _f64p]
        L.ko_search_rows.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, _f64p, ctypes.c_size_t,
                                     _u32p, ctypes.c_size_t, _f64p, _u32p]
        L.ko_knn_exact.restype = ctypes.c_int
        L.ko_knn_exact.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_size_t, _f64p, _u32p, _u64p]
        L.ko_knn_exact_range.restype = ctypes.c_int
        L.ko_knn_exact_range.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t,
                                         ctypes.c_size_t, ctypes.c_size_t, _f64p, _u32p, _u64p]
        L.ko_round_up.restype = ctypes.c_float
        L.ko_round_up.argtypes = [ctypes.c_double]
        L.ko_smooth_knn.argtypes = [_f32p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double, _f32p, _f32p]
        L.ko_connectivities.restype = ctypes.c_size_t
        L.ko_connectivities.argtypes = [ctypes.c_size_t, ctypes.c_size_t, _u32p, _f32p, _u64p, _u32p, _f32p,
                                        _f32p, _f32p]
        L.ko_max_threads.restype = ctypes.c_int
        L.ko_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


# ------------------------------------------------------------------ input layer --

def mix(v: int) -> int:
    """SplitMix64's output function (pca.rs:946-950)."""
    return int(lib().ko_mix(ctypes.c_uint64(v & 0xFFFFFFFFFFFFFFFF)))


def omega_row(seed: int, cell: int, l: int) -> np.ndarray:
    """Omega's row of a cell (pca.rs:936-943)."""
    out = np.empty(l, np.float64)
    lib().ko_omega_row(seed, cell, _p(out, _f64p), l)
    return out


def log1p32(x: np.ndarray) -> np.ndarray:
    """kyto's log1p (kernels/log1p.rs): ln(1 + x) in f64, rounded once to f32."""
    return np.log1p(np.asarray(x, np.float32).astype(np.float64)).astype(np.float32)


def normalize_log1p(indptr, counts, target_sum=1e4):
    """normalize_total(target_sum) then log1p (elementwise.rs:262-269, kernels.rs
    DivideRows, kernels/log1p.rs). Returns (normalized, log1p) f32 values."""
    indptr = np.ascontiguousarray(indptr, np.uint64)
    counts = np.ascontiguousarray(counts, np.float32)
    norm = np.empty_like(counts)
    out = np.empty_like(counts)
    lib().ko_normalize_log1p(_p(indptr, _u64p), _p(counts, _f32p), len(indptr) - 1, target_sum,
                             _p(norm, _f32p), _p(out, _f32p))
    return norm, out


# ------------------------------------------------------------------- dense FP64 --

def extend_basis(basis: np.ndarray, block: np.ndarray) -> np.ndarray:
    """pca_dense.rs:105-129: project the basis off twice (CGS2), thin SVD, keep the
    left singular vectors whose singular value exceeds DROP_TOLERANCE x ||block||_F
    (the norm taken before the projection)."""
    rows = block.shape[0]
    scale = float(np.linalg.norm(block))
    if scale == 0.0 or block.shape[1] == 0:
        return np.zeros((rows, 0))
    rest = block.copy()
    if basis.shape[1] > 0:
        for _ in range(2):
            rest -= basis @ (basis.T @ rest)
    u, s, _ = np.linalg.svd(rest, full_matrices=False)
    keep = 0
    while keep < len(s) and s[keep] > DROP_TOLERANCE * scale:
        keep += 1
    return np.ascontiguousarray(u[:, :keep])


def complement(columns: np.ndarray, count: int) -> np.ndarray:
    """pca_dense.rs:215-252: each time the unit vector of the gene least covered so
    far (the first on ties), the columns projected off twice."""
    allc = columns.copy()
    added = np.zeros((columns.shape[0], 0))
    for _ in range(count):
        covered = (allc * allc).sum(axis=1)
        gene = int(np.argmin(covered))
        v = np.zeros(allc.shape[0])
        v[gene] = 1.0
        for _ in range(2):
            v = v - allc @ (allc.T @ v)
        v = v / np.sqrt((v * v).sum())
        allc = np.hstack([allc, v[:, None]])
        added = np.hstack([added, v[:, None]])
    return added


def flip_signs(loadings: np.ndarray) -> None:
    """pca_dense.rs:256-273 (scikit-learn's svd_flip(u_based_decision=False)): each
    column's largest absolute value, the first on ties, made positive."""
    for c in range(loadings.shape[1]):
        col = loadings[:, c]
        at = int(np.argmax(np.abs(col)))  # first maximum
        if col[at] < 0:
            loadings[:, c] = -col


def rayleigh_ritz(basis: np.ndarray, products: np.ndarray, k: int):
    """pca_dense.rs:141-174: T = B^T (C B) symmetrised, its eigenpairs, the k
    largest (negatives clamped to 0), V = B U, completed by zero-variance directions,
    signs flipped."""
    rows = basis.shape[0]
    loadings = np.zeros((rows, 0))
    eig = []
    if basis.shape[1] > 0:
        t = basis.T @ products
        sym = 0.5 * (t + t.T)
        vals, vecs = np.linalg.eigh(sym)
        n = len(vals)
        taken = min(k, n)
        order = [n - 1 - i for i in range(taken)]
        eig = [max(vals[i], 0.0) for i in order]
        loadings = basis @ vecs[:, order]
    missing = k - len(eig)
    if missing > 0:
        loadings = np.hstack([loadings, complement(loadings, missing)])
        eig = eig + [0.0] * missing
    loadings = np.ascontiguousarray(loadings)
    flip_signs(loadings)
    return np.array(eig), loadings


# ------------------------------------------------------------------------- PCA --

@dataclass
class PcaResult:
    embedding: np.ndarray      # cells x k, f32
    loadings: np.ndarray       # genes x k, f32
    variance: np.ndarray       # k, f64
    variance_ratio: np.ndarray  # k, f64
    means: np.ndarray          # genes, f64
    eigenvalues: np.ndarray
    loadings64: np.ndarray
    passes: int                # product passes run


# kyto's default store tiling: cell blocks of 4,096 cells (crates/kyto-core/src/schema.rs:16).
CELL_BLOCK = 4096


def pca(indptr, indices, values, n_genes: int, n_components: int = 50, seed: int = 0,
        q: int = PCA_POWER_ITERATIONS, cells=None, leaf_rows: int | None = CELL_BLOCK) -> PcaResult:
    """kyto's streamed block-Krylov PCA (pca.rs): pass 0 sketch (:588-622), end_pass
    (:712-801), q product passes (:623-645), the Rayleigh-Ritz step (finish_basis
    :848-862), the rows pass (:649-677) and finish (:804-845). Each pass folds leaves
    of `leaf_rows` rows (kyto's store cell blocks) in parallel and combines them in
    kyto's fixed tree (reduction.rs:30-78); leaf_rows=None folds every row into one
    leaf, in cell order."""
    L = lib()
    indptr = np.ascontiguousarray(indptr, np.uint64)
    indices = np.ascontiguousarray(indices, np.uint32)
    values = np.ascontiguousarray(values, np.float32)
    n = len(indptr) - 1
    if cells is None:
        cells = np.arange(n, dtype=np.uint64)
    cells = np.ascontiguousarray(cells, np.uint64)
    if n < 2:
        raise ValueError("KY-E215: fewer than 2 cells reach pca")
    if n_components > min(n, n_genes):
        raise ValueError("KY-E215: n_components above min(cells, genes)")
    width = min(n_components + PCA_OVERSAMPLES, n_genes)
    a = (_p(indptr, _u64p), _p(indices, _u32p), _p(values, _f32p), n)
    tree = leaf_rows is not None and n > leaf_rows

    n_cells = np.zeros(1, np.uint64)
    sums = np.zeros(n_genes)
    omega_sums = np.zeros(width)
    sketch = np.zeros((n_genes, width))
    if tree:
        L.ko_pca_sketch_tree(*a, _p(cells, _u64p), seed, width, n_genes, leaf_rows, _p(n_cells, _u64p),
                             _p(sums, _f64p), _p(omega_sums, _f64p), _p(sketch, _f64p))
    else:
        L.ko_pca_sketch(*a, _p(cells, _u64p), seed, width, _p(n_cells, _u64p), _p(sums, _f64p),
                        _p(omega_sums, _f64p), _p(sketch, _f64p))
    means = sums / float(n)
    if not np.all(np.isfinite(means)):
        raise ValueError("KY-E215: a value that is not finite")
    sketch -= means[:, None] * omega_sums[None, :]
    basis = np.zeros((n_genes, 0))
    products = np.zeros((n_genes, 0))
    block = extend_basis(basis, sketch)
    passes = 0
    done = block.shape[1] == 0
    for p in range(1, q + 1):
        if done:
            break
        w = block.shape[1]
        block_means = means @ block
        blk = np.ascontiguousarray(block)
        prod = np.zeros((n_genes, w))
        wsum = np.zeros(w)
        if tree:
            L.ko_pca_product_tree(*a, _p(blk, _f64p), w, _p(block_means, _f64p), n_genes, leaf_rows,
                                  _p(prod, _f64p), _p(wsum, _f64p))
        else:
            L.ko_pca_product(*a, _p(blk, _f64p), w, _p(block_means, _f64p), _p(prod, _f64p),
                             _p(wsum, _f64p))
        passes += 1
        prod -= means[:, None] * wsum[None, :]
        basis = np.hstack([basis, block])
        products = np.hstack([products, prod])
        if p == q:
            done = True
        else:
            block = extend_basis(basis, prod)
            done = block.shape[1] == 0
    eigenvalues, loadings = rayleigh_ritz(basis, products, n_components)
    mean_projection = means @ loadings
    k = n_components
    rows = np.zeros((n, k))
    squares = np.zeros(n_genes)
    stored = np.zeros(n_genes, np.uint64)
    lo = np.ascontiguousarray(loadings)
    if tree:
        L.ko_pca_rows_tree(*a, _p(lo, _f64p), k, _p(mean_projection, _f64p), _p(means, _f64p), n_genes, leaf_rows,
                           _p(rows, _f64p), _p(squares, _f64p), _p(stored, _u64p))
    else:
        L.ko_pca_rows(*a, _p(lo, _f64p), k, _p(mean_projection, _f64p), _p(means, _f64p),
                      _p(rows, _f64p), _p(squares, _f64p), _p(stored, _u64p))
    total = float(np.sum(squares + (float(n) - stored.astype(np.float64)) * means * means))
    total_variance = total / (n - 1.0)
    denom = n - 1.0
    constant = total_variance == 0.0
    variance = np.zeros(k) if constant else eigenvalues / denom
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = variance / total_variance
    return PcaResult(rows.astype(np.float32), loadings.astype(np.float32), variance, ratio,
                     means, eigenvalues, loadings, passes)


def pca_truth(indptr, indices, values, n_genes: int, n_components: int):
    """The FP64 truth: eigh of the dense centred covariance (genes x genes). Exact
    PCA, which kyto's result equals when the basis spans every gene
    (crates/kyto-exec/tests/it/pca.rs:284). Returns (variance, loadings, ratio)."""
    n = len(indptr) - 1
    x = np.zeros((n, n_genes))
    rows = np.repeat(np.arange(n), np.diff(np.asarray(indptr, np.int64)))
    x[rows, np.asarray(indices, np.int64)] = np.asarray(values, np.float64)
    mu = x.mean(axis=0)
    xc = x - mu
    cov = xc.T @ xc / (n - 1.0)
    vals, vecs = np.linalg.eigh(cov)
    order = np.argsort(vals)[::-1][:n_components]
    total = np.trace(cov)
    v = np.ascontiguousarray(vecs[:, order])
    flip_signs(v)
    return vals[order], v, vals[order] / total


def pca_truth_sparse(indptr, indices, values, n_genes: int, n_components: int, parts: int = 16):
    """The FP64 truth at scale: the centred covariance (X^T X - n mu mu^T) / (n - 1)
    from the CSR's Gram matrix (ko_gram, no densified rows), its eigh. Returns
    (variance, loadings genes x k with flip_signs' signs, ratio)."""
    indptr = np.ascontiguousarray(indptr, np.uint64)
    indices = np.ascontiguousarray(indices, np.uint32)
    values = np.ascontiguousarray(values, np.float32)
    n = len(indptr) - 1
    gram = np.zeros((n_genes, n_genes))
    sums = np.zeros(n_genes)
    lib().ko_gram(_p(indptr, _u64p), _p(indices, _u32p), _p(values, _f32p), n, n_genes, parts, _p(gram, _f64p),
                  _p(sums, _f64p))
    mu = sums / n
    cov = (gram - n * np.outer(mu, mu)) / (n - 1.0)
    vals, vecs = np.linalg.eigh(cov)
    order = np.argsort(vals)[::-1][:n_components]
    v = np.ascontiguousarray(vecs[:, order])
    flip_signs(v)
    return vals[order], v, vals[order] / np.trace(cov)


def subspace_angles(a, b):
    """Principal angles (radians, ascending) between the column spans of a and b
    (orthonormal columns)."""
    s = np.linalg.svd(np.asarray(a, np.float64).T @ np.asarray(b, np.float64), compute_uv=False)
    return np.sort(np.arccos(np.clip(s, -1.0, 1.0)))


# ------------------------------------------------------------------- neighbours --

def squared_distance(a, b) -> float:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return float(lib().ko_squared_distance(_p(a, _f32p), _p(b, _f32p), len(a)))


def nearest_offer_many(k, keys, rows):
    keys = np.ascontiguousarray(keys, np.float64)
    rows = np.ascontiguousarray(rows, np.uint32)
    ok = np.zeros(k)
    orow = np.zeros(k, np.uint32)
    n = lib().ko_nearest_offer_many(k, _p(keys, _f64p), _p(rows, _u32p), len(keys),
                                    _p(ok, _f64p), _p(orow, _u32p))
    return list(zip(ok[:n].tolist(), orow[:n].tolist()))


def search_rows(points, k, queries, metric="euclidean"):
    """`search_row` (neighbors.rs:521-536) for each query: keys (f64) and rows, ranked."""
    pts = np.ascontiguousarray(points, np.float32)
    n, w = pts.shape
    qs = np.ascontiguousarray(queries, np.uint32)
    keys = np.zeros((len(qs), k))
    rows = np.zeros((len(qs), k), np.uint32)
    norms = None
    if metric == "cosine":
        norms = np.zeros(n)
        lib().ko_norms(_p(pts, _f32p), n, w, _p(norms, _f64p))
    lib().ko_search_rows(_p(pts, _f32p), n, w, None if norms is None else _p(norms, _f64p), k,
                         _p(qs, _u32p), len(qs), _p(keys, _f64p), _p(rows, _u32p))
    return keys, rows


def knn_exact(points, k, metric="euclidean", threads=None):
    """neighbors_exact.rs:50-355: the f32-filter exact search (kyto's CPU path).
    Returns (keys f64 n x k, rows u32 n x k, fallback row count)."""
    L = lib()
    if threads is not None:
        L.ko_set_threads(int(threads))
    pts = np.ascontiguousarray(points, np.float32)
    n, w = pts.shape
    assert k <= 64
    keys = np.zeros((n, k))
    rows = np.zeros((n, k), np.uint32)
    fb = np.zeros(1, np.uint64)
    L.ko_knn_exact(_p(pts, _f32p), n, w, 1 if metric == "cosine" else 0, k, _p(keys, _f64p),
                   _p(rows, _u32p), _p(fb, _u64p))
    return keys, rows, int(fb[0])


def knn_exact_range(points, k, q_first, q_count, threads=None):
    """The filtered exact search for queries [q_first, q_first + q_count) only (the
    CPU baseline's bounded sample). Returns (keys, rows, fallback count)."""
    L = lib()
    if threads is not None:
        L.ko_set_threads(int(threads))
    pts = np.ascontiguousarray(points, np.float32)
    n, w = pts.shape
    keys = np.zeros((q_count, k))
    rows = np.zeros((q_count, k), np.uint32)
    fb = np.zeros(1, np.uint64)
    L.ko_knn_exact_range(_p(pts, _f32p), n, w, 0, k, q_first, q_count, _p(keys, _f64p), _p(rows, _u32p),
                         _p(fb, _u64p))
    return keys, rows, int(fb[0])


def knn_exact_rows(points, k, queries, metric="euclidean", threads=None):
    """The filtered exact search for the listed query rows only (sampled rows of a
    large embedding). Returns (keys f64 len x k, rows u32 len x k, fallback count)."""
    L = lib()
    if threads is not None:
        L.ko_set_threads(int(threads))
    pts = np.ascontiguousarray(points, np.float32)
    n, w = pts.shape
    qs = np.ascontiguousarray(queries, np.uint32)
    keys = np.zeros((len(qs), k))
    rows = np.zeros((len(qs), k), np.uint32)
    fb = np.zeros(1, np.uint64)
    L.ko_knn_exact_list(_p(pts, _f32p), n, w, 1 if metric == "cosine" else 0, k, _p(qs, _u32p), 0, len(qs),
                        _p(keys, _f64p), _p(rows, _u32p), _p(fb, _u64p))
    return keys, rows, int(fb[0])


def nearest_keys(queries, points, m, key="squared_euclidean", skip=None, query_norms=None, point_norms=None):
    """DeviceKernels::nearest_keys (crates/kyto-gpu/src/launch_matrix.rs:240-286) with
    its kernels (kernels_matrix.rs:140-257): squared_distances sums (q - p)^2 in
    component order in f32, dot_products sums q p likewise and cosine_keys makes
    1 - dot / (|q| |p|) in f32, clipped to [0, 2], 1 for a zero norm; `smallest` keeps
    each query's m smallest by (key, point), the point skip[q] left out. Each f32
    operation rounded on its own (no fused multiply-add). Returns (keys, points)."""
    q = np.ascontiguousarray(queries, np.float32)
    p = np.ascontiguousarray(points, np.float32)
    nq, n = q.shape[0], p.shape[0]
    keys = np.empty((nq, m), np.float32)
    rows = np.empty((nq, m), np.uint32)
    for i in range(nq):
        acc = np.zeros(n, np.float32)
        for c in range(q.shape[1]):
            if key == "squared_euclidean":
                t = np.float32(q[i, c]) - p[:, c]
                acc = acc + t * t
            else:
                acc = acc + np.float32(q[i, c]) * p[:, c]
        if key == "cosine":
            qn = np.float32(query_norms[i])
            pn = np.asarray(point_norms, np.float32)
            with np.errstate(divide="ignore", invalid="ignore"):
                k = np.float32(1.0) - acc / (qn * pn)
            k = np.clip(k, np.float32(0.0), np.float32(2.0))
            acc = np.where((qn == 0) | (pn == 0), np.float32(1.0), k).astype(np.float32)
        order = np.lexsort((np.arange(n), acc))
        if skip is not None and skip[i] != 0xFFFFFFFF:
            order = order[order != skip[i]]
        keys[i] = acc[order[:m]]
        rows[i] = order[:m]
    return keys, rows


def graph(keys, rows, metric="euclidean"):
    """`Space::graph` (neighbors.rs:412-435): per row, entries sorted by index,
    distance = f32(sqrt(key)) (Euclidean) or f32(key) (cosine)."""
    n, k = rows.shape
    order = np.argsort(rows, axis=1, kind="stable")
    idx = np.take_along_axis(rows, order, axis=1)
    kk = np.take_along_axis(keys, order, axis=1)
    dist = (np.sqrt(kk) if metric == "euclidean" else kk).astype(np.float32)
    indptr = np.arange(0, n * k + 1, k, dtype=np.uint64)
    return indptr, idx.astype(np.uint32), dist


def neighbors(points, n_neighbors=15, metric="euclidean", n_components=None, threads=None):
    """`neighbors::search` (neighbors.rs:226-312) with the checks it makes, then the
    exact CPU search and the CSR graph."""
    pts = np.asarray(points, np.float32)
    if pts.ndim != 2 or pts.shape[1] == 0:
        raise ValueError("KY-E215: embedding is not rows of components")
    n, width = pts.shape
    used = width if n_components is None else n_components
    if not 1 <= used <= width:
        raise ValueError("KY-E215: n_components")
    if n_neighbors < 2 or n_neighbors > n:
        raise ValueError("KY-E215: n_neighbors")
    pts = np.ascontiguousarray(pts[:, :used])
    if not np.all(np.isfinite(pts)):
        raise ValueError("KY-E215: embedding not finite")
    keys, rows, _ = knn_exact(pts, n_neighbors - 1, metric, threads)
    return graph(keys, rows, metric)


@dataclass
class Connectivities:
    """kyto's ConnectivityGraph (connectivities.rs:33-72) and each row's scales."""
    indptr: np.ndarray   # u64, n + 1
    indices: np.ndarray  # u32, columns increasing in each row
    weights: np.ndarray  # f32 in (0, 1]
    sigma: np.ndarray    # f32, n
    rho: np.ndarray      # f32, n


def smooth_knn(ranked, n_neighbors):
    """`smooth_knn_distance` (connectivities.rs:125-168) of one row of ranked distances
    (the cell's own 0 left out): (sigma, rho) as f32."""
    r = np.ascontiguousarray(ranked, np.float32)
    sigma = np.zeros(1, np.float32)
    rho = np.zeros(1, np.float32)
    lib().ko_smooth_knn(_p(r, _f32p), len(r), n_neighbors, float(np.log2(n_neighbors)), _p(sigma, _f32p),
                        _p(rho, _f32p))
    return sigma[0], rho[0]


def connectivities(indices, distances, n_neighbors) -> Connectivities:
    """scanpy's connectivities of a neighbour graph (connectivities.rs:76-123 and the
    fuzzy union :190-256): rows of n_neighbors - 1 entries, columns increasing."""
    idx = np.ascontiguousarray(indices, np.uint32).ravel()
    dist = np.ascontiguousarray(distances, np.float32).ravel()
    k = n_neighbors - 1
    n = len(idx) // k
    indptr = np.zeros(n + 1, np.uint64)
    out_idx = np.zeros(2 * n * k, np.uint32)
    out_w = np.zeros(2 * n * k, np.float32)
    sigma = np.zeros(n, np.float32)
    rho = np.zeros(n, np.float32)
    nnz = lib().ko_connectivities(n, n_neighbors, _p(idx, _u32p), _p(dist, _f32p), _p(indptr, _u64p),
                                  _p(out_idx, _u32p), _p(out_w, _f32p), _p(sigma, _f32p), _p(rho, _f32p))
    return Connectivities(indptr, out_idx[:nnz].copy(), out_w[:nnz].copy(), sigma, rho)
