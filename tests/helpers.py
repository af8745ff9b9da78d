"""Shared test data: the golden fixtures' PCA input and seeded embeddings."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def synthetic_hvg_csr():
    """The 500 x 400 synthetic fixture's log1p CSR restricted to its 100 seurat HVGs
    (the input of scanpy's PCA in synthetic_pca.npz)."""
    g = load("synthetic_log1p.npz")
    indptr, indices, data, hvg = g["indptr"], g["indices"], g["data"], g["hvg"]
    pos = np.cumsum(hvg) - 1
    keep = hvg[indices]
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))[keep]
    ip = np.zeros(len(indptr), np.uint64)
    np.cumsum(np.bincount(rows, minlength=len(indptr) - 1), out=ip[1:])
    return ip, pos[indices[keep]].astype(np.uint32), data[keep].astype(np.float32), int(hvg.sum())


def scanpy_rows(name, n):
    """scanpy's exact kNN graph rows (indices sorted within a row) and distances."""
    nb = load(f"{name}_neighbors.npz")
    idx = nb["indices"].reshape(n, -1)
    dist = nb["distances"].reshape(n, -1)
    o = np.argsort(idx, axis=1)
    return np.take_along_axis(idx, o, 1), np.take_along_axis(dist, o, 1)


class XorShift:
    """The seeded stream of neighbors_exact.rs tests (xorshift, values in [0, 1))."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = s
        return np.float32((s >> 40) / float(1 << 24))


def clustered(seed, n, width, clusters, offset):
    """Rows around `clusters` centres (neighbors_exact.rs tests `clustered`)."""
    st = XorShift(seed)
    centres = np.array([10.0 * st.next() - 5.0 for _ in range(clusters * width)], np.float32)
    out = np.empty((n, width), np.float32)
    for r in range(n):
        c = r % clusters
        for j in range(width):
            noise = np.float32(st.next() + st.next() + st.next() - np.float32(1.5))
            out[r, j] = np.float32(centres[c * width + j] + noise + np.float32(offset))
    return out


def scanpy_graph(name):
    """scanpy's neighbour graph as umap receives it (indices u32 and its float64
    distances as float32, n x 14) and scanpy's connectivities (indptr, indices,
    weights)."""
    nb = load(f"{name}_neighbors.npz")
    n = len(nb["indptr"]) - 1
    idx = nb["indices"].astype(np.uint32).reshape(n, -1)
    dist = nb["distances"].astype(np.float32).reshape(n, -1)
    return idx, dist, (nb["conn_indptr"].astype(np.uint64), nb["conn_indices"].astype(np.uint32),
                       nb["conn_weights"])


# The neighbour graphs of connectivities.rs's unit tests (:266-397): rows of
# (column, distance), n_neighbors = row length + 1.
KNOWN_GRAPHS = {
    "line": [[(1, 1.0), (2, 3.0)], [(0, 1.0), (2, 2.0)], [(0, 3.0), (1, 2.0)]],
    "duplicates": [[(1, 0.0), (2, 1.0)], [(0, 0.0), (2, 1.0)], [(1, 1.0), (2, 0.0)]],
    "same": [[(1, 0.0), (2, 0.0)], [(0, 0.0), (2, 0.0)], [(0, 0.0), (1, 0.0)]],
    "far": [[(1, 1.0), (2, 1.0), (3, 1e6)], [(0, 1.0), (2, 1.0), (3, 1.0)], [(0, 1.0), (1, 1.0), (3, 1.0)],
            [(0, 1e6), (1, 1.0), (2, 1.0)]],
}


def known_graph(name):
    rows = KNOWN_GRAPHS[name]
    idx = np.array([[c for c, _ in r] for r in rows], np.uint32)
    dist = np.array([[d for _, d in r] for r in rows], np.float32)
    return idx, dist, idx.shape[1] + 1


def random_graph(n, k, seed, ties=False):
    """A random neighbour graph (columns increasing, no self entries), distances drawn
    so rows hold ties and zeros when `ties`."""
    rng = np.random.default_rng(seed)
    idx = np.empty((n, k), np.uint32)
    for r in range(n):
        c = rng.choice(n - 1, size=k, replace=False)
        c[c >= r] += 1
        idx[r] = np.sort(c)
    dist = rng.gamma(2.0, 1.0, size=(n, k)).astype(np.float32)
    if ties:
        dist = np.round(dist * 4) / 4
        dist[rng.random((n, k)) < 0.05] = 0.0
        dist[rng.random(n) < 0.01] = 0.0  # rows of zero distances
    return idx, dist.astype(np.float32)


def dense_union(idx, w, n):
    """w + w^T - w o w^T on a dense float32 matrix, as scipy orders it."""
    W = np.zeros((n, n), np.float32)
    rows = np.repeat(np.arange(n), idx.shape[1])
    W[rows, idx.ravel()] = w.ravel()
    return (W + W.T) - W * W.T


class Lcg:
    """The deterministic generator of kyto's PCA tests (crates/kyto-exec/tests/it/
    pca.rs:32-45): uniform in [0, 1) from the top 53 bits of a 64-bit LCG."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return (self.s >> 11) / float(1 << 53)


def lcg_rows(n_cells, n_genes, rank, noise, seed):
    """kyto's `data` (pca.rs tests :46-86): scores in [-1, 1) (cells x rank), then
    loadings (rank x genes) in [-1, 1) times 4 / (1 + component), then each value the
    f64 sum of score x loading over the components plus noise x U[-1, 1), clamped
    at 0 and rounded to f32. Draws in that order."""
    r = Lcg(seed)
    scores = [[r.next() * 2.0 - 1.0 for _ in range(rank)] for _ in range(n_cells)]
    loadings = [[(r.next() * 2.0 - 1.0) * (4.0 / (1.0 + c)) for _ in range(n_genes)] for c in range(rank)]
    out = np.empty((n_cells, n_genes), np.float32)
    for i in range(n_cells):
        for g in range(n_genes):
            v = 0.0
            for c in range(rank):
                v += scores[i][c] * loadings[c][g]
            v += noise * (r.next() * 2.0 - 1.0)
            out[i, g] = np.float32(max(v, 0.0))
    return out


def dense_to_csr(x):
    """A dense f32 matrix as CSR (u64 indptr, u32 indices, f32 values), zeros not stored."""
    x = np.asarray(x, np.float32)
    rows, cols = np.nonzero(x)
    indptr = np.zeros(x.shape[0] + 1, np.uint64)
    np.cumsum(np.bincount(rows, minlength=x.shape[0]), out=indptr[1:])
    return indptr, cols.astype(np.uint32), x[rows, cols].astype(np.float32)
