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
