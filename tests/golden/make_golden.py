"""Writes the golden fixtures of tests/golden/ from the reference's own data files.

Reads only `.npz` data under the reference checkout (tests/fixtures/reference/,
made by the reference's tests/fixtures/make_reference.py with scanpy 1.12.5; their
SHA-256 digests are in its manifest.json) with numpy.load(allow_pickle=False), and
keeps the arrays the pca() + neighbors() parity tests need:

- synthetic_log1p.npz: the 500 x 400 synthetic fixture's `target` normalize_total
  (1e4) CSR and its log1p CSR (normalize_log1p.npz), and the seurat HVG mask (hvg.npz)
  selecting the 100 genes scanpy's PCA ran on;
- synthetic_pca.npz / pbmc3k_pca.npz: scanpy's arpack PCA (embedding, components,
  variance, variance_ratio) on those HVGs (pca.npz);
- synthetic_neighbors.npz / pbmc3k_neighbors.npz: scanpy's exact kNN distance graph
  (n_neighbors=15, euclidean) on the PCA embedding (neighbors.npz), as
  int32 indices / float64 distances, and its connectivities graph (UMAP's fuzzy
  weights on that graph) as conn_indptr / conn_indices (int32) / conn_weights
  (float32).

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

SRC = "/root/reference/tests/fixtures/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def load(path):
    return np.load(os.path.join(SRC, path), allow_pickle=False)


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} is not here: the fixtures are already committed")
    nl = load("synthetic/normalize_log1p.npz")
    hvg = load("synthetic/hvg.npz")
    np.savez_compressed(
        os.path.join(OUT, "synthetic_log1p.npz"),
        indptr=nl["target/log1p/indptr"].astype(np.int64),
        indices=nl["target/log1p/indices"].astype(np.int32),
        data=nl["target/log1p/data"].astype(np.float32),
        normalized=nl["target/normalized/data"].astype(np.float32),
        normalized_indices=nl["target/normalized/indices"].astype(np.int32),
        shape=nl["target/log1p/shape"].astype(np.int64),
        cell_sum=nl["target/log1p/cell_sum"],
        gene_sum=nl["target/log1p/gene_sum"],
        hvg=hvg["seurat/highly_variable"].astype(bool),
    )
    for name in ("synthetic", "pbmc3k"):
        p = load(f"{name}/pca.npz")
        np.savez_compressed(
            os.path.join(OUT, f"{name}_pca.npz"),
            embedding=p["embedding"].astype(np.float32),
            components=p["components"].astype(np.float32),
            variance=p["variance"],
            variance_ratio=p["variance_ratio"],
        )
        nb = load(f"{name}/neighbors.npz")
        np.savez_compressed(
            os.path.join(OUT, f"{name}_neighbors.npz"),
            indptr=nb["distances/indptr"].astype(np.int64),
            indices=nb["distances/indices"].astype(np.int32),
            distances=nb["distances/data"].astype(np.float64),
            conn_indptr=nb["connectivities/indptr"].astype(np.int64),
            conn_indices=nb["connectivities/indices"].astype(np.int32),
            conn_weights=nb["connectivities/data"].astype(np.float32),
        )


if __name__ == "__main__":
    main()
