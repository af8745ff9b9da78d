"""Synthetic single-cell layers of the benchmark workloads (csrc/synth.c).

Negative-binomial counts (dispersion 0.3), gene means LogNormal(-1, 1.5), cell size
factors LogNormal(0, 0.5), T cell types each with N(0, 1) log-fold changes on 5% of
the genes; stored as CSR counts, then normalize_total(1e4) and log1p (the PCA input).
The gene means carry one scale factor, found by bisection on the model's expected
density, so the layer reaches the asked density; the measured density is returned.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class Workload:
    name: str
    n_cells: int
    n_genes: int
    n_types: int
    density: float
    seed: int = 0


WORKLOADS = {
    "pbmc3k_shaped": Workload("pbmc3k_shaped", 2_700, 2_000, 8, 0.10),
    "100k": Workload("100k", 100_000, 2_000, 12, 0.08),
    "1m": Workload("1m", 1_000_000, 2_000, 20, 0.08),
    "4m": Workload("4m", 4_000_000, 3_000, 40, 0.07),
}


def _spec(n_cells, n_genes, n_types, seed, scale):
    return N.SynthSpec(n_cells, n_genes, n_types, seed, -1.0, 1.5, 0.5, 0.3, 0.05, scale)


def calibrate(n_genes: int, n_types: int, density: float, seed: int = 0, sample_cells: int = 4096) -> float:
    """The gene-mean scale whose expected density (over `sample_cells` cells) is `density`."""
    lib = N.synth_lib()
    lo, hi = 1e-4, 1e3
    for _ in range(60):
        mid = (lo * hi) ** 0.5
        d = lib.ks_expected_density(ctypes_ref(_spec(sample_cells, n_genes, n_types, seed, mid)), sample_cells)
        if d < density:
            lo = mid
        else:
            hi = mid
    return (lo * hi) ** 0.5


def ctypes_ref(spec):
    import ctypes
    return ctypes.byref(spec)


@dataclass
class Layer:
    indptr: np.ndarray   # u64
    indices: np.ndarray  # u32
    values: np.ndarray   # f32 (counts, or log1p after normalize_log1p)
    n_genes: int
    mean_scale: float
    cell_types: np.ndarray | None = None

    @property
    def n_cells(self) -> int:
        return len(self.indptr) - 1

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    @property
    def density(self) -> float:
        return self.nnz / (self.n_cells * self.n_genes)


def counts(n_cells: int, n_genes: int, n_types: int, density: float, seed: int = 0,
           mean_scale: float | None = None, with_types: bool = False) -> Layer:
    lib = N.synth_lib()
    scale = calibrate(n_genes, n_types, density, seed) if mean_scale is None else mean_scale
    spec = _spec(n_cells, n_genes, n_types, seed, scale)
    indptr = np.zeros(n_cells + 1, np.uint64)
    nnz = lib.ks_count(ctypes_ref(spec), N.ptr(indptr, N.u64p))
    indices = np.empty(max(nnz, 1), np.uint32)
    values = np.empty(max(nnz, 1), np.float32)
    lib.ks_fill(ctypes_ref(spec), N.ptr(indptr, N.u64p), N.ptr(indices, N.u32p), N.ptr(values, N.f32p))
    types = None
    if with_types:
        types = np.empty(n_cells, np.uint32)
        lib.ks_cell_types(ctypes_ref(spec), N.ptr(types, N.u32p))
    return Layer(indptr, indices[:nnz], values[:nnz], n_genes, scale, types)


def normalize_log1p(layer: Layer, target_sum: float = 1e4) -> Layer:
    """kyto's normalize_total(target_sum) then log1p, in place (returns the layer)."""
    N.synth_lib().ks_normalize_log1p(N.ptr(layer.indptr, N.u64p), layer.n_cells, float(target_sum),
                                     N.ptr(layer.values, N.f32p))
    return layer


def log1p_layer(n_cells: int, n_genes: int, n_types: int, density: float, seed: int = 0,
                with_types: bool = False) -> Layer:
    return normalize_log1p(counts(n_cells, n_genes, n_types, density, seed, with_types=with_types))


def workload(name: str, with_types: bool = False) -> Layer:
    w = WORKLOADS[name]
    return log1p_layer(w.n_cells, w.n_genes, w.n_types, w.density, w.seed, with_types)


def block_reader(n_cells: int, n_genes: int, n_types: int, density: float, seed: int = 0):
    """A reader of the synthetic counts matrix by row blocks (for Device.stream_counts):
    the same matrix as counts(), generated block by block into the caller's buffers."""
    import ctypes

    lib = N.synth_lib()
    spec = _spec(n_cells, n_genes, n_types, seed, calibrate(n_genes, n_types, density, seed))

    def read(first, count, indptr, indices, values, capacity):
        return lib.ks_block(ctypes_ref(spec), first, count, ctypes.cast(indptr, ctypes.c_void_p),
                            ctypes.cast(indices, ctypes.c_void_p), ctypes.cast(values, ctypes.c_void_p), capacity)

    return read
