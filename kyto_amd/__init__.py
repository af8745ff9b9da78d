"""kyto_amd: the MI355X embedding path of kyto — pca() then neighbors() — behind
kyto's plan API (crates/kyto-py/src/atlas.rs:822-966, 1622-1720).

    import kyto_amd as ky
    out = (ky.scan(counts)                      # a CSR layer (cells x genes)
             .normalize_total(1e4).log1p()      # atlas.rs:578-632, made on the device
             .pca(n_components=50, seed=0)      # atlas.rs:822-856
             .neighbors(n_neighbors=15)         # atlas.rs:857-966
             .on(device="gpu")                  # atlas.rs:1622-1636
             .collect(memory_limit="32GB"))     # atlas.rs:1691-1720
    out["X_pca"], out["distances"], out["connectivities"]

A kyto store on disk works the same way: scan("pbmc.kyto") reads its schema, pca()
streams the layer's tiles through the pinned buffers (kyto_amd.store), and
sink_zarr(path) writes the outputs as a new store (X_pca, X_pca_loadings, the
distances and connectivities graphs, the layers).

Chained methods only build a plan; collect() runs it on the device. There is no CPU
path: on(device="cpu") is refused with KY-E501 and a missing native library raises.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field, replace
from typing import Any

import numpy as np

from ._native import KytoError, KY_E215, KY_E401, KY_E501
from . import store as _store
from .device import ConnectivityGraph, CsrLayer, DenseMatrix, Device, Graph, IvfIndex, PcaOutput, device_count

__all__ = ["scan", "Atlas", "Device", "KytoError", "Graph", "ConnectivityGraph", "PcaOutput", "device_count", "CsrLayer", "DenseMatrix",
           "IvfIndex"]

# crates/kyto-plan/src/physical_plan.rs:73
PCA_POWER_ITERATIONS = 9


def _piece(nbytes: int) -> int:
    """A device buffer's bytes as the library rounds them (256-byte pieces)."""
    return (int(nbytes) + 255) // 256 * 256


def _invalid(function, parameter, value, problem, accepted) -> KytoError:
    return KytoError(KY_E215, f"{function}: {parameter}={value!r}: {problem} (accepted: {accepted})")


def _integer(function, parameter, value, minimum):
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
        raise _invalid(function, parameter, value, "not an int", f"an int of at least {minimum}")
    if value < minimum:
        raise _invalid(function, parameter, value, f"below {minimum}", f"an int of at least {minimum}")
    return int(value)


_UNITS = {"b": 1, "kb": 1e3, "mb": 1e6, "gb": 1e9, "tb": 1e12, "kib": 2**10, "mib": 2**20, "gib": 2**30,
          "tib": 2**40}


def parse_memory(limit) -> int | None:
    """`memory_limit=` as kyto's terminals take it: bytes, or '32GB' / '512MiB'."""
    if limit is None:
        return None
    if isinstance(limit, (int, np.integer)) and not isinstance(limit, bool):
        if limit <= 0:
            raise _invalid("collect", "memory_limit", limit, "not positive", "a positive size")
        return int(limit)
    m = re.fullmatch(r"\s*([0-9]*\.?[0-9]+)\s*([a-zA-Z]*)\s*", str(limit))
    if not m or m.group(2).lower() not in _UNITS or float(m.group(1)) <= 0:
        raise _invalid("collect", "memory_limit", limit, "not a size", "an int of bytes or a string like '32GB'")
    return int(float(m.group(1)) * _UNITS[m.group(2).lower()])


@dataclass(frozen=True)
class _Layer:
    indptr: np.ndarray
    indices: np.ndarray
    values: np.ndarray
    n_genes: int

    @property
    def n_cells(self):
        return len(self.indptr) - 1

    @property
    def nnz(self):
        return int(self.indptr[-1])


@dataclass(frozen=True)
class _StoreRef:
    """A layer of a kyto store on disk, read when a step needs it."""
    store: str
    name: str
    n_cells: int
    n_genes: int
    nnz: int
    dtype: str
    cell_block: int
    gene_block: int

    def host(self) -> _Layer:
        with _store.StoreLayer(self.store, self.name) as layer:
            ip, ix, v = layer.read_rows()
        return _Layer(ip, ix, v, self.n_genes)


def _as_layer(x) -> _Layer:
    if hasattr(x, "indptr") and hasattr(x, "indices"):
        values = getattr(x, "values", None)
        if values is None:
            values = getattr(x, "data")
        n_genes = getattr(x, "n_genes", None)
        if n_genes is None:
            n_genes = x.shape[1]
        return _Layer(np.ascontiguousarray(x.indptr, np.uint64), np.ascontiguousarray(x.indices, np.uint32),
                      np.ascontiguousarray(values, np.float32), int(n_genes))
    if isinstance(x, tuple) and len(x) == 4:
        ip, ix, v, g = x
        return _Layer(np.ascontiguousarray(ip, np.uint64), np.ascontiguousarray(ix, np.uint32),
                      np.ascontiguousarray(v, np.float32), int(g))
    raise _invalid("scan", "layer", type(x).__name__, "not a CSR layer",
                   "an object with indptr/indices/data (or values) and shape/n_genes, or (indptr, indices, values, n_genes)")


@dataclass(frozen=True)
class _Map:
    """normalize_total(target_sum) or log1p over a layer (atlas.rs:578-630)."""
    kind: str          # "normalize_total" or "log1p"
    layer: str
    into: str
    target_sum: float | None = None


@dataclass(frozen=True)
class _Pca:
    layer: str | None
    into: str
    n_components: int
    seed: int
    genes: Any = None
    power_iterations: int = PCA_POWER_ITERATIONS


@dataclass(frozen=True)
class _Neighbors:
    embedding: str | None
    n_neighbors: int
    n_components: int | None
    metric: str
    search: str
    seed: int
    into: str | None
    precision: str


@dataclass(frozen=True)
class Atlas:
    """A lazy plan over CSR layers (kyto's Atlas for the embedding path)."""

    layers: dict
    default_layer: str
    steps: tuple = ()
    device: Any = "auto"

    # ----------------------------------------------------------------- steps --

    def _latest_layer(self) -> str:
        """The layer the latest normalize_total / log1p of the plan wrote, else the
        default layer (atlas.rs:610-620)."""
        maps = [s for s in self.steps if isinstance(s, _Map)]
        return maps[-1].into if maps else self.default_layer

    def _layer_names(self) -> set:
        return set(self.layers) | {s.into for s in self.steps if isinstance(s, _Map)}

    def _add_map(self, kind, layer, into, default_into, target_sum=None) -> "Atlas":
        name = layer if layer is not None else self._latest_layer()
        if name not in self._layer_names():
            raise KytoError(109, f"{kind}: layer {name!r} is not a layer of the plan")
        if into is None:
            if default_into in self._layer_names():
                raise KytoError(220, f"{kind}: the plan already has the layer {default_into!r}; pass into= to name "
                                     f"the output")
            into = default_into
        return replace(self, steps=self.steps + (_Map(kind, name, into, target_sum),))

    def normalize_total(self, target_sum=None, *, layer=None, into=None) -> "Atlas":
        """atlas.rs:578-591: each cell's values divided by its total over target_sum
        (kyto's f32 steps, on the device). The median target (target_sum=None) needs a
        statistics pass over every cell that this path does not run: KY-E501."""
        if target_sum is None:
            raise KytoError(KY_E501, "normalize_total: the median target (target_sum=None) is not on the device "
                                     "path; pass target_sum=")
        if isinstance(target_sum, bool) or not isinstance(target_sum, (int, float, np.integer, np.floating)):
            raise _invalid("normalize_total", "target_sum", target_sum, "not a number", "a float above 0")
        target_sum = float(target_sum)
        if not (np.isfinite(target_sum) and target_sum > 0):
            raise _invalid("normalize_total", "target_sum", target_sum, "not a finite number above 0",
                           "a float above 0")
        return self._add_map("normalize_total", layer, into, "normalized", target_sum)

    def log1p(self, *, layer=None, into=None) -> "Atlas":
        """atlas.rs:628-632: log(1 + x) of every value (kyto's log1p, on the device)."""
        return self._add_map("log1p", layer, into, "log1p")

    def pca(self, *, layer=None, into=None, n_components=None, seed=None, genes=None,
            power_iterations=None) -> "Atlas":
        """atlas.rs:822-856: n_components defaults to 50, seed is required; layer None
        reads the layer the latest normalize_total / log1p wrote, else the default.
        power_iterations (not in kyto's signature, whose plan fixes it at 9,
        physical_plan.rs:73) sets the block-Krylov passes for a workload that names
        another count."""
        if seed is None:
            raise _invalid("pca", "seed", None, "a seed is required (R7)", "an int of at least 0")
        seed = _integer("pca", "seed", seed, 0)
        q = PCA_POWER_ITERATIONS if power_iterations is None else _integer("pca", "power_iterations",
                                                                          power_iterations, 1)
        n_components = 50 if n_components is None else _integer("pca", "n_components", n_components, 1)
        name = layer if layer is not None else self._latest_layer()
        if name not in self._layer_names():
            raise KytoError(109, f"pca: layer {name!r} is not a layer of the plan")
        if genes is not None:
            genes = np.asarray(genes, bool)
            if genes.shape != (self.layers[self._source(name)[0]].n_genes,):
                raise _invalid("pca", "genes", "mask", "not one flag per gene", "a boolean mask over the genes")
            if not genes.any():
                raise _invalid("pca", "genes", "mask", "the predicate is true for none of the genes reaching pca",
                               f"a predicate true for at least {n_components} genes (n_components)")
        return replace(self, steps=self.steps + (_Pca(name, into or "X_pca", n_components, seed, genes, q),))

    def neighbors(self, embedding=None, /, *, n_neighbors=None, n_components=None, metric="euclidean",
                  search="exact", seed=None, into=None, precision="bf16") -> "Atlas":
        """atlas.rs:857-966: n_neighbors (default 15) counts the cell itself; search
        'exact' is exact (candidates at `precision`, re-ranked and certified exactly),
        'approximate' is the learned IVF router."""
        n_neighbors = 15 if n_neighbors is None else _integer("neighbors", "n_neighbors", n_neighbors, 0)
        if n_neighbors < 2:
            raise _invalid("neighbors", "n_neighbors", n_neighbors, "it counts the cell itself",
                           "an int of at least 2")
        if n_components is not None:
            n_components = _integer("neighbors", "n_components", n_components, 1)
        seed = 0 if seed is None else _integer("neighbors", "seed", seed, 0)
        if metric not in ("euclidean", "cosine"):
            raise _invalid("neighbors", "metric", metric, "unknown", "'euclidean' or 'cosine'")
        if search not in ("exact", "approximate"):
            raise _invalid("neighbors", "search", search, "unknown", "'exact' or 'approximate'")
        if precision not in ("fp32", "bf16"):
            raise _invalid("neighbors", "precision", precision, "unknown", "'fp32' or 'bf16'")
        if embedding is None:
            pcas = [s for s in self.steps if isinstance(s, _Pca)]
            embedding = pcas[-1].into if pcas else "X_pca"
        if not any(isinstance(s, _Pca) and s.into == embedding for s in self.steps):
            raise KytoError(109, f"neighbors: embedding {embedding!r} is not a cell embedding of the plan")
        return replace(self, steps=self.steps + (
            _Neighbors(embedding, n_neighbors, n_components, metric, search, seed, into, precision),))

    def _source(self, name):
        """The scanned layer a layer is made from and the maps in order."""
        maps = []
        while name not in self.layers:
            step = next(s for s in reversed(self.steps) if isinstance(s, _Map) and s.into == name)
            maps.append(step)
            name = step.layer
        return name, maps[::-1]

    def on(self, device="auto") -> "Atlas":
        """atlas.rs:1622-1636: 'auto', 'gpu', 'gpu:N', an int N, or 'cpu'."""
        if isinstance(device, (int, np.integer)) and not isinstance(device, bool):
            ordinal = int(device)
        elif isinstance(device, str) and device in ("auto", "gpu", "cuda", "cpu"):
            ordinal = device
        elif isinstance(device, str) and re.fullmatch(r"(gpu|cuda):\d+", device):
            ordinal = int(device.split(":")[1])
        else:
            raise _invalid("on", "device", device, "unknown device", "'auto', 'gpu', 'gpu:N', 'cpu' or an int")
        return replace(self, device=ordinal)

    def explain(self) -> str:
        lines = []
        for s in reversed(self.steps):
            if isinstance(s, _Map):
                extra = f" target_sum={s.target_sum:g}" if s.kind == "normalize_total" else ""
                lines.append(f'{s.kind.upper()} "{s.layer}" -> "{s.into}"{extra} [device: {self.device}]')
            elif isinstance(s, _Pca):
                lines.append(f'PCA "{s.layer}" -> "{s.into}" n_components={s.n_components} seed={s.seed} '
                             f"[device: {self.device}]")
            else:
                dist, conn = _graph_names(s.into)
                lines.append(f'NEIGHBORS "{s.embedding}" -> "{dist}", "{conn}" n_neighbors={s.n_neighbors} '
                             f"metric={s.metric} search={s.search} precision={s.precision} [device: {self.device}]")
        lines.append(f"SCAN ({', '.join(self.layers)})")
        return "\n".join(lines)

    # -------------------------------------------------------------- terminal --

    def collect(self, *, memory_limit=None, threads=None, progress=None) -> dict:
        """Run the plan on the device; returns a dict of the outputs (embeddings,
        loadings, variances, graphs)."""
        return self._run(memory_limit, threads, progress, keep_layers=False)

    def sink_zarr(self, path, /, *, overwrite=False, memory_limit=None, threads=None, progress=None,
                  cell_block=None, gene_block=None, zstd_level=_store.ZSTD_LEVEL) -> None:
        """Run the plan and write its output as a new kyto store at `path` (the
        reference's Atlas.sink_zarr, python/kyto/_kyto.pyi:1323; format.md §10): the
        scanned layers (copied from a scanned store with its tables), the layer a pca
        read when the maps made it (the last map's), each pca's embedding and
        `<into>_loadings`, each neighbors step's distances and connectivities graphs.
        The store is written next to `path` and renamed into place; KY-E118 if `path`
        exists unless overwrite=True (and then only a kyto store is replaced)."""
        out = self._run(memory_limit, threads, progress, keep_layers=True)
        made = out.pop("__made_layers__")
        first = next(iter(self.layers.values()))
        n_cells, n_genes = first.n_cells, first.n_genes
        from_store = isinstance(first, _StoreRef)
        cb = cell_block or (first.cell_block if from_store else 4096)
        gb = gene_block or (first.gene_block if from_store else 1024)
        layers = {}
        if not from_store:
            for name, layer in self.layers.items():
                layers[name] = _store.LayerData(layer.indptr, layer.indices, layer.values)
        for name, (source, values) in made.items():
            src = self.layers[source]
            host = src.host() if isinstance(src, _StoreRef) else src
            ip, ix = host.indptr, host.indices
            keep = values != 0  # a map can make a zero (format.md §4.6: never stored)
            if not keep.all():
                rows = np.repeat(np.arange(n_cells), np.diff(np.asarray(ip, np.int64)))[keep]
                ip = np.zeros(n_cells + 1, np.uint64)
                np.cumsum(np.bincount(rows, minlength=n_cells), out=ip[1:])
                ix, values = ix[keep], values[keep]
            layers[name] = _store.LayerData(ip, ix, values)
        embeddings, graphs = {}, {}
        for step in self.steps:
            if isinstance(step, _Pca):
                embeddings[step.into] = ("cells", out[step.into])
                embeddings[f"{step.into}_loadings"] = ("genes", out[f"{step.into}_loadings"])
            elif isinstance(step, _Neighbors):
                key, conn = _graph_names(step.into)
                g, c = out[key], out[conn]
                graphs[key] = (g.indptr, g.indices, g.distances)
                graphs[conn] = (c.indptr, c.indices, c.weights)
        _store.create_store(path, n_cells=n_cells, n_genes=n_genes, cell_block=cb, gene_block=gb, layers=layers,
                            embeddings=embeddings, graphs=graphs,
                            tables_from=first.store if from_store else None,
                            copy_layers=list(self.layers) if from_store else None, overwrite=overwrite,
                            zstd_level=zstd_level)

    def _memory_plan(self, limit) -> dict:
        """kyto's within_memory_limit for this path (memory_limit.rs:119-175): each
        stage's device memory (the resident embeddings of earlier pca steps, plus a
        pca's layer and passes, or a neighbors step's search and connectivities) from
        the library's models; a neighbors search takes the largest round of queries
        that fits. KY-E401, before anything runs, with the smallest peak the plan can
        run in, when a stage does not fit. Returns each neighbors step's round."""
        from .device import connectivities_bytes, neighbors_bytes, neighbors_plan, pca_bytes

        rounds: dict = {}
        if limit is None:
            return rounds
        resident, required = 0, 0
        shapes: dict = {}
        for i, step in enumerate(self.steps):
            if isinstance(step, _Pca):
                source, _ = self._source(step.layer)
                layer = self.layers[source]
                n, g, nnz = layer.n_cells, layer.n_genes, layer.nnz
                required = max(required, resident + pca_bytes(n, g, nnz, step.n_components, step.power_iterations))
                resident += _piece(4 * n * step.n_components)
                shapes[step.into] = (n, step.n_components)
            elif isinstance(step, _Neighbors):
                n, cols = shapes[step.embedding]
                precision = step.precision
                conn = connectivities_bytes(n, step.n_neighbors)
                smallest = neighbors_bytes(n, cols, step.n_neighbors, step.n_components, step.metric, precision,
                                           256 if precision == "bf16" else 64)
                need = resident + max(smallest, conn)
                required = max(required, need)
                if need <= limit:
                    rounds[i], _ = neighbors_plan(n, cols, limit - resident, step.n_neighbors, step.n_components,
                                                  step.metric, precision)
        if required > limit:
            raise KytoError(KY_E401, f"the plan needs at least {required} bytes but memory_limit is {limit} bytes")
        return rounds

    def _total_rows(self) -> int:
        """Cell rows the run computes, summed over its passes (kyto's Progress.total):
        a pca reads its layer once and makes 2 + q passes, a neighbors step one."""
        total, cells = 0, {}
        for step in self.steps:
            if isinstance(step, _Pca):
                n = self.layers[self._source(step.layer)[0]].n_cells
                total += n * (step.power_iterations + 3)
                cells[step.into] = n
            elif isinstance(step, _Neighbors):
                total += cells.get(step.embedding, 0)
        return total

    def _run(self, memory_limit, threads, progress, keep_layers) -> dict:
        """Run the plan on the device: on a worker thread, while the calling thread
        reports progress(done, total) as the run moves and turns Ctrl-C into a cancel
        (the run stops at its next pass, block or round with KY-E503)."""
        import threading

        from .device import set_threads

        limit = parse_memory(memory_limit)
        if threads is not None:
            threads = _integer("collect", "threads", threads, 1)
        if progress is not None and not callable(progress):
            raise _invalid("collect", "progress", type(progress).__name__, "not callable",
                           "a callable taking (done, total)")
        if self.device == "cpu":
            raise KytoError(KY_E501, "this build runs pca and neighbors on the GPU only: on(device='gpu')")
        rounds = self._memory_plan(limit)
        ordinal = 0 if self.device in ("auto", "gpu", "cuda") else int(self.device)
        dev = Device(ordinal)
        outcome: dict = {}

        def work():
            try:
                outcome["out"] = self._steps(dev, rounds, limit is not None, keep_layers)
            except BaseException as e:  # noqa: BLE001 - handed to the calling thread
                outcome["error"] = e

        try:
            set_threads(threads or 0)
            dev.run_begin(self._total_rows())
            worker = threading.Thread(target=work, name="kyto_amd-run", daemon=True)
            worker.start()
            seen = None
            try:
                while worker.is_alive():
                    worker.join(0.01)
                    if progress is not None:
                        now = dev.progress()
                        if now != seen:
                            seen = now
                            progress(*now)
            except KeyboardInterrupt as interrupt:
                # Ctrl-C (or a progress callback raising it) cancels the run cleanly: it
                # stops at its next pass, block or round, its buffers freed.
                dev.cancel()
                worker.join()
                error = outcome.get("error")
                if isinstance(error, KytoError) and error.status == 503:
                    raise error from interrupt
                raise
            if "error" in outcome:
                raise outcome["error"]
            if progress is not None:
                total = dev.progress()[1]
                if seen != (total, total):
                    progress(total, total)
            return outcome["out"]
        finally:
            set_threads(0)
            dev.close()

    def _steps(self, dev: Device, rounds: dict, limited: bool, keep_layers: bool) -> dict:
        out: dict = {}
        made: dict = {}
        resident: dict[str, DenseMatrix] = {}
        try:
            for i, step in enumerate(self.steps):
                if isinstance(step, _Map):
                    continue  # applied on the device by the step that reads the layer
                if isinstance(step, _Pca):
                    source, maps = self._source(step.layer)
                    layer = self.layers[source]
                    res, emb = _run_pca(dev, layer, step, _device_maps(maps), keep_layer=keep_layers)
                    if keep_layers and res.made_values is not None:
                        made[step.layer] = (source, res.made_values)
                    resident[step.into] = emb
                    out[step.into] = res.embedding
                    out[f"{step.into}_loadings"] = res.loadings
                    out[f"{step.into}_variance"] = res.variance
                    out[f"{step.into}_variance_ratio"] = res.variance_ratio
                    out[f"{step.into}_means"] = res.means
                else:
                    emb = resident[step.embedding]
                    g = _run_neighbors(dev, emb, step, rounds.get(i, 0))
                    key, conn = _graph_names(step.into)
                    out[key] = g
                    # The connectivities graph the step declares (atlas.rs:857-862).
                    out[conn] = dev.connectivities(g)
                if limited:
                    dev.trim()  # the stage's cached blocks back to the device before the next
        finally:
            for m in resident.values():
                m.free()
            dev.trim()
        if keep_layers:
            out["__made_layers__"] = made
        return out


def _graph_names(into):
    """The distances and connectivities graphs a neighbors step writes (atlas.rs:857-862)."""
    if into is None:
        return "distances", "connectivities"
    return f"{into}_distances", f"{into}_connectivities"


def _device_maps(maps):
    """The maps before a step as the device producer takes them: an optional
    normalize_total then an optional log1p (ky_csr_upload_counts), or None."""
    kinds = [m.kind for m in maps]
    if not kinds:
        return None
    if kinds not in (["normalize_total"], ["log1p"], ["normalize_total", "log1p"]):
        raise KytoError(KY_E501, f"the device makes a layer by normalize_total then log1p, not {' then '.join(kinds)}")
    target = maps[0].target_sum if kinds[0] == "normalize_total" else None
    return target, kinds[-1] == "log1p"


def _upload(dev: Device, layer, maps):
    if isinstance(layer, _StoreRef):
        target, log1p = maps if maps is not None else (None, False)
        with _store.StoreLayer(layer.store, layer.name) as opened:
            return dev.stream_layer(opened, target_sum=target, log1p=log1p)
    if maps is None:
        return dev.upload_csr(layer.indptr, layer.indices, layer.values, layer.n_genes)
    target, log1p = maps
    return dev.upload_counts(layer.indptr, layer.indices, layer.values, layer.n_genes, target_sum=target, log1p=log1p)


def _run_pca(dev: Device, layer, step: _Pca, maps=None, keep_layer=False):
    """The PCA of a step; with keep_layer, the made layer's values (the maps' output)
    come back as res.made_values."""
    if step.genes is None:
        csr = _upload(dev, layer, maps)
        try:
            res = dev.pca(csr, step.n_components, step.seed, step.power_iterations)
            res.made_values = csr.values() if keep_layer and maps is not None else None
        finally:
            csr.free()
        return res, res.device_embedding
    if isinstance(layer, _StoreRef):
        layer = layer.host()
    made_values = None
    if maps is not None:
        # The maps see every gene of a cell (its total), then the gene set is taken.
        made = _upload(dev, layer, maps)
        try:
            made_values = made.values()
            layer = _Layer(layer.indptr, layer.indices, made_values, layer.n_genes)
        finally:
            made.free()
    # A gene set: the layer's columns of the set, in order (pca.rs: the sums run in
    # the order a filter_genes before the node gives them), zero loadings elsewhere.
    flags = step.genes
    pos = np.cumsum(flags) - 1
    keep = flags[layer.indices]
    rows = np.repeat(np.arange(layer.n_cells), np.diff(layer.indptr.astype(np.int64)))[keep]
    indptr = np.zeros(layer.n_cells + 1, np.uint64)
    np.cumsum(np.bincount(rows, minlength=layer.n_cells), out=indptr[1:])
    n_set = int(flags.sum())
    if step.n_components > min(layer.n_cells, n_set):
        raise _invalid("pca", "n_components", step.n_components,
                       f"{layer.n_cells} cells and {n_set} genes (those of its gene set, `genes`) reach pca",
                       f"an int from 1 to {min(layer.n_cells, n_set)}")
    csr = dev.upload_csr(indptr, pos[layer.indices[keep]].astype(np.uint32), layer.values[keep], n_set)
    try:
        res = dev.pca(csr, step.n_components, step.seed, step.power_iterations)
    finally:
        csr.free()
    full = np.zeros((layer.n_genes, step.n_components), np.float32)
    full[flags] = res.loadings
    res.loadings = full
    res.made_values = made_values if keep_layer else None
    return res, res.device_embedding


def _run_neighbors(dev: Device, emb: DenseMatrix, step: _Neighbors, round_rows: int = 0) -> Graph:
    if step.search == "exact":
        return dev.neighbors(emb, step.n_neighbors, step.n_components, step.metric, step.precision, round_rows)
    if step.n_components is not None and step.n_components != emb.cols:
        raise KytoError(KY_E501, "neighbors: the approximate search uses every component")
    n_lists = max(1, min(4096, emb.rows // 256))
    index = dev.ivf_build(emb, n_lists=n_lists, seed=step.seed, label_k=step.n_neighbors - 1)
    try:
        return index.search(emb, step.n_neighbors)
    finally:
        index.free()


def _scan_store(path, default_layer) -> Atlas:
    """A plan over the layers of a kyto store (the reference's scan(store),
    python/kyto/_kyto.pyi:3820): its schema read now, no tile read until a step runs."""
    path = os.fspath(path)
    schema = _store.read_schema(path)
    tiling = schema["tiling"]
    refs = {m["name"]: _StoreRef(path, m["name"], int(schema["n_cells"]), int(schema["n_genes"]), int(m["nnz"]),
                                 m["dtype"], int(tiling["cell_block_size"]), int(tiling["gene_block_size"]))
            for m in schema.get("layers", [])}
    if not refs:
        raise KytoError(109, f"scan: the store at '{path}' has no layer")
    if default_layer is None:
        default_layer = "counts" if "counts" in refs else sorted(refs)[0]
    if default_layer not in refs:
        raise KytoError(109, f"scan: no layer named '{default_layer}' in the store at '{path}' (it has: "
                             f"{', '.join(sorted(refs))})")
    return Atlas(refs, default_layer)


def scan(layers, default_layer: str | None = None) -> Atlas:
    """A plan over a kyto store (a path), one CSR layer (named `default_layer`, "X" by
    default) or a dict of them (the default layer: the only one, else "X")."""
    if isinstance(layers, (str, os.PathLike)):
        return _scan_store(layers, default_layer)
    if default_layer is None:
        default_layer = next(iter(layers)) if isinstance(layers, dict) and len(layers) == 1 else "X"
    if isinstance(layers, dict):
        parsed = {k: _as_layer(v) for k, v in layers.items()}
    else:
        parsed = {default_layer: _as_layer(layers)}
    return Atlas(parsed, default_layer)
