"""kyto stores on disk (the reference's docs/format.md, store format "0.1") for the
embedding path: a tiled layer opened for reading (its rows decoded by the native
reader, crates/kyto-store/src/tile_reader.rs), the native writers of layers,
embeddings and graphs (tile_encode.rs, embedding_io.rs, graph_io.rs), and
`create_store`, which writes a whole store the way kyto's StoreWriter does
(store_writer.rs, format.md §10): a temporary directory next to the target, tables,
members, provenance.json, schema.json, the root zarr.json last, then a rename.

Every byte of the Zarr members (chunks, shards, zstd, crc32c, zarr.json) is written and
read by libkyto_gpu.so; this module holds the JSON files and the table files around
them.
"""
from __future__ import annotations

import json
import os
import secrets
import shutil
from ctypes import byref, c_int, c_uint32, c_uint64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _native as N
from ._native import KytoError

FORMAT_VERSION = "0.1"
# Writer defaults of kyto_store (write_options.rs:9-15).
NNZ_CHUNK = 1 << 16
NNZ_SHARD = 1 << 20
ZSTD_LEVEL = 3
AXES = {"cells": 0, "genes": 1}


def _b(path) -> bytes:
    return os.fsencode(os.fspath(path))


def read_schema(store) -> dict:
    """The store's schema.json (format.md §8), after the root's format_version check
    (KY-E116 not a store, KY-E115 another version)."""
    store = os.fspath(store)
    root = os.path.join(store, "zarr.json")
    if not os.path.exists(root):
        raise KytoError(116, f"'{store}' is not a kyto store: its root zarr.json is missing")
    with open(root, encoding="utf-8") as f:
        version = (json.load(f).get("attributes") or {}).get("kyto", {}).get("format_version")
    if version is None:
        raise KytoError(116, f"'{store}' is not a kyto store: its root zarr.json has no attributes.kyto.format_version")
    if version != FORMAT_VERSION:
        raise KytoError(115, f"the store at '{store}' has format_version '{version}', but this library reads only "
                             f"{FORMAT_VERSION}")
    with open(os.path.join(store, "schema.json"), encoding="utf-8") as f:
        return json.load(f)


class StoreLayer:
    """A layer of a kyto store opened for reading (ky_store_layer_open)."""

    def __init__(self, store, name: str):
        self.store = os.fspath(store)
        self.name = name
        self._lib = N.gpu_lib()
        h = c_void_p()
        N.check(self._lib.ky_store_layer_open(_b(self.store), name.encode(), byref(h)))
        self._h = h
        n, g, nnz, cb, gb, counts = c_uint64(), c_uint32(), c_uint64(), c_uint64(), c_uint64(), c_int()
        N.check(self._lib.ky_store_layer_shape(h, byref(n), byref(g), byref(nnz), byref(cb), byref(gb), byref(counts)))
        self.n_cells, self.n_genes, self.nnz = int(n.value), int(g.value), int(nnz.value)
        self.cell_block, self.gene_block = int(cb.value), int(gb.value)
        self.counts = bool(counts.value)

    @property
    def handle(self):
        if self._h is None:
            raise KytoError(N.KY_E215, f"layer '{self.name}' of '{self.store}' is closed")
        return self._h

    def read_rows(self, first: int = 0, n: int | None = None):
        """Rows [first, first + n) as host arrays (indptr u64 from 0, global gene
        indices u32, values f32; uint32 counts converted)."""
        n = self.n_cells - first if n is None else n
        indptr = np.zeros(n + 1, np.uint64)
        nnz = c_uint64()
        empty_i, empty_v = np.zeros(1, np.uint32), np.zeros(1, np.float32)
        N.check(self._lib.ky_store_layer_rows(self.handle, first, n, N.ptr(indptr, N.u64p), N.ptr(empty_i, N.u32p),
                                              N.ptr(empty_v, N.f32p), 0, byref(nnz)))
        got = nnz.value
        indices, values = np.empty(max(got, 1), np.uint32), np.empty(max(got, 1), np.float32)
        N.check(self._lib.ky_store_layer_rows(self.handle, first, n, N.ptr(indptr, N.u64p), N.ptr(indices, N.u32p),
                                              N.ptr(values, N.f32p), indices.size, byref(nnz)))
        return indptr, indices[:got], values[:got]

    def close(self):
        if self._h is not None:
            N.check(self._lib.ky_store_layer_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def write_layer(store, name, indptr, indices, values, n_genes: int, cell_block: int, gene_block: int, *,
                nnz_chunk: int = NNZ_CHUNK, nnz_shard: int = NNZ_SHARD, zstd_level: int = ZSTD_LEVEL) -> None:
    """Layer `name` of a store under construction (ky_store_write_layer); uint32 values
    make a counts layer, anything else is written as float32."""
    indptr = np.ascontiguousarray(indptr, np.uint64)
    indices = np.ascontiguousarray(indices, np.uint32)
    counts = np.asarray(values).dtype == np.uint32
    values = np.ascontiguousarray(values, np.uint32 if counts else np.float32)
    N.check(N.gpu_lib().ky_store_write_layer(_b(store), name.encode(), len(indptr) - 1, int(n_genes),
                                             N.ptr(indptr, N.u64p), N.ptr(indices, N.u32p), values.ctypes.data_as(c_void_p),
                                             1 if counts else 0, int(cell_block), int(gene_block), int(nnz_chunk),
                                             int(nnz_shard), int(zstd_level)))


def write_embedding(store, name, values, *, axis: str = "cells", chunk_rows: int, zstd_level: int = ZSTD_LEVEL) -> None:
    """Embedding `name` (rows x width float32) of a store under construction
    (ky_store_write_embedding): chunks of [chunk_rows, width]."""
    if axis not in AXES:
        raise KytoError(N.KY_E215, f"axis={axis!r}: an embedding's axis is 'cells' or 'genes'")
    values = np.ascontiguousarray(values, np.float32)
    if values.ndim != 2:
        raise KytoError(N.KY_E215, "an embedding is a 2-D array")
    N.check(N.gpu_lib().ky_store_write_embedding(_b(store), name.encode(), AXES[axis], N.ptr(values, N.f32p),
                                                 values.shape[0], values.shape[1], int(chunk_rows), int(zstd_level)))


def read_embedding(store, name) -> np.ndarray:
    lib = N.gpu_lib()
    rows, width = c_uint64(), c_uint32()
    N.check(lib.ky_store_read_embedding(_b(store), name.encode(), byref(rows), byref(width), None, 0))
    out = np.empty((rows.value, width.value), np.float32)
    N.check(lib.ky_store_read_embedding(_b(store), name.encode(), byref(rows), byref(width), N.ptr(out, N.f32p),
                                        out.size))
    return out


def write_graph(store, name, indptr, indices, weights, *, nnz_chunk: int = NNZ_CHUNK, nnz_shard: int = NNZ_SHARD,
                zstd_level: int = ZSTD_LEVEL) -> None:
    """Cell graph `name` (CSR, float32 weights) of a store under construction
    (ky_store_write_graph)."""
    indptr = np.ascontiguousarray(indptr, np.uint64)
    indices = np.ascontiguousarray(indices, np.uint32)
    weights = np.ascontiguousarray(weights, np.float32)
    if indices.size != int(indptr[-1]) or weights.size != indices.size:
        raise KytoError(N.KY_E215, f"{weights.size} weights and {indices.size} indices for {int(indptr[-1])} edges")
    if indices.size == 0:
        indices, weights = np.zeros(1, np.uint32), np.zeros(1, np.float32)
    N.check(N.gpu_lib().ky_store_write_graph(_b(store), name.encode(), len(indptr) - 1, N.ptr(indptr, N.u64p),
                                             N.ptr(indices, N.u32p), N.ptr(weights, N.f32p), int(nnz_chunk),
                                             int(nnz_shard), int(zstd_level)))


def read_graph(store, name):
    """(indptr u64, indices u32, weights f32) of a float32 cell graph."""
    lib = N.gpu_lib()
    n, e = c_uint64(), c_uint64()
    N.check(lib.ky_store_read_graph(_b(store), name.encode(), byref(n), byref(e), None, None, None))
    indptr = np.empty(n.value + 1, np.uint64)
    indices, weights = np.empty(max(e.value, 1), np.uint32), np.empty(max(e.value, 1), np.float32)
    N.check(lib.ky_store_read_graph(_b(store), name.encode(), byref(n), byref(e), N.ptr(indptr, N.u64p),
                                    N.ptr(indices, N.u32p), N.ptr(weights, N.f32p)))
    return indptr, indices[:e.value], weights[:e.value]


@dataclass
class LayerData:
    """A host CSR for create_store (uint32 values: counts)."""
    indptr: np.ndarray
    indices: np.ndarray
    values: np.ndarray


def _fsync_tree(path: str) -> None:
    for root, dirs, files in os.walk(path):
        for name in files:
            fd = os.open(os.path.join(root, name), os.O_RDONLY)
            try:
                os.fsync(fd)
            finally:
                os.close(fd)
        fd = os.open(root, os.O_RDONLY)
        try:
            os.fsync(fd)
        finally:
            os.close(fd)


def _write_json(path: str, obj) -> None:
    with open(path, "w", encoding="utf-8") as f:
        json.dump(obj, f, indent=2)
        f.write("\n")


def _write_ids(path: str, column: str, ids) -> None:
    import pyarrow as pa
    import pyarrow.parquet as pq
    pq.write_table(pa.table({column: pa.array([str(i) for i in ids], pa.string())}), path, compression="zstd")


def _is_store(path: str) -> bool:
    try:
        with open(os.path.join(path, "zarr.json"), encoding="utf-8") as f:
            return "format_version" in ((json.load(f).get("attributes") or {}).get("kyto") or {})
    except (OSError, ValueError):
        return False


def create_store(path, *, n_cells: int, n_genes: int, cell_block: int, gene_block: int, layers=None,
                 embeddings=None, graphs=None, cell_ids=None, gene_ids=None, tables_from=None, copy_layers=None,
                 overwrite: bool = False, nnz_chunk: int = NNZ_CHUNK, nnz_shard: int = NNZ_SHARD,
                 zstd_level: int = ZSTD_LEVEL, provenance=None) -> None:
    """Write a new kyto store at `path` (format.md §10): layers {name: LayerData},
    embeddings {name: (axis, array)}, graphs {name: (indptr, indices, weights)}; the
    cell and gene tables hold the IDs (`c0`, `g0`, ... when not given), or are copied
    from the store `tables_from`, as are the layers named in `copy_layers`. Written in
    `.<name>.kyto-tmp-<16 hex>` next to `path` and renamed into place; an existing
    kyto store at `path` is replaced only with overwrite=True (KY-E118 otherwise)."""
    path = os.path.abspath(os.fspath(path))
    parent, base = os.path.split(path)
    if os.path.lexists(path):
        if not overwrite:
            raise KytoError(118, f"'{path}' already exists; pass overwrite=True to replace a kyto store")
        if os.path.islink(path) or not _is_store(path):
            raise KytoError(118, f"'{path}' exists and is not a kyto store; it is never replaced")
    tmp = os.path.join(parent, f".{base}.kyto-tmp-{secrets.token_hex(8)}")
    os.makedirs(tmp)
    try:
        layers, embeddings, graphs = dict(layers or {}), dict(embeddings or {}), dict(graphs or {})
        copy_layers = list(copy_layers or [])
        for group in ("layers", "embeddings", "graphs"):
            os.makedirs(os.path.join(tmp, group), exist_ok=True)
        if tables_from is not None:
            for table in ("cells.parquet", "genes.parquet"):
                shutil.copyfile(os.path.join(os.fspath(tables_from), table), os.path.join(tmp, table))
        else:
            _write_ids(os.path.join(tmp, "cells.parquet"), "cell_id",
                       cell_ids if cell_ids is not None else (f"c{i}" for i in range(n_cells)))
            _write_ids(os.path.join(tmp, "genes.parquet"), "gene_id",
                       gene_ids if gene_ids is not None else (f"g{j}" for j in range(n_genes)))
        schema_layers, schema_emb, schema_graphs = [], [], []
        source_schema = read_schema(tables_from) if copy_layers else None
        for name in copy_layers:
            entry = next(m for m in source_schema["layers"] if m["name"] == name)
            shutil.copytree(os.path.join(os.fspath(tables_from), "layers", name), os.path.join(tmp, "layers", name))
            schema_layers.append(dict(entry))
        for name, layer in layers.items():
            write_layer(tmp, name, layer.indptr, layer.indices, layer.values, n_genes, cell_block, gene_block,
                        nnz_chunk=nnz_chunk, nnz_shard=nnz_shard, zstd_level=zstd_level)
            dtype = "uint32" if np.asarray(layer.values).dtype == np.uint32 else "float32"
            schema_layers.append({"name": name, "dtype": dtype, "nnz": int(layer.indptr[-1])})
        for name, (axis, values) in embeddings.items():
            expect = n_cells if axis == "cells" else n_genes
            if values.shape[0] != expect:
                raise KytoError(N.KY_E215, f"embedding '{name}' has {values.shape[0]} rows, not the store's {expect} {axis}")
            write_embedding(tmp, name, values, axis=axis, chunk_rows=cell_block if axis == "cells" else gene_block,
                            zstd_level=zstd_level)
            schema_emb.append({"name": name, "axis": axis, "width": int(values.shape[1]), "dtype": "float32"})
        for name, (indptr, indices, weights) in graphs.items():
            if len(indptr) != n_cells + 1:
                raise KytoError(N.KY_E215, f"graph '{name}' has {len(indptr) - 1} rows, not the store's {n_cells} cells")
            write_graph(tmp, name, indptr, indices, weights, nnz_chunk=nnz_chunk, nnz_shard=nnz_shard,
                        zstd_level=zstd_level)
            schema_graphs.append({"name": name, "axis": "cells", "dtype": "float32", "nnz": int(indptr[-1])})
        for group in ("layers", "embeddings", "graphs"):
            _write_json(os.path.join(tmp, group, "zarr.json"), {"zarr_format": 3, "node_type": "group", "attributes": {}})
        _write_json(os.path.join(tmp, "provenance.json"), provenance or {
            "kyto_version": "kyto_amd", "dependencies": {"zstd": _zstd_version()}, "plan": None, "source": None,
            "seeds": {}, "device": "mi355x", "devices": None, "run": {"started_at": None, "finished_at": None,
                                                                      "timings": {}}})
        if tables_from is not None:
            src = read_schema(tables_from)
            cells, genes = src["cells"], src["genes"]
        else:
            cells = {"columns": [{"name": "cell_id", "dtype": "String"}]}
            genes = {"columns": [{"name": "gene_id", "dtype": "String"}]}
        by_name = lambda m: m["name"].encode()  # noqa: E731 - byte order (format.md §8)
        _write_json(os.path.join(tmp, "schema.json"), {
            "n_cells": int(n_cells), "n_genes": int(n_genes),
            "tiling": {"cell_block_size": int(cell_block), "gene_block_size": int(gene_block)},
            "cells": cells, "genes": genes,
            "layers": sorted(schema_layers, key=by_name), "embeddings": sorted(schema_emb, key=by_name),
            "graphs": sorted(schema_graphs, key=by_name)})
        _write_json(os.path.join(tmp, "zarr.json"), {"zarr_format": 3, "node_type": "group",
                                                     "attributes": {"kyto": {"format_version": FORMAT_VERSION}}})
        _fsync_tree(tmp)
        if os.path.lexists(path):
            old = os.path.join(parent, f".{base}.kyto-old-{secrets.token_hex(8)}")
            os.rename(path, old)
            os.rename(tmp, path)
            shutil.rmtree(old, ignore_errors=True)
        else:
            os.rename(tmp, path)
        fd = os.open(parent, os.O_RDONLY)
        try:
            os.fsync(fd)
        finally:
            os.close(fd)
    except BaseException:
        shutil.rmtree(tmp, ignore_errors=True)
        raise


def _zstd_version() -> str:
    import ctypes
    v = ctypes.CDLL("libzstd.so.1").ZSTD_versionNumber()
    return f"{v // 10000}.{v // 100 % 100}.{v % 100}"
