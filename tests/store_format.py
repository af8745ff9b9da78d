"""An independent, pure-Python restatement of kyto's store format 0.1 (the
reference's docs/format.md §2-§8, zarrs' Zarr v3 sharding_indexed codec) for the tests
of the native store reader and writers: it writes stores the way kyto's StoreWriter
lays them out (store_writer.rs, zarr_arrays.rs:57-140, tile_encode.rs) and reads Zarr
arrays back, with pyarrow's zstd and a bitwise CRC-32C. It shares no code with
libkyto_gpu.so; the tests compare the two."""
from __future__ import annotations

import json
import os

import numpy as np
import pyarrow as pa

MISSING = 2**64 - 1


def crc32c(data: bytes) -> int:
    """CRC-32C, bit by bit (RFC 3720 B.4: check value 0xE3069283 for b'123456789')."""
    c = 0xFFFFFFFF
    for byte in data:
        c ^= byte
        for _ in range(8):
            c = (c >> 1) ^ (0x82F63B78 if c & 1 else 0)
    return c ^ 0xFFFFFFFF


def zstd(data: bytes, level: int = 3) -> bytes:
    return pa.Codec("zstd", compression_level=level).compress(data, asbytes=True)


def unzstd(data: bytes, size: int) -> bytes:
    return pa.Codec("zstd").decompress(data, decompressed_size=size, asbytes=True)


def _codecs(level):
    return [{"name": "bytes", "configuration": {"endian": "little"}},
            {"name": "zstd", "configuration": {"level": level, "checksum": False}}]


def _sharding(inner, level):
    return [{"name": "sharding_indexed", "configuration": {
        "chunk_shape": list(inner), "codecs": _codecs(level),
        "index_codecs": [{"name": "bytes", "configuration": {"endian": "little"}}, {"name": "crc32c"}],
        "index_location": "end"}}]


def _meta(path, shape, dtype, chunk, codecs, dims, attributes=None):
    os.makedirs(path, exist_ok=True)
    meta = {"zarr_format": 3, "node_type": "array", "shape": list(shape), "data_type": dtype,
            "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": list(chunk)}},
            "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
            "fill_value": 0, "codecs": codecs, "attributes": attributes or {}, "dimension_names": dims}
    with open(os.path.join(path, "zarr.json"), "w") as f:
        json.dump(meta, f)


def _group(path, attributes=None):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "zarr.json"), "w") as f:
        json.dump({"zarr_format": 3, "node_type": "group", "attributes": attributes or {}}, f)


def _shard(chunks: list, level: int) -> bytes | None:
    """A shard of inner chunks (raw little-endian bytes each, None or all zero:
    absent), index (offset, nbytes) u64 pairs + crc32c at the end; None when every
    inner chunk is absent."""
    body, index = b"", []
    for raw in chunks:
        if raw is None or not any(raw):
            index += [MISSING, MISSING]
            continue
        z = zstd(raw, level)
        index += [len(body), len(z)]
        body += z
    if not body:
        return None
    idx = np.array(index, "<u8").tobytes()
    return body + idx + crc32c(idx).to_bytes(4, "little")


def _write(path, data: bytes | None):
    if data is None:
        return
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(data)


def write_entries(path, values: np.ndarray, dtype: str, K: int, S: int, level: int):
    n = values.size
    _meta(path, [n], dtype, [S], _sharding([K], level), ["entry"])
    item = values.dtype.itemsize
    for s in range(-(-n // S)):
        chunks = []
        for j in range(S // K):
            lo = s * S + j * K
            if lo >= n:
                chunks.append(None)
                continue
            part = np.zeros(K, values.dtype)
            part[:min(K, n - lo)] = values[lo:lo + K]
            chunks.append(part.tobytes())
        _write(os.path.join(path, "c", str(s)), _shard(chunks, level))
    assert item == 4


def write_1d(path, values: np.ndarray, dtype: str, chunk: int, level: int, dims):
    _meta(path, [values.size], dtype, [chunk], _codecs(level), dims)
    for i in range(-(-values.size // chunk)):
        part = np.zeros(chunk, values.dtype)
        part[:min(chunk, values.size - i * chunk)] = values[i * chunk:(i + 1) * chunk]
        if part.any():
            _write(os.path.join(path, "c", str(i)), zstd(part.tobytes(), level))


def write_layer(store, name, indptr, indices, values, n_genes, C, G, K, S, level=3):
    """format.md §4 from a CSR (global gene indices)."""
    indptr = np.asarray(indptr, np.int64)
    n = len(indptr) - 1
    ncb, ngb = -(-n // C), -(-n_genes // G)
    tiles_idx, tiles_val, offsets = [], [], [0]
    tile_ptrs = np.zeros((ncb, ngb, C + 1), np.uint32)
    for b in range(ncb):
        rows = range(b * C, min(n, b * C + C))
        for g in range(ngb):
            lo, hi = g * G, min(n_genes, g * G + G)
            count = 0
            for r in rows:
                seg = indices[indptr[r]:indptr[r + 1]]
                keep = (seg >= lo) & (seg < hi)
                tiles_idx.append((seg[keep] - lo).astype(np.uint32))
                tiles_val.append(values[indptr[r]:indptr[r + 1]][keep])
                count += int(keep.sum())
                tile_ptrs[b, g, r - b * C + 1] = count
            tile_ptrs[b, g, len(rows) + 1:] = count
            offsets.append(offsets[-1] + count)
    local = np.concatenate(tiles_idx) if tiles_idx else np.zeros(0, np.uint32)
    vals = np.concatenate(tiles_val) if tiles_val else np.zeros(0, values.dtype)
    d = os.path.join(store, "layers", name)
    _group(d, {"kyto": {"kind": "tiled_csr"}})
    _meta(os.path.join(d, "indptr"), [ncb, ngb, C + 1], "uint32", [1, max(1, ngb), C + 1], _sharding([1, 1, C + 1], level),
          ["cell_block", "gene_block", "row"])
    for b in range(ncb):
        chunks = [tile_ptrs[b, g].tobytes() for g in range(ngb)]
        _write(os.path.join(d, "indptr", "c", str(b), "0", "0"), _shard(chunks, level) if chunks else None)
    write_1d(os.path.join(d, "tile_offsets"), np.array(offsets, np.uint64), "uint64", 65536, level, ["tile"])
    write_entries(os.path.join(d, "indices"), local, "uint32", K, S, level)
    write_entries(os.path.join(d, "data"), vals, "uint32" if vals.dtype == np.uint32 else "float32", K, S, level)
    return {"name": name, "dtype": "uint32" if vals.dtype == np.uint32 else "float32", "nnz": int(offsets[-1])}


def write_store(store, n_cells, n_genes, C, G, layers: dict, K=1 << 16, S=1 << 20, level=3, version="0.1"):
    """A store with `layers` {name: (indptr, indices, values)}; IDs tables omitted
    (the native reader does not read them)."""
    os.makedirs(store, exist_ok=True)
    for group in ("layers", "embeddings", "graphs"):
        _group(os.path.join(store, group))
    entries = [write_layer(store, name, *csr, n_genes, C, G, K, S, level) for name, csr in layers.items()]
    schema = {"n_cells": n_cells, "n_genes": n_genes, "tiling": {"cell_block_size": C, "gene_block_size": G},
              "cells": {"columns": [{"name": "cell_id", "dtype": "String"}]},
              "genes": {"columns": [{"name": "gene_id", "dtype": "String"}]},
              "layers": sorted(entries, key=lambda m: m["name"]), "embeddings": [], "graphs": []}
    with open(os.path.join(store, "schema.json"), "w") as f:
        json.dump(schema, f)
    _group(store, {"kyto": {"format_version": version}})


# ----------------------------------------------------------------- reading --

_DT = {"uint32": np.uint32, "uint64": np.uint64, "float32": np.float32, "float64": np.float64}


def read_array(path):
    """(metadata, values) of a Zarr v3 array as kyto writes it, read whole."""
    with open(os.path.join(path, "zarr.json")) as f:
        meta = json.load(f)
    shape = meta["shape"]
    dt = np.dtype(_DT[meta["data_type"]]).newbyteorder("<")
    chunk = meta["chunk_grid"]["configuration"]["chunk_shape"]
    codecs = meta["codecs"]
    out = np.zeros(shape, dt)
    grid = [-(-s // c) for s, c in zip(shape, chunk)]
    sharded = codecs[0]["name"] == "sharding_indexed"
    inner = codecs[0]["configuration"]["chunk_shape"] if sharded else chunk
    n_inner_grid = [c // i for c, i in zip(chunk, inner)]
    for coords in np.ndindex(*grid):
        key = os.path.join(path, "c", *map(str, coords))
        if not os.path.exists(key):
            continue
        with open(key, "rb") as f:
            raw = f.read()
        block = np.zeros(chunk, dt)
        if sharded:
            n_in = int(np.prod(n_inner_grid))
            idx = raw[-(n_in * 16 + 4):-4]
            assert crc32c(idx) == int.from_bytes(raw[-4:], "little"), key
            index = np.frombuffer(idx, "<u8").reshape(-1, 2)
            for lin, ic in enumerate(np.ndindex(*n_inner_grid)):
                off, nb = index[lin]
                if off == MISSING and nb == MISSING:
                    continue
                part = np.frombuffer(unzstd(raw[off:off + nb], int(np.prod(inner)) * dt.itemsize), dt).reshape(inner)
                sl = tuple(slice(a * i, (a + 1) * i) for a, i in zip(ic, inner))
                block[sl] = part
        else:
            block = np.frombuffer(unzstd(raw, int(np.prod(chunk)) * dt.itemsize), dt).reshape(chunk)
        sl = tuple(slice(a * c, min(s, (a + 1) * c)) for a, c, s in zip(coords, chunk, shape))
        out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]
    return meta, out
