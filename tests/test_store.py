"""The store reader and writers of libkyto_gpu.so (host code: no GPU) against an
independent pure-Python restatement of kyto's store format 0.1 (tests/store_format.py,
after the reference's docs/format.md §2-§10): stores written by one are read by the
other, bit for bit, over ragged cell and gene blocks, empty tiles, absent (all-zero)
chunks, inner chunks that straddle tiles and shards, uint32 and float32 values, and
windows of rows that cut cell blocks; corrupt stores and bad writer arguments give
kyto's storage error codes."""
import json
import os

import numpy as np
import pytest

from kyto_amd import _native as N
from kyto_amd import store as ks
from tests import store_format as sf


def _csr(n, g, density, seed, counts=False):
    rng = np.random.default_rng(seed)
    mask = rng.random((n, g)) < density
    mask[rng.integers(0, n, max(1, n // 10))] = False  # empty rows
    mask[:, g // 3: g // 3 + max(1, g // 10)] = False  # empty columns: empty tiles
    indptr = np.zeros(n + 1, np.uint64)
    indptr[1:] = np.cumsum(mask.sum(1))
    _, cols = np.nonzero(mask)
    if counts:
        values = rng.integers(1, 5000, cols.size).astype(np.uint32)
    else:
        values = ((rng.random(cols.size) + 0.01) * rng.choice([-1.0, 1.0], cols.size)).astype(np.float32)
        values[::97] = np.nan  # NaN is a value (format.md §4.6)
    return indptr, cols.astype(np.uint32), values


def _rows(csr, first, n):
    ip, ix, v = csr
    lo, hi = int(ip[first]), int(ip[first + n])
    return ip[first:first + n + 1] - ip[first], ix[lo:hi], v[lo:hi]


def _same(got, want, counts):
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1])
    w = want[2].astype(np.float32) if counts else want[2]
    assert np.array_equal(got[2].view(np.uint32), w.view(np.uint32))


CONFIGS = [
    # n, genes, C, G, K, S: ragged cell and gene blocks, chunks smaller than tiles
    (1000, 300, 128, 64, 256, 1024),
    # one gene block (G > genes): chunks decode straight into the output
    (777, 100, 256, 512, 64, 64),
    # one tile, the writer defaults
    (500, 50, 500, 50, 1 << 16, 1 << 20),
    # inner chunks larger than cell blocks' runs
    (300, 40, 16, 7, 4096, 8192),
]


def test_crc32c_check_value():
    assert sf.crc32c(b"123456789") == 0xE3069283


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("counts", [False, True])
def test_reader_of_python_written_store(tmp_path, cfg, counts):
    n, g, C, G, K, S = cfg
    csr = _csr(n, g, 0.08, seed=n + g, counts=counts)
    sf.write_store(str(tmp_path / "s"), n, g, C, G, {"x": csr}, K=K, S=S)
    with ks.StoreLayer(tmp_path / "s", "x") as layer:
        assert (layer.n_cells, layer.n_genes, layer.nnz) == (n, g, int(csr[0][-1]))
        assert (layer.cell_block, layer.gene_block, layer.counts) == (C, G, counts)
        _same(layer.read_rows(), csr, counts)
        for first, count in [(0, C), (C, min(C, n - C)), (37, min(300, n - 37)), (n - 1, 1), (5, 0)]:
            _same(layer.read_rows(first, count), _rows(csr, first, count), counts)


@pytest.mark.parametrize("cfg", CONFIGS)
def test_native_layer_writer_against_python_reader(tmp_path, cfg):
    n, g, C, G, K, S = cfg
    csr = _csr(n, g, 0.08, seed=7 + n, counts=True)
    sf.write_store(str(tmp_path / "ref"), n, g, C, G, {"x": csr}, K=K, S=S)
    os.makedirs(tmp_path / "nat")
    ks.write_layer(tmp_path / "nat", "x", *csr, n_genes=g, cell_block=C, gene_block=G, nnz_chunk=K, nnz_shard=S)
    for array in ("indptr", "tile_offsets", "indices", "data"):
        m_ref, v_ref = sf.read_array(str(tmp_path / "ref" / "layers" / "x" / array))
        m_nat, v_nat = sf.read_array(str(tmp_path / "nat" / "layers" / "x" / array))
        assert np.array_equal(v_ref, v_nat), array
        for key in ("shape", "data_type", "chunk_grid", "chunk_key_encoding", "dimension_names"):
            assert m_ref[key] == m_nat[key], (array, key)
        assert m_ref["codecs"] == m_nat["codecs"], array
    with open(tmp_path / "nat" / "layers" / "x" / "zarr.json") as f:
        assert json.load(f)["attributes"] == {"kyto": {"kind": "tiled_csr"}}
    # Absent chunks are the same ones: empty tiles have no indptr inner chunk.
    for root, _, files in os.walk(tmp_path / "ref" / "layers" / "x"):
        rel = os.path.relpath(root, tmp_path / "ref")
        assert sorted(f for f in files if f != "zarr.json") == sorted(
            f for f in os.listdir(tmp_path / "nat" / rel) if f != "zarr.json" and os.path.isfile(tmp_path / "nat" / rel / f)), rel


def test_absent_all_zero_chunks(tmp_path):
    # Every cell stores gene 0 only: the indices' inner chunks are all zero, so the
    # writer stores none (format.md §4.4) and the reader reads them as zeros.
    n, g = 700, 20
    indptr = np.arange(n + 1, dtype=np.uint64)
    indices = np.zeros(n, np.uint32)
    values = np.arange(1, n + 1, dtype=np.uint32)
    sf.write_store(str(tmp_path / "s"), n, g, 64, 10, {"x": (indptr, indices, values)}, K=128, S=256)
    assert not os.path.exists(tmp_path / "s" / "layers" / "x" / "indices" / "c")
    with ks.StoreLayer(tmp_path / "s", "x") as layer:
        _same(layer.read_rows(), (indptr, indices, values), True)


def _store(tmp_path, **kw):
    n, g = 600, 90
    csr = _csr(n, g, 0.1, seed=3)
    sf.write_store(str(tmp_path / "s"), n, g, 128, 32, {"x": csr}, K=512, S=2048, **kw)
    return tmp_path / "s", csr


def _expect(code, fn):
    with pytest.raises(N.KytoError, match=code):
        fn()


def test_refusals_of_the_store(tmp_path):
    s, _ = _store(tmp_path)
    _expect("KY-E109", lambda: ks.StoreLayer(s, "counts"))
    _expect("KY-E116", lambda: ks.StoreLayer(tmp_path, "x"))
    s2, _ = _store(tmp_path / "v", version="0.2")
    _expect("KY-E115", lambda: ks.StoreLayer(s2, "x"))


def _flip(path, at):
    with open(path, "r+b") as f:
        f.seek(at, os.SEEK_END if at < 0 else os.SEEK_SET)
        b = f.read(1)
        f.seek(-1, os.SEEK_CUR)
        f.write(bytes([b[0] ^ 0x40]))


def test_corrupt_shard_index(tmp_path):
    s, _ = _store(tmp_path)
    _flip(s / "layers" / "x" / "indices" / "c" / "0", -10)  # inside the index
    _expect("KY-E107.*checksum", lambda: ks.StoreLayer(s, "x"))


def test_corrupt_tile_offsets(tmp_path):
    s, csr = _store(tmp_path)
    schema = json.loads((s / "schema.json").read_text())
    schema["layers"][0]["nnz"] += 1
    (s / "schema.json").write_text(json.dumps(schema))
    _expect("KY-E107", lambda: ks.StoreLayer(s, "x"))


def test_corrupt_tile_indptr(tmp_path):
    s, csr = _store(tmp_path)
    # Tile (0, 0)'s indptr ending one entry short of what tile_offsets gives.
    _, tp = sf.read_array(str(s / "layers" / "x" / "indptr"))
    chunks = [tp[0, g].copy() for g in range(tp.shape[1])]
    chunks[0][-1] -= 1
    with open(s / "layers" / "x" / "indptr" / "c" / "0" / "0" / "0", "wb") as f:
        f.write(sf._shard([c.tobytes() for c in chunks], 3))
    with ks.StoreLayer(s, "x") as layer:
        _expect("KY-E107.*tile \\(0, 0\\)", lambda: layer.read_rows(0, 10))
        layer.read_rows(128, 100)  # other cell blocks read as before


def test_corrupt_local_index(tmp_path):
    s, csr = _store(tmp_path)
    _, local = sf.read_array(str(s / "layers" / "x" / "indices"))
    local = local.copy()
    local[0] = 31 + 40  # beyond gene block 0's width of 32
    sf.write_entries(str(s / "layers" / "x" / "indices"), local, "uint32", 512, 2048, 3)
    with ks.StoreLayer(s, "x") as layer:
        _expect("KY-E107.*not below", lambda: layer.read_rows(0, 128))


def test_chunk_of_the_wrong_size(tmp_path):
    s, csr = _store(tmp_path)
    _, offsets = sf.read_array(str(s / "layers" / "x" / "tile_offsets"))
    with open(s / "layers" / "x" / "tile_offsets" / "c" / "0", "wb") as f:
        f.write(sf.zstd(offsets[:100].tobytes()))
    _expect("KY-E107.*decodes to", lambda: ks.StoreLayer(s, "x"))


def test_embedding_writer_and_reader(tmp_path):
    rng = np.random.default_rng(0)
    emb = rng.normal(size=(1000, 50)).astype(np.float32)
    emb[128:256] = 0.0  # a chunk of the fill value: not stored
    os.makedirs(tmp_path / "s")
    ks.write_embedding(tmp_path / "s", "X_pca", emb, axis="cells", chunk_rows=128)
    meta, got = sf.read_array(str(tmp_path / "s" / "embeddings" / "X_pca"))
    assert np.array_equal(got.view(np.uint32), emb.view(np.uint32))
    assert meta["shape"] == [1000, 50] and meta["data_type"] == "float32"
    assert meta["chunk_grid"]["configuration"]["chunk_shape"] == [128, 50]
    assert [c["name"] for c in meta["codecs"]] == ["bytes", "zstd"]
    assert meta["attributes"] == {"kyto": {"axis": "cells"}}
    assert meta["dimension_names"] == ["cell", "component"] and meta["fill_value"] == 0
    assert not os.path.exists(tmp_path / "s" / "embeddings" / "X_pca" / "c" / "1")
    assert os.path.exists(tmp_path / "s" / "embeddings" / "X_pca" / "c" / "7" / "0")  # the padded last chunk
    ks.write_embedding(tmp_path / "s", "X_pca_loadings", emb[:90, :20], axis="genes", chunk_rows=32)
    meta, got = sf.read_array(str(tmp_path / "s" / "embeddings" / "X_pca_loadings"))
    assert meta["dimension_names"] == ["gene", "component"] and np.array_equal(got, emb[:90, :20])


def test_graph_writer_and_reader(tmp_path):
    rng = np.random.default_rng(1)
    n, k = 3000, 14
    idx = np.sort(np.stack([rng.choice(n, k, replace=False) for _ in range(n)]), axis=1).astype(np.uint32)
    w = rng.random((n, k)).astype(np.float32)
    indptr = np.arange(0, n * k + 1, k, dtype=np.uint64)
    os.makedirs(tmp_path / "s")
    ks.write_graph(tmp_path / "s", "distances", indptr, idx.ravel(), w.ravel(), nnz_chunk=1000, nnz_shard=8000)
    g = tmp_path / "s" / "graphs" / "distances"
    with open(g / "zarr.json") as f:
        assert json.load(f)["attributes"] == {"kyto": {"kind": "csr", "axis": "cells"}}
    m, ip = sf.read_array(str(g / "indptr"))
    assert np.array_equal(ip, indptr) and m["chunk_grid"]["configuration"]["chunk_shape"] == [65536]
    assert m["dimension_names"] == ["row"]
    m, ix = sf.read_array(str(g / "indices"))
    assert np.array_equal(ix, idx.ravel()) and m["codecs"][0]["configuration"]["chunk_shape"] == [1000]
    assert m["chunk_grid"]["configuration"]["chunk_shape"] == [8000]
    _, d = sf.read_array(str(g / "data"))
    assert np.array_equal(d, w.ravel())
    # An empty graph: no chunks of indices or data.
    ks.write_graph(tmp_path / "s", "empty", np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    assert not os.path.exists(tmp_path / "s" / "graphs" / "empty" / "indices" / "c")


def test_create_store_round_trip(tmp_path):
    n, g = 1500, 120
    counts = _csr(n, g, 0.1, seed=11, counts=True)
    rng = np.random.default_rng(2)
    emb = rng.normal(size=(n, 50)).astype(np.float32)
    load = rng.normal(size=(g, 50)).astype(np.float32)
    k = 9
    idx = np.sort(np.stack([rng.choice(n, k, replace=False) for _ in range(n)]), axis=1).astype(np.uint32)
    w = rng.random(n * k).astype(np.float32)
    ip = np.arange(0, n * k + 1, k, dtype=np.uint64)
    path = tmp_path / "out.kyto"
    ks.create_store(path, n_cells=n, n_genes=g, cell_block=512, gene_block=64,
                    layers={"counts": ks.LayerData(*counts)},
                    embeddings={"X_pca": ("cells", emb), "X_pca_loadings": ("genes", load)},
                    graphs={"distances": (ip, idx.ravel(), w)})
    schema = ks.read_schema(path)
    assert schema["n_cells"] == n and schema["tiling"] == {"cell_block_size": 512, "gene_block_size": 64}
    assert schema["layers"] == [{"name": "counts", "dtype": "uint32", "nnz": int(counts[0][-1])}]
    assert [e["name"] for e in schema["embeddings"]] == ["X_pca", "X_pca_loadings"]
    assert schema["graphs"] == [{"name": "distances", "axis": "cells", "dtype": "float32", "nnz": n * k}]
    assert [f for f in os.listdir(tmp_path) if f.startswith(".")] == []  # no temporary left
    with ks.StoreLayer(path, "counts") as layer:
        _same(layer.read_rows(), counts, True)
    assert np.array_equal(ks.read_embedding(path, "X_pca"), emb)
    assert np.array_equal(ks.read_embedding(path, "X_pca_loadings"), load)
    gi, gx, gw = ks.read_graph(path, "distances")
    assert np.array_equal(gi, ip) and np.array_equal(gx, idx.ravel()) and np.array_equal(gw, w)
    import pyarrow.parquet as pq
    assert pq.read_table(path / "cells.parquet").column("cell_id").to_pylist()[:2] == ["c0", "c1"]
    _expect("KY-E118", lambda: ks.create_store(path, n_cells=n, n_genes=g, cell_block=512, gene_block=64))
    ks.create_store(path, n_cells=n, n_genes=g, cell_block=512, gene_block=64, overwrite=True,
                    tables_from=None, embeddings={"X_pca": ("cells", emb)})
    assert ks.read_schema(path)["layers"] == []
    _expect("KY-E109", lambda: ks.read_graph(path, "distances"))


def test_writer_refusals(tmp_path):
    os.makedirs(tmp_path / "s")
    emb = np.ones((10, 3), np.float32)
    for bad in ("", "CON", "lpt1.txt", ".hidden", "__x", "a:b", "end.", "x" * 65):
        _expect("KY-E102", lambda: ks.write_embedding(tmp_path / "s", bad, emb, chunk_rows=4))
    ks.write_embedding(tmp_path / "s", "Emb", emb, chunk_rows=4)
    _expect("KY-E103", lambda: ks.write_embedding(tmp_path / "s", "emb", emb, chunk_rows=4))
    _expect("KY-E104", lambda: ks.write_embedding(tmp_path / "s", "e2", emb, chunk_rows=4, zstd_level=0))
    ip = np.array([0, 1, 2], np.uint64)
    for K, S in ((0, 8), (4, 6)):
        _expect("KY-E104", lambda: ks.write_graph(tmp_path / "s", "g", ip, np.array([1, 0], np.uint32),
                                                  np.ones(2, np.float32), nnz_chunk=K, nnz_shard=S))
    _expect("KY-E215", lambda: ks.write_graph(tmp_path / "s", "g", ip, np.array([2, 0], np.uint32),
                                              np.ones(2, np.float32)))
    _expect("KY-E106", lambda: ks.write_layer(tmp_path / "s", "x", ip, np.array([0, 1], np.uint32),
                                              np.array([1.0, 0.0], np.float32), 3, 2, 2))
    _expect("KY-E106", lambda: ks.write_layer(tmp_path / "s", "x", ip, np.array([0, 5], np.uint32),
                                              np.array([1.0, 2.0], np.float32), 3, 2, 2))


def test_scan_of_a_store(tmp_path):
    import kyto_amd
    counts = _csr(400, 60, 0.1, seed=4, counts=True)
    path = tmp_path / "p.kyto"
    ks.create_store(path, n_cells=400, n_genes=60, cell_block=128, gene_block=32,
                    layers={"counts": ks.LayerData(*counts), "other": ks.LayerData(*_csr(400, 60, 0.1, seed=5))})
    atlas = kyto_amd.scan(path)
    assert atlas.default_layer == "counts" and set(atlas.layers) == {"counts", "other"}
    plan = atlas.normalize_total(1e4).log1p().pca(n_components=10, seed=0).neighbors(n_neighbors=5)
    assert 'PCA "log1p"' in plan.explain() and "SCAN (counts, other)" in plan.explain()
    assert kyto_amd.scan(path, default_layer="other").default_layer == "other"
    _expect("KY-E109", lambda: kyto_amd.scan(path, default_layer="x"))
    _expect("KY-E109", lambda: atlas.pca(layer="x", seed=0))
    ref = atlas.layers["counts"].host()
    _same((ref.indptr, ref.indices, ref.values), counts, True)


def test_sink_writes_the_plan_outputs(tmp_path, monkeypatch):
    # The sink's assembly on the CPU: collect's outputs stood in for (the device run
    # is tests/test_gpu_store.py), the store written and read back.
    import kyto_amd
    from kyto_amd.device import ConnectivityGraph, Graph
    n, g = 300, 40
    counts = _csr(n, g, 0.1, seed=6, counts=True)
    src = tmp_path / "in.kyto"
    ks.create_store(src, n_cells=n, n_genes=g, cell_block=128, gene_block=32, layers={"counts": ks.LayerData(*counts)})
    plan = kyto_amd.scan(src).normalize_total(1e4).log1p().pca(n_components=5, seed=0).neighbors(n_neighbors=4)
    rng = np.random.default_rng(0)
    made = counts[2].astype(np.float32) * 0.5
    made[3] = 0.0  # a made zero is not stored
    emb, load = rng.normal(size=(n, 5)).astype(np.float32), rng.normal(size=(g, 5)).astype(np.float32)
    idx = np.sort(np.stack([rng.choice(n, 3, replace=False) for _ in range(n)]), axis=1).astype(np.uint32).ravel()
    ip = np.arange(0, 3 * n + 1, 3, dtype=np.uint64)
    dist, w = rng.random(3 * n).astype(np.float32), rng.random(3 * n).astype(np.float32)

    def fake_run(self, memory_limit, threads, progress, keep_layers):
        assert keep_layers
        return {"X_pca": emb, "X_pca_loadings": load, "distances": Graph(4, ip, idx, dist),
                "connectivities": ConnectivityGraph(ip, idx, w, None, None), "__made_layers__": {"log1p": ("counts", made)}}

    monkeypatch.setattr(kyto_amd.Atlas, "_run", fake_run)
    out = tmp_path / "out.kyto"
    plan.sink_zarr(out)
    schema = ks.read_schema(out)
    assert [m["name"] for m in schema["layers"]] == ["counts", "log1p"]
    assert schema["layers"][1] == {"name": "log1p", "dtype": "float32", "nnz": int(counts[0][-1]) - 1}
    assert schema["tiling"] == {"cell_block_size": 128, "gene_block_size": 32}
    with ks.StoreLayer(out, "counts") as layer:
        _same(layer.read_rows(), counts, True)
    with ks.StoreLayer(out, "log1p") as layer:
        lip, lix, lv = layer.read_rows()
    keep = made != 0
    assert np.array_equal(lix, counts[1][keep]) and np.array_equal(lv, made[keep]) and int(lip[-1]) == keep.sum()
    assert np.array_equal(ks.read_embedding(out, "X_pca"), emb)
    assert np.array_equal(ks.read_embedding(out, "X_pca_loadings"), load)
    assert np.array_equal(ks.read_graph(out, "distances")[2], dist)
    assert np.array_equal(ks.read_graph(out, "connectivities")[2], w)
    assert (out / "cells.parquet").read_bytes() == (src / "cells.parquet").read_bytes()
