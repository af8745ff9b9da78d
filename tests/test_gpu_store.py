"""A kyto store on disk into the device layer and back out: the store's counts layer
streamed cell block by cell block through the pinned buffers (ky_csr_stream_layer)
equals the layer that the in-memory counts make on the device, bit for bit, for
aligned and unaligned blocks; a corrupt tile stops the stream with the store's own
error code; the plan scan(store) ... sink_zarr(path) writes the outputs that collect()
gives for the same data in memory."""
import json
import os

import numpy as np
import pytest

import kyto_amd
from kyto_amd import _native as N
from kyto_amd import store as ks
from kyto_amd import synthetic as S
from tests import store_format as sf

pytestmark = pytest.mark.gpu


def _counts(n=20_000, g=500, seed=5):
    c = S.counts(n, g, 6, 0.1, seed)
    return c.indptr, c.indices, np.asarray(c.values).astype(np.uint32), g


@pytest.fixture(scope="module")
def counts_store(tmp_path_factory):
    ip, ix, v, g = _counts()
    path = tmp_path_factory.mktemp("stores") / "counts.kyto"
    ks.create_store(path, n_cells=len(ip) - 1, n_genes=g, cell_block=2048, gene_block=128,
                    layers={"counts": ks.LayerData(ip, ix, v)}, nnz_chunk=4096, nnz_shard=65536)
    return path, (ip, ix, v, g)


@pytest.mark.parametrize("block_rows", [0, 3000])
def test_streamed_store_layer_equals_upload(device, counts_store, block_rows):
    path, (ip, ix, v, g) = counts_store
    want = device.upload_counts(ip, ix, v.astype(np.float32), g, target_sum=1e4, log1p=True)
    with ks.StoreLayer(path, "counts") as layer:
        got = device.stream_layer(layer, block_rows=block_rows, target_sum=1e4, log1p=True)
    try:
        assert got.stream_info["nnz"] == int(ip[-1])
        assert np.array_equal(got.values().view(np.uint32), want.values().view(np.uint32))
        a = device.pca(got, 20, seed=0, power_iterations=2)
        b = device.pca(want, 20, seed=0, power_iterations=2)
        assert np.array_equal(a.embedding.view(np.uint32), b.embedding.view(np.uint32))
        a.device_embedding.free()
        b.device_embedding.free()
    finally:
        got.free()
        want.free()


def test_corrupt_tile_stops_the_stream(device, tmp_path):
    ip, ix, v, g = _counts(6000, 200, seed=8)
    path = tmp_path / "bad.kyto"
    ks.create_store(path, n_cells=len(ip) - 1, n_genes=g, cell_block=1024, gene_block=64,
                    layers={"counts": ks.LayerData(ip, ix, v)}, nnz_chunk=4096, nnz_shard=65536)
    # Cell block 3's first tile: its indptr ends one short of tile_offsets.
    _, tp = sf.read_array(str(path / "layers" / "counts" / "indptr"))
    chunks = [tp[3, j].copy() for j in range(tp.shape[1])]
    chunks[0][-1] -= 1
    with open(path / "layers" / "counts" / "indptr" / "c" / "3" / "0" / "0", "wb") as f:
        f.write(sf._shard([c.tobytes() for c in chunks], 3))
    with ks.StoreLayer(path, "counts") as layer:
        with pytest.raises(N.KytoError, match="KY-E107.*tile \\(3, 0\\)"):
            device.stream_layer(layer, block_rows=1024, target_sum=1e4, log1p=True)
    # The context is usable afterwards.
    ok = device.upload_counts(ip, ix, v.astype(np.float32), g)
    ok.free()


def test_plan_from_store_to_store(device, counts_store, tmp_path):
    path, (ip, ix, v, g) = counts_store

    def plan(src):
        return (kyto_amd.scan(src).normalize_total(1e4).log1p().pca(n_components=20, seed=0)
                .neighbors(n_neighbors=10).on("gpu"))

    mem = plan({"counts": (ip, ix, v.astype(np.float32), g)}).collect()
    out = tmp_path / "out.kyto"
    plan(path).sink_zarr(out)
    schema = ks.read_schema(out)
    assert [m["name"] for m in schema["layers"]] == ["counts", "log1p"]
    assert [m["name"] for m in schema["embeddings"]] == ["X_pca", "X_pca_loadings"]
    assert [m["name"] for m in schema["graphs"]] == ["connectivities", "distances"]
    assert np.array_equal(ks.read_embedding(out, "X_pca").view(np.uint32), mem["X_pca"].view(np.uint32))
    assert np.array_equal(ks.read_embedding(out, "X_pca_loadings"), mem["X_pca_loadings"])
    gi, gx, gw = ks.read_graph(out, "distances")
    assert np.array_equal(gx, mem["distances"].indices) and np.array_equal(gw, mem["distances"].distances)
    ci, cx, cw = ks.read_graph(out, "connectivities")
    assert np.array_equal(ci, mem["connectivities"].indptr) and np.array_equal(cw, mem["connectivities"].weights)
    with ks.StoreLayer(out, "counts") as layer:
        cip, cix, cv = layer.read_rows()
    assert np.array_equal(cix, ix) and np.array_equal(cv, v.astype(np.float32))
    # The log1p layer the device made, as the device made it.
    with ks.StoreLayer(out, "log1p") as layer:
        _, lix, lv = layer.read_rows()
    made = device.upload_counts(ip, ix, v.astype(np.float32), g, target_sum=1e4, log1p=True)
    want = made.values()
    made.free()
    assert np.array_equal(lix, ix) and np.array_equal(lv.view(np.uint32), want.view(np.uint32))
    with open(out / "cells.parquet", "rb") as f:
        assert f.read(4) == b"PAR1"


def test_store_reader_as_a_block_reader(device, counts_store):
    # The C ABI a Rust caller binds: ky_store_layer_read passed to ky_csr_stream as its
    # ky_block_reader, the ky_layer as its user pointer.
    import ctypes
    from kyto_amd.device import CsrLayer
    path, (ip, ix, v, g) = counts_store
    lib = N.gpu_lib()
    reader = ctypes.cast(lib.ky_store_layer_read, N.BlockReader)
    want = device.upload_counts(ip, ix, v.astype(np.float32), g, target_sum=1e4, log1p=True)
    with ks.StoreLayer(path, "counts") as layer:
        h = ctypes.c_void_p()
        info = N.StreamInfo()
        N.check(lib.ky_csr_stream(device._h, layer.n_cells, layer.n_genes, 5000, reader, layer.handle, 0, 1e4, 1,
                                  ctypes.byref(h), ctypes.byref(info)))
    got = CsrLayer._adopt(device, h, len(ip) - 1, g, int(info.nnz))
    try:
        assert int(info.nnz) == int(ip[-1])
        assert np.array_equal(got.values().view(np.uint32), want.values().view(np.uint32))
    finally:
        got.free()
        want.free()
