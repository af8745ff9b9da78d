"""The input layer made on the device (ky_csr_upload_counts): kyto's normalize_total
then log1p, bit for bit the oracle's (kyto's CPU arithmetic) and, for log1p, scanpy's
fixture values; the plan API's normalize_total().log1p().pca() equal to PCA on the
host-made layer."""
import numpy as np
import pytest

import kyto_amd
import oracle
from kyto_amd import synthetic as S
from tests.helpers import load

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_normalize_log1p_equals_oracle(device):
    c = S.counts(6000, 700, 6, 0.1, seed=3)
    # An empty row (divisor 0 -> 1) and a row of one value.
    ip = np.concatenate([c.indptr[:101], [c.indptr[100]], c.indptr[101:]]).astype(np.uint64)
    norm, log = oracle.normalize_log1p(ip, c.values)
    layer = device.upload_counts(ip, c.indices, c.values, 700, target_sum=1e4, log1p=True)
    try:
        assert np.array_equal(_bits(layer.values()), _bits(log))
    finally:
        layer.free()
    only_norm = device.upload_counts(ip, c.indices, c.values, 700, target_sum=1e4, log1p=False)
    try:
        assert np.array_equal(_bits(only_norm.values()), _bits(norm))
    finally:
        only_norm.free()


def test_log1p_equals_scanpy_fixture(device):
    g = load("synthetic_log1p.npz")
    ip = g["indptr"].astype(np.uint64)
    layer = device.upload_counts(ip, g["indices"], g["normalized"], int(g["shape"][1]), target_sum=None, log1p=True)
    try:
        assert np.array_equal(_bits(layer.values()), _bits(g["data"]))
    finally:
        layer.free()


def test_plan_normalize_log1p_pca_equals_host_layer():
    c = S.counts(8000, 400, 5, 0.08, seed=4)
    host = S.Layer(c.indptr.copy(), c.indices.copy(), c.values.copy(), c.n_genes, c.mean_scale)
    S.normalize_log1p(host)
    plan = kyto_amd.scan((c.indptr, c.indices, c.values, c.n_genes)).normalize_total(1e4).log1p()
    assert "NORMALIZE_TOTAL" in plan.explain() and 'LOG1P "normalized" -> "log1p"' in plan.explain()
    a = plan.pca(n_components=20, seed=0).on("gpu").collect()
    b = kyto_amd.scan((host.indptr, host.indices, host.values, host.n_genes)).pca(n_components=20, seed=0) \
        .on("gpu").collect()
    assert np.array_equal(a["X_pca"], b["X_pca"])
    assert np.array_equal(a["X_pca_variance"], b["X_pca_variance"])


def test_streamed_blocks_equal_one_upload(device):
    # 70,000 cells in blocks of 32,768 (two full, one partial); dense enough that the
    # first pinned capacity is too small, and a tiny nnz hint so the resident arrays
    # grow while blocks arrive.
    n, g = 70_000, 300
    c = S.counts(n, g, 5, 0.2, seed=6)
    read = S.block_reader(n, g, 5, 0.2, seed=6)
    streamed = device.stream_counts(n, g, read, block_rows=32_768, nnz_hint=10, target_sum=1e4, log1p=True)
    whole = device.upload_counts(c.indptr, c.indices, c.values, g, target_sum=1e4, log1p=True)
    try:
        assert streamed.stream_info["blocks"] == 3 and streamed.nnz == c.nnz
        assert np.array_equal(_bits(streamed.values()), _bits(whole.values()))
        a = device.pca(streamed, 20, seed=0, power_iterations=4)
        b = device.pca(whole, 20, seed=0, power_iterations=4)
        assert np.array_equal(a.embedding, b.embedding)
        a.device_embedding.free()
        b.device_embedding.free()
    finally:
        streamed.free()
        whole.free()


def test_stream_reader_failure_is_e505(device):
    from kyto_amd import _native as N

    def bad(first, count, indptr, indices, values, capacity):
        raise RuntimeError("store unreadable")

    with pytest.raises(N.KytoError, match="KY-E505|505"):
        device.stream_counts(1000, 10, bad, block_rows=256)
