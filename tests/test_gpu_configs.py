"""The small configurations end to end against the oracle: the PBMC-3k-shaped layer
(2,700 x 2,000, 8 types, ~10%) and the 100,000 x 2,000 layer (12 types, ~8%), both
from the counts CSR: normalize_total(1e4) + log1p on the device, pca(50) with block
60 and 4 power iterations, then neighbors(15) at FP32 and BF16 candidate precision.
PCA within the north star's tolerance of the oracle's; the neighbour graph of the
device's embedding equal, index for index and bit for bit, to the oracle's exact graph
of that same embedding; its connectivities equal to the oracle's."""
import numpy as np
import pytest

import oracle
from kyto_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["pbmc3k_shaped", "100k"])
def test_config_counts_to_graph(device, name):
    w = S.WORKLOADS[name]
    c = S.counts(w.n_cells, w.n_genes, w.n_types, w.density, w.seed)
    layer = device.upload_counts(c.indptr, c.indices, c.values, w.n_genes, target_sum=1e4, log1p=True)
    try:
        res = device.pca(layer, 50, seed=0, power_iterations=4)
    finally:
        layer.free()
    _, values = oracle.normalize_log1p(c.indptr, c.values, 1e4)
    ref = oracle.pca(c.indptr, c.indices, values, w.n_genes, 50, seed=0, q=4)
    # PCA: explained variance within 1e-4 relative, components up to sign.
    assert (np.abs(res.variance - ref.variance) / ref.variance).max() < 1e-4
    assert (np.abs(res.variance_ratio - ref.variance_ratio) / ref.variance_ratio).max() < 1e-4
    lead = 10
    err = np.abs(res.loadings[:, :lead] - ref.loadings[:, :lead]).max(axis=0)
    assert (err / np.abs(ref.loadings[:, :lead]).max(axis=0)).max() < 1e-3
    s = np.linalg.svd(res.loadings.astype(np.float64).T @ ref.loadings64, compute_uv=False)
    assert np.arccos(np.clip(s.min(), -1, 1)) < 1e-2
    # kNN of the device's embedding: the oracle's exact graph of the same embedding.
    emb = res.embedding
    _, want_idx, want_dist = oracle.neighbors(emb, 15)
    for precision in ("fp32", "bf16"):
        g = device.neighbors(res.device_embedding, 15, precision=precision)
        assert np.array_equal(g.rows(), want_idx), precision
        assert np.array_equal(g.distances.reshape(want_idx.shape).view(np.uint32), want_dist.view(np.uint32))
    c_dev = device.connectivities(g)
    c_ref = oracle.connectivities(want_idx, want_dist, 15)
    assert np.array_equal(c_dev.indptr, c_ref.indptr) and np.array_equal(c_dev.indices, c_ref.indices)
    assert np.array_equal(c_dev.weights.view(np.uint32), c_ref.weights.view(np.uint32))
    res.device_embedding.free()
