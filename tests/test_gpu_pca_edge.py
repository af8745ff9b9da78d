"""pca() on the device at the edges the reference's GPU tests cover: large gene means
(crates/kyto-exec/tests/it/gpu_matrix.rs:612-650: offsets up to 1e6 against the CPU
path within pca.embedding), values outside the range f32 products keep exact (1e-20
and 1e30 refused with KY-E215, the CPU path accepting them), and pca over a gene set
(pca.rs tests: the bits of a gene filter before it, zero loadings outside the set)."""
import numpy as np
import pytest

import kyto_amd
import oracle
from kyto_amd import KytoError
from kyto_amd import synthetic as S
from tests.helpers import dense_to_csr, lcg_rows

pytestmark = pytest.mark.gpu

# tests/parity/tolerances.toml: pca.embedding, pca.loadings (each column relative to
# its largest value), pca.variance.
EMBEDDING = 1e-3


def _column_error(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float((np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-300)).max())


@pytest.mark.parametrize("offset", [0.0, 1e3, 1e5, 1e6])
def test_large_means_match_the_cpu_path(device, offset):
    rows = lcg_rows(200, 100, 4, 1.0, 6) + np.float32(offset)
    ip, ix, v = dense_to_csr(rows)
    csr = device.upload_csr(ip, ix, v, 100)
    try:
        res = device.pca(csr, 5, seed=0, power_iterations=9)
    finally:
        csr.free()
    res.device_embedding.free()
    ref = oracle.pca(ip, ix, v, 100, 5, seed=0, q=9)
    worst = max(float((np.abs(res.variance - ref.variance) / ref.variance).max()),
                _column_error(res.embedding, ref.embedding), _column_error(res.loadings, ref.loadings))
    print(f"offset {offset:g}: worst {worst:.1e} from the cpu path")
    assert worst <= EMBEDDING, (offset, worst)


@pytest.mark.parametrize("scale", [1e-20, 1e30])
def test_values_beyond_exact_f32_products_are_refused(device, scale):
    rows = (lcg_rows(100, 20, 3, 1.0, 7).astype(np.float64) * scale).astype(np.float32)
    ip, ix, v = dense_to_csr(rows)
    oracle.pca(ip, ix, v, 20, 3, seed=0, q=9)  # the CPU path takes them
    with pytest.raises(KytoError) as e:
        device.upload_csr(ip, ix, v, 20)
    assert e.value.code == "KY-E215" and "1e-10 to 1e10" in str(e.value)


def test_pca_over_a_gene_set_is_a_gene_filter_before_it(device):
    L = S.log1p_layer(6000, 300, 6, 0.08, seed=21)
    genes = np.zeros(300, bool)
    genes[np.random.default_rng(3).choice(300, 120, replace=False)] = True
    out = kyto_amd.scan((L.indptr, L.indices, L.values, 300)).pca(n_components=20, seed=0, genes=genes).collect()
    # The oracle on the layer's columns of the set (a gene filter before pca).
    pos = np.cumsum(genes) - 1
    keep = genes[L.indices]
    rows = np.repeat(np.arange(L.n_cells), np.diff(L.indptr.astype(np.int64)))[keep]
    ip = np.zeros(L.n_cells + 1, np.uint64)
    np.cumsum(np.bincount(rows, minlength=L.n_cells), out=ip[1:])
    ref = oracle.pca(ip, pos[L.indices[keep]].astype(np.uint32), L.values[keep], int(genes.sum()), 20, seed=0,
                     q=kyto_amd.PCA_POWER_ITERATIONS)
    assert (np.abs(out["X_pca_variance"] - ref.variance) / ref.variance).max() < 1e-4
    loadings = out["X_pca_loadings"]
    assert loadings.shape == (300, 20)
    assert not loadings[~genes].any()  # zero outside the set
    assert _column_error(loadings[genes][:, :10], ref.loadings[:, :10]) < 1e-3
    assert _column_error(out["X_pca"][:, :10], ref.embedding[:, :10]) < 1e-3
    # An empty set, and a set smaller than n_components, are KY-E215.
    with pytest.raises(KytoError) as e:
        kyto_amd.scan((L.indptr, L.indices, L.values, 300)).pca(n_components=5, seed=0, genes=np.zeros(300, bool))
    assert e.value.code == "KY-E215"
    few = np.zeros(300, bool)
    few[:4] = True
    with pytest.raises(KytoError) as e:
        kyto_amd.scan((L.indptr, L.indices, L.values, 300)).pca(n_components=5, seed=0, genes=few).collect()
    assert e.value.code == "KY-E215"
