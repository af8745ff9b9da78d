"""connectivities() on the device against the oracle (kyto's connectivities.rs
restated): the graph must equal the oracle's bit for bit (edges, float32 weights,
each row's sigma and rho) on scanpy's fixture graphs, the reference's own test
graphs, random graphs with ties, zero rows and self entries, every register width of
the kernel, a 200,000-row graph from the device's own search, and the refusals."""
import numpy as np
import pytest

import kyto_amd
import oracle
from kyto_amd import _native as N
from kyto_amd import synthetic as S
from kyto_amd.device import Graph
from tests.helpers import KNOWN_GRAPHS, known_graph, random_graph, scanpy_graph

pytestmark = pytest.mark.gpu


def _graph(idx, dist, n_neighbors):
    n = idx.shape[0]
    k = n_neighbors - 1
    return Graph(n_neighbors, np.arange(0, n * k + 1, k, dtype=np.uint64), np.ascontiguousarray(idx, np.uint32).ravel(),
                 np.ascontiguousarray(dist, np.float32).ravel())


def _assert_oracle(device, idx, dist, n_neighbors):
    got = device.connectivities(_graph(idx, dist, n_neighbors))
    ref = oracle.connectivities(idx, dist, n_neighbors)
    assert np.array_equal(got.sigma.view(np.uint32), ref.sigma.view(np.uint32))
    assert np.array_equal(got.rho.view(np.uint32), ref.rho.view(np.uint32))
    assert np.array_equal(got.indptr, ref.indptr)
    assert np.array_equal(got.indices, ref.indices)
    assert np.array_equal(got.weights.view(np.uint32), ref.weights.view(np.uint32))
    return got


@pytest.mark.parametrize("name", ["synthetic", "pbmc3k"])
def test_scanpy_fixture_graphs(device, name):
    idx, dist, (ip, ix, w) = scanpy_graph(name)
    got = _assert_oracle(device, idx, dist, 15)
    # scanpy's own weights: the same edges, within connectivities.weights = 1e-5
    # (tolerances.toml:173-178); the float32 exp of the fixture platform parts from the
    # correctly rounded one on 0.5% of the weights, by one or two units in the last place.
    assert np.array_equal(got.indptr, ip) and np.array_equal(got.indices, ix)
    assert np.abs(got.weights.astype(np.float64) - w).max() <= 2.5e-7
    assert (got.weights == w).mean() >= 0.99


@pytest.mark.parametrize("name", sorted(KNOWN_GRAPHS))
def test_reference_unit_graphs(device, name):
    _assert_oracle(device, *known_graph(name))


@pytest.mark.parametrize("k,ties", [(1, False), (4, True), (14, False), (14, True), (29, True), (33, False),
                                    (64, True)])
def test_random_graphs(device, k, ties):
    # Every register width of the row kernel (16, 32, 64 entries), ties, zeros and rows
    # of zero distances (the global floor, a host pass over every ranked distance).
    idx, dist = random_graph(1500, k, seed=k, ties=ties)
    got = _assert_oracle(device, idx, dist, k + 1)
    if ties:
        assert got.info["zero_rho_rows"] > 0


def test_self_entries_are_dropped(device):
    # A row naming itself (scanpy's graph can): membership 0 there.
    idx, dist = random_graph(400, 6, seed=9)
    idx[7] = np.sort(np.concatenate([[7], idx[7][:5]]))
    _assert_oracle(device, idx, dist, 7)


def test_device_search_graph(device):
    # The device's own kNN graph of a 200,000-row clustered embedding.
    rng = np.random.default_rng(3)
    centres = rng.normal(0, 4, size=(40, 30))
    pts = (centres[rng.integers(0, 40, 200_000)] + rng.normal(0, 1, size=(200_000, 30))).astype(np.float32)
    m = device.upload_dense(pts)
    try:
        g = device.neighbors(m, 15, precision="bf16")
    finally:
        m.free()
    got = _assert_oracle(device, g.rows(), g.distances.reshape(-1, 14), 15)
    # Symmetric: the transpose holds the same weights.
    rows = np.repeat(np.arange(got.n_cells), np.diff(got.indptr.astype(np.int64)))
    a = rows.astype(np.int64) * got.n_cells + got.indices
    b = got.indices.astype(np.int64) * got.n_cells + rows
    o = np.argsort(b)
    assert np.array_equal(a, b[o]) and np.array_equal(got.weights, got.weights[o])


def test_plan_writes_connectivities(device):
    layer = S.log1p_layer(3000, 200, 6, 0.1, seed=2)
    out = (kyto_amd.scan((layer.indptr, layer.indices, layer.values, 200)).pca(n_components=20, seed=0).neighbors(n_neighbors=10, into="knn")
           .on("gpu").collect())
    g = out["knn_distances"]
    c = out["knn_connectivities"]
    ref = oracle.connectivities(g.rows(), g.distances.reshape(-1, 9), 10)
    assert np.array_equal(c.indices, ref.indices)
    assert np.array_equal(c.weights.view(np.uint32), ref.weights.view(np.uint32))


def test_refusals(device):
    idx, dist, nn = known_graph("line")
    bad = idx.copy()
    bad[1] = [2, 0]  # not increasing
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.connectivities(_graph(bad, dist, nn))
    bad = idx.copy()
    bad[2, 1] = 3  # beyond the rows
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.connectivities(_graph(bad, dist, nn))
    for v in (-1.0, np.nan, np.inf):
        d = dist.copy()
        d[0, 1] = v
        with pytest.raises(N.KytoError, match="KY-E215"):
            device.connectivities(_graph(idx, d, nn))
    wide = np.tile(np.arange(1, 66, dtype=np.uint32), (70, 1))
    with pytest.raises(N.KytoError, match="KY-E501"):
        device.connectivities(_graph(wide, np.ones((70, 65), np.float32), 66))
