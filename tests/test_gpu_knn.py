"""neighbors() on the device against the oracle: the graph must be kyto's exact
graph bit for bit (indices and f32 distances), whatever the candidate precision,
on the golden embeddings, clustered / far / tied / duplicated / tiny inputs, and
the argument errors."""
import numpy as np
import pytest

import oracle
from kyto_amd import _native as N
from tests.helpers import clustered, load, scanpy_rows

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp32", "bf16"]


def _exact(device, pts, n_neighbors, precision, n_components=None):
    m = device.upload_dense(pts)
    try:
        return device.neighbors(m, n_neighbors, n_components=n_components, precision=precision)
    finally:
        m.free()


def _assert_oracle_graph(g, pts, n_neighbors, metric="euclidean"):
    _, idx, dist = oracle.neighbors(pts, n_neighbors, metric=metric)
    got = g.rows()
    assert np.array_equal(got, idx)
    assert np.array_equal(g.distances.reshape(idx.shape).view(np.uint32), dist.view(np.uint32))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_pbmc3k_embedding_equals_oracle_and_scanpy(device, precision):
    emb = load("pbmc3k_pca.npz")["embedding"][:, :50].copy()
    g = _exact(device, emb, 15, precision)
    _assert_oracle_graph(g, emb, 15)
    ri, rd = scanpy_rows("pbmc3k", 2700)
    assert np.array_equal(g.rows(), ri)
    d = g.distances.reshape(2700, 14).astype(np.float64)
    assert (np.abs(d - rd) / np.maximum(rd, 1e-12)).max() <= 2e-7


@pytest.mark.parametrize("precision", PRECISIONS)
def test_synthetic_embedding_equals_oracle(device, precision):
    emb = load("synthetic_pca.npz")["embedding"].copy()
    g = _exact(device, emb, 15, precision)
    _assert_oracle_graph(g, emb, 15)
    # n_components: the leading 10 of the 20 columns.
    g10 = _exact(device, emb, 15, precision, n_components=10)
    _assert_oracle_graph(g10, np.ascontiguousarray(emb[:, :10]), 15)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_clusters_far_ties_and_duplicates(device, precision):
    cases = [
        (clustered(1, 3000, 50, 8, 0.0), 15),
        (clustered(2, 1500, 7, 8, 0.0), 6),
        (clustered(4, 2000, 20, 5, 1000.0), 15),
        (np.full((300, 3), 0.25, np.float32), 8),
    ]
    rng = np.random.default_rng(5)
    cases.append((np.floor(rng.random((6000, 2)) * 4).astype(np.float32), 15))  # 16 distinct rows
    cases.append((np.floor(rng.random((2000, 3)) * 6).astype(np.float32), 11))
    for pts, nn in cases:
        g = _exact(device, pts, nn, precision)
        _assert_oracle_graph(g, pts, nn)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_small_inputs(device, precision):
    for pts, nn in [
        (np.array([[0.0], [1.0]], np.float32), 2),
        (np.array([[0.0], [1.0], [3.0], [7.0]], np.float32), 3),
        (clustered(6, 16, 4, 2, 0.0), 16),
        (clustered(7, 1024 + 128 + 3, 9, 3, 0.0), 5),
        (clustered(8, 65, 64, 3, 0.0), 65),
    ]:
        g = _exact(device, pts, nn, precision)
        _assert_oracle_graph(g, pts, nn)
    g = _exact(device, np.array([[0.0], [1.0], [3.0], [7.0]], np.float32), 3, precision)
    assert list(g.indices) == [1, 2, 0, 2, 0, 1, 1, 2]
    assert list(g.distances) == [1.0, 3.0, 1.0, 2.0, 3.0, 2.0, 6.0, 4.0]


def test_huge_norms_searched_exactly(device):
    pts = clustered(9, 400, 3, 2, 0.0) * np.float32(1e20)
    g = _exact(device, pts, 6, "fp32")
    _assert_oracle_graph(g, pts, 6)


def test_refusals(device):
    pts = np.zeros((3, 2), np.float32)
    m = device.upload_dense(pts)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.neighbors(m, 1)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.neighbors(m, 4)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.neighbors(m, 2, n_components=3)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.neighbors(m, 2, metric="manhattan")
    bad = device.upload_dense(np.array([[0.0, np.nan], [1.0, 2.0]], np.float32))
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.neighbors(bad, 2)


@pytest.mark.parametrize("key", ["squared_euclidean", "cosine"])
def test_nearest_keys_equal_the_reference_kernels(device, key):
    # DeviceKernels::nearest_keys: a query matrix of its own, a skip per query, the
    # f32 keys of kernels_matrix.rs in component order: the oracle's bits, ranked by
    # (key, point).
    pts = clustered(3, 2100, 16, 9, 0.0)
    pts[7] = 0.0  # a zero row: cosine key 1
    rng = np.random.default_rng(4)
    queries = np.concatenate([pts[100:150], rng.standard_normal((20, 16)).astype(np.float32), pts[7:8]])
    skip = np.full(len(queries), 0xFFFFFFFF, np.uint32)
    skip[:50] = np.arange(100, 150)  # the queries that are rows of the points skip themselves
    qn = pn = None
    if key == "cosine":
        qn = np.sqrt((queries.astype(np.float64) ** 2).sum(axis=1)).astype(np.float32)
        pn = np.sqrt((pts.astype(np.float64) ** 2).sum(axis=1)).astype(np.float32)
    m = device.upload_dense(pts)
    try:
        keys, rows = device.nearest_keys(queries, m, 20, key=key, skip=skip, query_norms=qn, point_norms=pn)
        with pytest.raises(N.KytoError, match="KY-E215"):
            device.nearest_keys(queries[:2], m, 2100, skip=skip[:2])  # m above the points left after skip
        with pytest.raises(N.KytoError, match="KY-E215"):
            device.nearest_keys(queries[:2], m, 5, key="cosine")  # no norms
    finally:
        m.free()
    want_keys, want_rows = oracle.nearest_keys(queries, pts, 20, key=key, skip=skip, query_norms=qn,
                                               point_norms=pn)
    assert np.array_equal(rows, want_rows)
    assert np.array_equal(keys.view(np.uint32), want_keys.view(np.uint32))
    # The Euclidean keys are within f32 rounding of the exact distances.
    if key == "squared_euclidean":
        exact = ((queries[:, None, :].astype(np.float64) - pts[rows].astype(np.float64)) ** 2).sum(axis=2)
        assert np.abs(keys - exact).max() <= 1e-5 * (1 + exact.max())


def _large_clustered(n, d, clusters, seed):
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((clusters, d)).astype(np.float32) * 3
    lab = rng.integers(0, clusters, n)
    return (centres[lab] + rng.standard_normal((n, d)).astype(np.float32)).astype(np.float32)


def test_large_input_threshold_estimate(device):
    # Above 2^17 rows the BF16 scan starts every query at a sampled threshold: the graph
    # must still be the FP32 path's (kyto's exact graph), and the oracle's on sampled rows.
    pts = _large_clustered(140_000, 50, 12, seed=21)
    g16 = _exact(device, pts, 15, "bf16")
    g32 = _exact(device, pts, 15, "fp32")
    assert np.array_equal(g16.rows(), g32.rows())
    assert np.array_equal(g16.distances.view(np.uint32), g32.distances.view(np.uint32))
    first, count = 70_000, 300
    keys, rows, _ = oracle.knn_exact_range(pts, 14, first, count)
    o = np.argsort(rows, axis=1)
    want_rows = np.take_along_axis(rows, o, 1)
    want_dist = np.sqrt(np.take_along_axis(keys, o, 1)).astype(np.float32)
    got = g16.rows()[first:first + count]
    assert np.array_equal(got, want_rows)
    assert np.array_equal(g16.distances.reshape(-1, 14)[first:first + count], want_dist)


def test_scan_tail_split_over_point_segments(device):
    # Rounds of 782 query blocks (and 781) against 768 resident ones, so the last few
    # blocks' queries of each round are scanned as point segments and their lists
    # merged. The graph must be the one-round (symmetric) scan's on every row, and the
    # oracle's on sampled rows.
    pts = _large_clustered(400_000, 50, 16, seed=23)
    m = device.upload_dense(pts)
    try:
        whole = device.neighbors(m, 15, precision="bf16")
        g16 = device.neighbors(m, 15, precision="bf16", round_rows=782 * 256)
    finally:
        m.free()
    assert g16.info["rounds"] == 2
    assert np.array_equal(g16.rows(), whole.rows())
    assert np.array_equal(g16.distances.view(np.uint32), whole.distances.view(np.uint32))
    assert g16.info["fallback"] < 1000
    for first in (199_000, 399_600):
        keys, rows, _ = oracle.knn_exact_range(pts, 14, first, 400)
        o = np.argsort(rows, axis=1)
        assert np.array_equal(g16.rows()[first:first + 400], np.take_along_axis(rows, o, 1))


def test_short_rows_are_scanned_again_without_thresholds(device):
    # 2,000 copies of one row: each copy has 1,999 rows at key 0, far more than a row's
    # list of earlier rows holds in the symmetric scan, so most copies end their round
    # with fewer than m candidates; they are scanned again as a list from twice their
    # starting thresholds (point segments over the device), and the ties then go to the
    # exact search. A second group of 300 copies stays within the lists. The graph is the FP32
    # path's on every row and the oracle's on sampled rows, copies included.
    pts = _large_clustered(160_000, 50, 10, seed=41)
    pts[10_000:12_000] = pts[5]
    pts[90_000:90_300] = pts[7]
    m = device.upload_dense(pts)
    try:
        g16 = device.neighbors(m, 15, precision="bf16")
        g32 = device.neighbors(m, 15, precision="fp32")
    finally:
        m.free()
    assert g16.info["rounds"] == 1
    # (The earliest copies hold few earlier copies in their lists; the others overflow.)
    assert g16.info["rescanned"] >= 1_500
    assert np.array_equal(g16.rows(), g32.rows())
    assert np.array_equal(g16.distances.view(np.uint32), g32.distances.view(np.uint32))
    rows = np.unique(np.concatenate([np.arange(0, 160_000, 4001), np.arange(9_990, 10_020),
                                     np.arange(11_990, 12_010), np.arange(90_290, 90_310)])).astype(np.uint32)
    keys, want, _ = oracle.knn_exact_rows(pts, 14, rows)
    o = np.argsort(want, axis=1, kind="stable")
    assert np.array_equal(g16.rows()[rows], np.take_along_axis(want, o, 1))
    want_dist = np.sqrt(np.take_along_axis(keys, o, 1)).astype(np.float32)
    assert np.array_equal(g16.distances.reshape(-1, 14)[rows].view(np.uint32), want_dist.view(np.uint32))


def test_symmetric_scan_equals_the_exact_graph(device):
    # One round over every row (above 2^17 rows) keys each pair of rows once, for both
    # rows: each row's candidates join its own scan of the later rows with the keys the
    # earlier rows' blocks found at or below its threshold. 150,100 rows end in a partial
    # query block whose one tile is its own; duplicated rows tie keys across the join.
    # The graph is the FP32 path's on every row, the rounds' (the one-sided scan) on
    # every row, and the oracle's on sampled rows.
    pts = _large_clustered(150_100, 50, 12, seed=37)
    pts[1::1001] = pts[::1001][:len(pts[1::1001])]
    pts[149_990:] = pts[70_000:70_110]
    m = device.upload_dense(pts)
    try:
        whole = device.neighbors(m, 15, precision="bf16")
        rounds = device.neighbors(m, 15, precision="bf16", round_rows=76_800)
        g32 = device.neighbors(m, 15, precision="fp32")
    finally:
        m.free()
    assert whole.info["rounds"] == 1 and rounds.info["rounds"] == 2
    assert whole.info["fallback"] < 200
    for g in (rounds, g32):
        assert np.array_equal(g.rows(), whole.rows())
        assert np.array_equal(g.distances.view(np.uint32), whole.distances.view(np.uint32))
    rows = np.unique(np.concatenate([np.arange(0, 150_100, 1499), np.arange(149_980, 150_100)])).astype(np.uint32)
    keys, want, _ = oracle.knn_exact_rows(pts, 14, rows)
    o = np.argsort(want, axis=1, kind="stable")
    assert np.array_equal(whole.rows()[rows], np.take_along_axis(want, o, 1))
    want_dist = np.sqrt(np.take_along_axis(keys, o, 1)).astype(np.float32)
    assert np.array_equal(whole.distances.reshape(-1, 14)[rows].view(np.uint32), want_dist.view(np.uint32))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scale", [1e-21, 1e-12])
def test_tiny_embeddings_stay_exact(device, precision, scale):
    # Products of components near 1e-21 fall below f32's normal range: the BF16 and
    # FP32 key bounds carry an absolute term for it, so such rows go to the exact
    # search rather than being certified on flushed keys. The graph is kyto's exact
    # graph bit for bit.
    pts = (clustered(5, 1200, 20, 6, 0.0).astype(np.float64) * scale).astype(np.float32)
    g = _exact(device, pts, 15, precision)
    _assert_oracle_graph(g, pts, 15)


def _cosine_cases():
    rng = np.random.default_rng(11)
    pts = clustered(12, 2500, 30, 7, 0.0)
    with_zeros = pts.copy()
    with_zeros[::97] = 0.0                       # zero rows: key 1 to every row
    with_zeros[5::131] = with_zeros[4::131][:len(with_zeros[5::131])]  # duplicated rows
    return [
        (load("pbmc3k_pca.npz")["embedding"][:, :50].copy(), 15),
        (with_zeros, 15),
        ((pts.astype(np.float64) * 1e-20).astype(np.float32), 10),  # cosine ignores the scale
        (clustered(13, 1500, 8, 4, 50.0), 6),        # far from the origin: keys near 0
        (np.floor(rng.random((3000, 3)) * 3).astype(np.float32), 12),  # ties, zero rows
    ]


@pytest.mark.parametrize("precision", PRECISIONS)
def test_cosine_graph_equals_oracle(device, precision):
    # metric='cosine' (neighbors.rs:440-445): kyto's f64 key 1 - dot / (|q| |p|) in its
    # lane order, clipped to [0, 2], 1 for a zero row; distances f32(key).
    for pts, nn in _cosine_cases():
        m = device.upload_dense(pts)
        try:
            g = device.neighbors(m, nn, metric="cosine", precision=precision)
        finally:
            m.free()
        _assert_oracle_graph(g, pts, nn, metric="cosine")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_rounds_give_the_same_graph(device, precision):
    # The candidate search in rounds (a memory limit's smaller scratch): the same graph,
    # bit for bit, and the scratch the call reports is the one planned for it.
    pts = _large_clustered(140_000, 50, 12, seed=31)
    m = device.upload_dense(pts)
    try:
        whole = device.neighbors(m, 15, precision=precision)
        for round_rows in (50_000, 12_800):
            g = device.neighbors(m, 15, precision=precision, round_rows=round_rows)
            assert g.info["rounds"] == -(-140_000 // round_rows)
            assert np.array_equal(g.rows(), whole.rows())
            assert np.array_equal(g.distances.view(np.uint32), whole.distances.view(np.uint32))
            assert g.info["scratch_bytes"] == device.neighbors_bytes(140_000, 50, 15, precision=precision,
                                                                     round_rows=round_rows)
        assert whole.info["scratch_bytes"] > g.info["scratch_bytes"]
        cos_whole = device.neighbors(m, 15, metric="cosine", precision=precision)
        cos_rounds = device.neighbors(m, 15, metric="cosine", precision=precision, round_rows=30_000)
        assert np.array_equal(cos_rounds.rows(), cos_whole.rows())
        rows = np.arange(0, 140_000, 997, dtype=np.uint32)
        keys, want, _ = oracle.knn_exact_rows(pts, 14, rows, metric="cosine")
        o = np.argsort(want, axis=1)
        assert np.array_equal(cos_whole.rows()[rows], np.take_along_axis(want, o, 1))
    finally:
        m.free()
