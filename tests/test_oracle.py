"""The oracle pinned against the reference's golden vectors, its own known-answer
tests and the mathematics (no GPU). Each pin is chosen so a plausible mistake (a
dropped centring term, a wrong lane order, a transposed operand, a lost tie rule)
fails one of them."""
import numpy as np
import pytest

import oracle
from tests.helpers import clustered, load, scanpy_rows, synthetic_hvg_csr


def test_log1p_layer_bit_exact_against_scanpy_fixture():
    # synthetic/normalize_log1p.npz: scanpy's normalize_total(1e4) output and its log1p.
    g = load("synthetic_log1p.npz")
    out = oracle.log1p32(g["normalized"])
    assert np.array_equal(out.view(np.uint32), g["data"].view(np.uint32))


def test_log1p_known_values():
    # kernels/log1p.rs:55-58 doc examples and special values.
    x = np.array([0.0, 1.0, -1.0, np.inf, 1e-45], np.float32)
    out = oracle.log1p32(x)
    assert out[0] == 0.0 and out[1] == np.float32(np.log(2.0)) and out[2] == -np.inf
    assert out[3] == np.inf and out[4] == np.float32(1e-45)


def test_omega_rows_are_uniform_and_depend_on_seed_and_cell():
    # pca.rs:957-970.
    row = oracle.omega_row(0, 7, 1000)
    assert np.all((row >= -1.0) & (row < 1.0))
    assert abs(row.mean()) < 0.1
    assert not np.array_equal(row, oracle.omega_row(0, 8, 1000))
    assert not np.array_equal(row, oracle.omega_row(1, 7, 1000))
    assert np.array_equal(row, oracle.omega_row(0, 7, 1000))


def test_omega_closed_form():
    # SplitMix64 restated in numpy uint64 arithmetic gives the same bits.
    def mix(v):
        v = np.uint64(v)
        v = (v ^ (v >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        v = (v ^ (v >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return v ^ (v >> np.uint64(31))

    phi = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        for seed, cell in [(0, 0), (3, 12345), (2**40 + 7, 2**33)]:
            state = mix(np.uint64(seed) ^ mix(np.uint64(cell) + phi))
            want = []
            for _ in range(5):
                state = state + phi
                want.append(float(mix(state) >> np.uint64(11)) * 2.0**-52 - 1.0)
            assert np.array_equal(oracle.omega_row(seed, cell, 5), np.array(want))


def test_pca_matches_scanpy_fixture_and_exact_truth():
    # synthetic/pca.npz: scanpy arpack, 20 components over the 100 HVGs. kyto measured
    # 1.5e-6 relative (pca_parity.rs:18-19); the tolerance file's bound is 1e-3.
    ip, ix, v, genes = synthetic_hvg_csr()
    ref = load("synthetic_pca.npz")
    for q in (4, 9):
        r = oracle.pca(ip, ix, v, genes, 20, seed=0, q=q)
        rel = np.abs(r.variance - ref["variance"]) / ref["variance"]
        assert rel.max() < 1e-5
        rrel = np.abs(r.variance_ratio - ref["variance_ratio"]) / ref["variance_ratio"]
        assert rrel.max() < 1e-5
        # Loadings: scanpy's svd_flip signs, so equal up to f32 noise column by column.
        comp = ref["components"].astype(np.float64)
        err = np.abs(r.loadings - comp).max(axis=0) / np.abs(comp).max(axis=0)
        assert err.max() < 1e-3
        emb = ref["embedding"].astype(np.float64)
        eerr = np.abs(r.embedding - emb).max(axis=0) / np.abs(emb).max(axis=0)
        assert eerr.max() < 1e-3
    # The basis spans all 100 genes after 4 passes of 30 columns: the exact PCA.
    tv, tvec, tratio = oracle.pca_truth(ip, ix, v, genes, 20)
    r = oracle.pca(ip, ix, v, genes, 20, seed=0, q=4)
    assert np.abs(r.eigenvalues - tv * (len(ip) - 2)).max() / tv.max() < 1e-10 * len(ip)
    assert np.abs(r.variance - tv).max() / tv.max() < 1e-12
    assert np.abs(np.abs(r.loadings64) - np.abs(tvec)).max() < 1e-8


def test_pca_without_row_centring_is_caught():
    # A planted mistake: the product pass without the mean term (P = X^T X Q) moves
    # the variances far beyond the tolerance, so the scanpy pin is not vacuous.
    ip, ix, v, genes = synthetic_hvg_csr()
    ref = load("synthetic_pca.npz")
    n = len(ip) - 1
    x = np.zeros((n, genes))
    rows = np.repeat(np.arange(n), np.diff(ip.astype(np.int64)))
    x[rows, ix.astype(np.int64)] = v
    vals = np.sort(np.linalg.eigvalsh(x.T @ x / (n - 1)))[::-1][:20]
    assert (np.abs(vals - ref["variance"]) / ref["variance"]).max() > 1e-2


def test_extend_basis_known_answers():
    # pca_dense.rs:319-334: only the second gene's direction is new.
    basis = np.array([[1.0], [0.0], [0.0]])
    block = np.array([[2.0, 4.0], [0.0, 1.0], [0.0, 0.0]])
    added = oracle.extend_basis(basis, block)
    assert added.shape[1] == 1 and abs(abs(added[1, 0]) - 1.0) < 1e-12
    assert oracle.extend_basis(np.hstack([basis, added]), block).shape[1] == 0
    assert oracle.extend_basis(basis, np.zeros((3, 2))).shape[1] == 0


def test_rayleigh_ritz_known_answers():
    # pca_dense.rs:336-355.
    vals, load_ = oracle.rayleigh_ritz(np.eye(3), np.diag([1.0, 3.0, 2.0]), 2)
    assert abs(vals[0] - 3) < 1e-12 and abs(vals[1] - 2) < 1e-12
    assert abs(load_[1, 0] - 1) < 1e-12 and abs(load_[2, 1] - 1) < 1e-12
    vals, load_ = oracle.rayleigh_ritz(np.array([[0.0], [-1.0], [0.0]]), np.array([[0.0], [-5.0], [0.0]]), 3)
    assert list(vals) == [5.0, 0.0, 0.0]
    assert np.abs(load_.T @ load_ - np.eye(3)).max() < 1e-12
    assert abs(load_[1, 0] - 1.0) < 1e-12


def test_nearest_keeps_the_k_smallest_ties_to_the_earlier_row():
    # neighbors.rs:555-561.
    got = oracle.nearest_offer_many(2, [3.0, 1.0, 1.0, 0.5, 1.0], [0, 5, 2, 9, 1])
    assert got == [(0.5, 9), (1.0, 1)]


def test_a_line_of_points():
    # neighbors.rs:563-576.
    indptr, idx, dist = oracle.neighbors(np.array([[0.0], [1.0], [3.0], [7.0]], np.float32), 3)
    assert list(indptr) == [0, 2, 4, 6, 8]
    assert list(idx.ravel()) == [1, 2, 0, 2, 0, 1, 1, 2]
    assert list(dist.ravel()) == [1.0, 3.0, 1.0, 2.0, 3.0, 2.0, 6.0, 4.0]


def test_exact_key_lane_order():
    # kernels.rs:422-448: component j in lane j % 8, the fixed tree. A left-to-right
    # sum gives other bits on some rows; the lane restatement equals numpy's
    # evaluation of the same tree.
    rng = np.random.default_rng(1)
    a = rng.standard_normal((200, 50)).astype(np.float32) * 3
    b = rng.standard_normal((200, 50)).astype(np.float32)
    differs = 0
    for i in range(200):
        d = a[i].astype(np.float64) - b[i].astype(np.float64)
        lanes = [0.0] * 8
        for j in range(50):
            lanes[j % 8] += d[j] * d[j]
        want = ((lanes[0] + lanes[4]) + (lanes[2] + lanes[6])) + ((lanes[1] + lanes[5]) + (lanes[3] + lanes[7]))
        got = oracle.squared_distance(a[i], b[i])
        assert got == want
        seq = 0.0
        for j in range(50):
            seq += d[j] * d[j]
        differs += seq != got
    assert differs > 0


@pytest.mark.parametrize("name,ncomp", [("pbmc3k", 50), ("synthetic", 20)])
def test_neighbors_against_scanpy_fixture(name, ncomp):
    # pbmc3k: every row equal, distances within 2e-7 (tolerances.toml neighbors.distance);
    # synthetic: 490 of 500 rows, the other 10 at duplicate cells (deviations.md:46).
    emb = load(f"{name}_pca.npz")["embedding"][:, :ncomp]
    n = emb.shape[0]
    indptr, idx, dist = oracle.neighbors(emb, 15)
    ref_idx, ref_dist = scanpy_rows(name, n)
    equal = (idx == ref_idx).all(axis=1)
    rel = np.abs(dist[equal] - ref_dist[equal]) / np.maximum(ref_dist[equal], 1e-12)
    assert rel.max() <= 2e-7
    if name == "pbmc3k":
        assert equal.all()
    else:
        assert equal.sum() == 490
        # Every difference is at the fixture's duplicate cells (4 identical rows):
        # the row of a duplicate (scanpy keeps a self loop, kyto excludes the cell by
        # position) or a tie among the duplicates (kyto takes the earlier row).
        _, inv, cnt = np.unique(emb, axis=0, return_inverse=True, return_counts=True)
        dups = set(np.where(cnt[inv] > 1)[0].tolist())
        assert len(dups) == 4
        for q in np.where(~equal)[0]:
            diff = set(ref_idx[q].tolist()) ^ set(idx[q].tolist())
            assert q in dups or diff <= dups


def _assert_filtered_equals_brute(points, k):
    keys, rows, _ = oracle.knn_exact(points, k)
    bk, br = oracle.search_rows(points, k, np.arange(points.shape[0]))
    assert np.array_equal(rows, br)
    assert np.array_equal(keys.view(np.uint64), bk.view(np.uint64))


def test_filtered_equals_brute_force_on_clusters_far_and_tied():
    # neighbors_exact.rs:417-460.
    _assert_filtered_equals_brute(clustered(2, 1500, 7, 8, 0.0), 5)
    _assert_filtered_equals_brute(clustered(4, 800, 20, 5, 1000.0), 14)
    rng = np.random.default_rng(5)
    tied = np.floor(rng.random((3000, 2)) * 4).astype(np.float32)
    _assert_filtered_equals_brute(tied, 14)
    _assert_filtered_equals_brute(np.full((300, 3), 0.25, np.float32), 7)


def test_huge_norms_use_brute_force():
    pts = clustered(9, 400, 3, 2, 0.0) * np.float32(1e20)
    keys, rows, fb = oracle.knn_exact(pts, 5)
    assert fb == 400
    bk, br = oracle.search_rows(pts, 5, np.arange(400))
    assert np.array_equal(rows, br)


def test_round_up_never_rounds_down():
    # neighbors_exact.rs:463-469.
    lib = oracle.lib()
    assert lib.ko_round_up(1.0) == 1.0
    assert lib.ko_round_up(1.0 + 1e-12) >= 1.0 + 1e-12
    assert lib.ko_round_up(float("inf")) == float("inf")


# ----------------------------------------------------------- connectivities ----

from tests.helpers import dense_union, known_graph, random_graph, scanpy_graph  # noqa: E402


@pytest.mark.parametrize("name", ["synthetic", "pbmc3k"])
def test_connectivities_against_scanpy_fixture(name):
    # connectivities_parity.rs: scanpy's own distances in, the same edges out, weights
    # within connectivities.weights = 1e-5 (tolerances.toml:173-178). Measured: every
    # edge equal; weights equal bit for bit but for 46 of 9,762 / 370 of 62,096, one or
    # two units in the last place, each at an exp(x) within 0.01 units of a rounding
    # midpoint (the fixture platform's float32 exp and the correctly rounded one part).
    idx, dist, (ip, ix, w) = scanpy_graph(name)
    c = oracle.connectivities(idx, dist, 15)
    assert np.array_equal(c.indptr, ip)
    assert np.array_equal(c.indices, ix)
    diff = np.abs(c.weights.astype(np.float64) - w)
    assert diff.max() <= 1e-5
    assert diff.max() <= 2.5e-7  # two units in the last place at 1.0
    assert (c.weights == w).mean() >= 0.99


def test_sigma_known_answers():
    # connectivities.rs:319-350: memberships sum to log2(k); the floor's rounding
    # (umap-learn 0.5.12 gives sigma bits 983688805).
    d = np.array([0.5, 0.9, 1.4, 2.0, 2.7], np.float32)
    sigma, rho = oracle.smooth_knn(d, 6)
    assert rho == np.float32(0.5)
    m = np.exp(-(np.maximum(d - rho, 0) / np.float64(sigma)))
    assert abs(m.sum() - np.log2(6)) < 1e-4
    r = np.array([0, 0, 0, 0, 2.4006681, 2.417103, 2.4605973, 2.4799116, 2.5937085], np.float32)
    sigma, rho = oracle.smooth_knn(r, 10)
    assert rho == np.float32(2.4006681)
    assert int(np.float32(sigma).view(np.uint32)) == 983688805


def _weight(c, r, col):
    s, e = int(c.indptr[r]), int(c.indptr[r + 1])
    j = np.searchsorted(c.indices[s:e], col)
    assert c.indices[s + j] == col
    return c.weights[s + j]


def test_connectivities_known_graphs():
    # connectivities.rs:266-397.
    c = oracle.connectivities(*known_graph("line"))
    assert c.indptr.tolist() == [0, 2, 4, 6] and c.indices.tolist() == [1, 2, 0, 2, 0, 1]
    for r in range(3):
        for col in range(3):
            if r != col:
                assert _weight(c, r, col) == _weight(c, col, r)
    assert _weight(c, 0, 1) == 1.0
    assert _weight(c, 1, 2) >= 1.0 - np.finfo(np.float32).eps
    assert 0.0 < _weight(c, 0, 2) < 1.0
    c = oracle.connectivities(*known_graph("duplicates"))
    assert c.indptr.tolist() == [0, 2, 4, 6] and c.indices.tolist() == [1, 2, 0, 2, 0, 1]
    assert c.weights[0] == 1.0 and np.all((c.weights > 0) & (c.weights <= 1))
    c = oracle.connectivities(*known_graph("same"))
    assert c.weights.tolist() == [1.0] * 6
    c = oracle.connectivities(*known_graph("far"))
    assert c.indptr.tolist() == [0, 2, 5, 8, 10]
    assert c.indices.tolist() == [1, 2, 0, 2, 3, 0, 1, 3, 1, 2]


@pytest.mark.parametrize("ties", [False, True])
def test_connectivities_union_equals_the_dense_formula(ties):
    # The sparse merge against (W + W^T) - W o W^T on a dense float32 matrix, with the
    # memberships restated in numpy from the oracle's sigma and rho (f32 differences
    # and quotients, exp rounded once from f64).
    n, k = 300, 9
    idx, dist = random_graph(n, k, seed=5, ties=ties)
    c = oracle.connectivities(idx, dist, k + 1)
    sig, rho = c.sigma[:, None], c.rho[:, None]
    diff = (dist - rho).astype(np.float32)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        x = (-(diff / sig)).astype(np.float32)
        w = np.where((diff <= 0) | (sig == 0), np.float32(1), np.exp(x.astype(np.float64)).astype(np.float32))
    w = np.where(idx == np.arange(n)[:, None], np.float32(0), w).astype(np.float32)
    U = dense_union(idx, w, n)
    nz = np.nonzero(U)
    assert np.array_equal(c.indptr, np.concatenate([[0], np.cumsum(np.bincount(nz[0], minlength=n))]))
    assert np.array_equal(c.indices, nz[1].astype(np.uint32))
    assert np.array_equal(c.weights.view(np.uint32), U[nz].view(np.uint32))
    # Rows not floored: the memberships of the bisected sigma sum to log2(k + 1).
    target = np.log2(k + 1)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        s = np.where(diff > 0, np.exp(-(diff.astype(np.float64) / c.sigma[:, None].astype(np.float64))), 1.0).sum(1)
    floor = 1e-3 * dist.sum(axis=1, dtype=np.float32).astype(np.float64) / (k + 1)
    free = (c.rho > 0) & (c.sigma.astype(np.float64) > floor * (1 + 1e-6))
    assert free.sum() > n // 2
    assert np.abs(s[free] - target).max() < 1e-4


def _block_order_tree(leaves):
    """reduction.rs:30-78 as a binary counter: push each leaf, combining two subtrees
    of one level as soon as the right one is complete; finish from the right."""
    stack = []
    for leaf in leaves:
        level, right = 0, leaf
        while stack and stack[-1][0] == level:
            _, left = stack.pop()
            right = left + right
            level += 1
        stack.append((level, right))
    _, right = stack.pop()
    while stack:
        _, left = stack.pop()
        right = left + right
    return right


def test_parallel_fold_is_the_leaves_in_kytos_tree():
    # Leaves of 500 rows (13 of them, a ragged last one), each folded in cell order by
    # the pinned one-leaf fold, combined in the binary-counter tree: bit for bit.
    from kyto_amd import synthetic as S
    L = S.log1p_layer(6100, 80, 4, 0.1, seed=5)
    n, g, w, leaf = L.n_cells, 80, 7, 500
    rng = np.random.default_rng(2)
    q = rng.standard_normal((g, w))
    bm = rng.standard_normal(w)
    lib, P = oracle.lib(), oracle._p
    leaves = []
    for r0 in range(0, n, leaf):
        r1 = min(r0 + leaf, n)
        prod, sums = np.zeros((g, w)), np.zeros(w)
        ip = np.ascontiguousarray(L.indptr[r0:r1 + 1])
        lib.ko_pca_product(P(ip, oracle._u64p), P(L.indices, oracle._u32p), P(L.values, oracle._f32p), r1 - r0,
                           P(q, oracle._f64p), w, P(bm, oracle._f64p), P(prod, oracle._f64p), P(sums, oracle._f64p))
        leaves.append(np.concatenate([prod.ravel(), sums]))
    want = _block_order_tree(leaves)
    prod, sums = np.zeros((g, w)), np.zeros(w)
    lib.ko_pca_product_tree(P(L.indptr, oracle._u64p), P(L.indices, oracle._u32p), P(L.values, oracle._f32p), n,
                            P(q, oracle._f64p), w, P(bm, oracle._f64p), g, leaf, P(prod, oracle._f64p),
                            P(sums, oracle._f64p))
    got = np.concatenate([prod.ravel(), sums])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    # A left-to-right combine (a wrong tree) differs in the last bits somewhere.
    seq = leaves[0].copy()
    for lf in leaves[1:]:
        seq = seq + lf
    assert not np.array_equal(seq.view(np.uint64), want.view(np.uint64))
    # The whole PCA through the tree stays within rounding of the one-leaf fold.
    a = oracle.pca(L.indptr, L.indices, L.values, g, 10, seed=0, q=4, leaf_rows=leaf)
    b = oracle.pca(L.indptr, L.indices, L.values, g, 10, seed=0, q=4, leaf_rows=None)
    assert np.abs(a.variance - b.variance).max() / b.variance.max() < 1e-12


def test_sparse_truth_is_the_dense_truth():
    # ko_gram's X^T X from the CSR against the dense product; the covariance's
    # eigenpairs against pca_truth's (which densifies the rows).
    from kyto_amd import synthetic as S
    L = S.log1p_layer(3000, 60, 4, 0.1, seed=8)
    x = np.zeros((L.n_cells, 60))
    x[np.repeat(np.arange(L.n_cells), np.diff(L.indptr.astype(np.int64))), L.indices.astype(np.int64)] = L.values
    v1, l1, r1 = oracle.pca_truth(L.indptr, L.indices, L.values, 60, 12)
    v2, l2, r2 = oracle.pca_truth_sparse(L.indptr, L.indices, L.values, 60, 12, parts=5)
    assert np.abs(v1 - v2).max() / v1.max() < 1e-12
    assert np.abs(r1 - r2).max() < 1e-12
    assert oracle.subspace_angles(l1[:, :5], l2[:, :5]).max() < 1e-6  # arccos near 1: ~sqrt(eps)
    gram = np.zeros((60, 60))
    sums = np.zeros(60)
    lib, P = oracle.lib(), oracle._p
    lib.ko_gram(P(L.indptr, oracle._u64p), P(L.indices, oracle._u32p), P(L.values, oracle._f32p), L.n_cells, 60, 3,
                P(gram, oracle._f64p), P(sums, oracle._f64p))
    assert np.abs(gram - x.T @ x).max() / np.abs(x.T @ x).max() < 1e-13
    assert np.allclose(sums, x.sum(axis=0), rtol=1e-13)


def test_listed_queries_equal_the_full_search():
    pts = clustered(31, 1500, 12, 9, 0.0)
    keys, rows, _ = oracle.knn_exact(pts, 14)
    qs = np.array([1499, 3, 700, 0, 42, 42, 1000], np.uint32)
    k2, r2, _ = oracle.knn_exact_rows(pts, 14, qs)
    assert np.array_equal(r2, rows[qs]) and np.array_equal(k2.view(np.uint64), keys[qs].view(np.uint64))


def test_nearest_keys_known_answers():
    # kernels_matrix.rs:140-257 on points whose keys are exact: a line (squared
    # distances, ties to the earlier point, the skipped point left out) and cosine
    # keys of axis vectors (0 same direction, 1 orthogonal or a zero row, 2 opposite).
    pts = np.array([[0.0], [1.0], [3.0], [7.0], [1.0]], np.float32)
    keys, rows = oracle.nearest_keys(np.array([[1.0], [5.0]], np.float32), pts, 3, skip=[1, 0xFFFFFFFF])
    assert rows.tolist() == [[4, 0, 2], [2, 3, 1]]
    assert keys.tolist() == [[0.0, 1.0, 4.0], [4.0, 4.0, 16.0]]
    axes = np.array([[1, 0], [0, 1], [-1, 0], [0, 0], [2, 0]], np.float32)
    norms = np.sqrt((axes ** 2).sum(axis=1)).astype(np.float32)
    keys, rows = oracle.nearest_keys(axes[:1], axes, 5, key="cosine", query_norms=norms[:1], point_norms=norms)
    assert rows.tolist() == [[0, 4, 1, 3, 2]] and keys.tolist() == [[0.0, 0.0, 1.0, 1.0, 2.0]]
