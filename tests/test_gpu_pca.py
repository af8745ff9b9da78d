"""pca() on the device against the oracle (kyto's CPU path restated), on the same
seeded inputs: each pass, the whole PCA, the scanpy fixture, determinism and the
error codes."""
import numpy as np
import pytest

import oracle
from kyto_amd import _native as N
from kyto_amd import synthetic as S
from tests.helpers import load, synthetic_hvg_csr

pytestmark = pytest.mark.gpu


def _layer(n=20000, genes=300, types=6, density=0.08, seed=11):
    L = S.log1p_layer(n, genes, types, density, seed)
    return L.indptr, L.indices, L.values, genes


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_sketch_pass_matches_oracle(device):
    ip, ix, v, g = _layer()
    csr = device.upload_csr(ip, ix, v, g)
    l = 40
    sketch, sums, osum = device.pca_sketch(csr, seed=5, l=l)
    n = len(ip) - 1
    cells = np.arange(n, dtype=np.uint64)
    ref = dict(n=np.zeros(1, np.uint64), sums=np.zeros(g), osum=np.zeros(l), sketch=np.zeros((g, l)))
    L = oracle.lib()
    P = oracle._p
    L.ko_pca_sketch(P(ip, oracle._u64p), P(ix, oracle._u32p), P(v, oracle._f32p), n, P(cells, oracle._u64p), 5, l,
                    P(ref["n"], oracle._u64p), P(ref["sums"], oracle._f64p), P(ref["osum"], oracle._f64p),
                    P(ref["sketch"], oracle._f64p))
    # Gene sums: exact products, f64 sums in another order.
    assert _rel(sums, ref["sums"]) < 1e-13
    # Omega's rows drawn on the device are the oracle's, rounded to f32.
    om32 = np.stack([oracle.omega_row(5, r, l) for r in range(n)]).astype(np.float32).astype(np.float64)
    assert _rel(osum, om32.sum(axis=0)) < 1e-12
    x = np.zeros((n, g))
    x[np.repeat(np.arange(n), np.diff(ip.astype(np.int64))), ix.astype(np.int64)] = v
    assert _rel(sketch, x.T @ om32) < 1e-12
    # Against the oracle's f64 Omega: within the f32 rounding of Omega.
    assert _rel(sketch, ref["sketch"]) < 1e-6


def test_product_and_project_passes_match_oracle(device):
    ip, ix, v, g = _layer(seed=12)
    n = len(ip) - 1
    csr = device.upload_csr(ip, ix, v, g)
    rng = np.random.default_rng(0)
    q = np.linalg.qr(rng.standard_normal((g, 60)))[0]
    means = rng.random(g) * 0.5
    bm = means @ q
    prod, sums = device.pca_product(csr, q, bm)
    rp, rs = np.zeros((g, 60)), np.zeros(60)
    L, P = oracle.lib(), oracle._p
    L.ko_pca_product(P(ip, oracle._u64p), P(ix, oracle._u32p), P(v, oracle._f32p), n, P(q, oracle._f64p), 60,
                     P(bm, oracle._f64p), P(rp, oracle._f64p), P(rs, oracle._f64p))
    # The device multiplies by Q rounded to f32 and rounds each w to f32.
    assert _rel(prod, rp) < 2e-6
    assert _rel(sums, rs) < 2e-6
    k = 50
    vv = np.linalg.qr(rng.standard_normal((g, k)))[0]
    mp = means @ vv
    emb, sq, st = device.pca_project(csr, vv, mp, means)
    rr, rsq, rst = np.zeros((n, k)), np.zeros(g), np.zeros(g, np.uint64)
    L.ko_pca_rows(P(ip, oracle._u64p), P(ix, oracle._u32p), P(v, oracle._f32p), n, P(vv, oracle._f64p), k,
                  P(mp, oracle._f64p), P(means, oracle._f64p), P(rr, oracle._f64p), P(rsq, oracle._f64p),
                  P(rst, oracle._u64p))
    assert _rel(emb, rr) < 1e-5
    assert _rel(sq, rsq) < 1e-12
    assert np.array_equal(st, rst)


@pytest.mark.parametrize("q", [4, 9])
def test_pca_matches_oracle(device, q):
    ip, ix, v, g = _layer(n=30000, genes=400, types=8, seed=13)
    csr = device.upload_csr(ip, ix, v, g)
    res = device.pca(csr, 50, seed=0, power_iterations=q)
    ref = oracle.pca(ip, ix, v, g, 50, seed=0, q=q)
    # north_star: explained variance within 1e-4 relative, components up to sign.
    assert (np.abs(res.variance - ref.variance) / ref.variance).max() < 1e-4
    assert (np.abs(res.variance_ratio - ref.variance_ratio) / ref.variance_ratio).max() < 1e-4
    assert res.passes == ref.passes
    assert _rel(res.means, ref.means) < 1e-13
    # Leading components (gap-separated): loadings and embedding columns, signs flipped alike.
    lead = 10
    err = np.abs(res.loadings[:, :lead] - ref.loadings[:, :lead]).max(axis=0) / np.abs(ref.loadings[:, :lead]).max(axis=0)
    assert err.max() < 1e-3
    e = res.embedding[:, :lead].astype(np.float64)
    r = ref.embedding[:, :lead].astype(np.float64)
    assert (np.abs(e - r).max(axis=0) / np.abs(r).max(axis=0)).max() < 1e-3
    # Subspace of all 50 components: principal angles small.
    s = np.linalg.svd(res.loadings.astype(np.float64).T @ ref.loadings64, compute_uv=False)
    assert np.arccos(np.clip(s.min(), -1, 1)) < 1e-2


def test_pca_on_scanpy_fixture(device):
    ip, ix, v, genes = synthetic_hvg_csr()
    ref = load("synthetic_pca.npz")
    csr = device.upload_csr(ip, ix, v, genes)
    res = device.pca(csr, 20, seed=0, power_iterations=9)
    assert (np.abs(res.variance - ref["variance"]) / ref["variance"]).max() < 1e-5
    comp = ref["components"].astype(np.float64)
    assert (np.abs(res.loadings - comp).max(axis=0) / np.abs(comp).max(axis=0)).max() < 1e-3
    orc = oracle.pca(ip, ix, v, genes, 20, seed=0, q=9)
    assert (np.abs(res.variance - orc.variance) / orc.variance).max() < 1e-6


def test_pca_is_deterministic(device):
    ip, ix, v, g = _layer(n=20000, genes=300, seed=14)
    csr = device.upload_csr(ip, ix, v, g)
    a = device.pca(csr, 30, seed=1, power_iterations=4)
    b = device.pca(csr, 30, seed=1, power_iterations=4)
    assert np.array_equal(a.variance, b.variance)
    assert np.array_equal(a.embedding, b.embedding)
    c = device.pca(csr, 30, seed=2, power_iterations=4)
    assert (np.abs(c.variance - a.variance) / a.variance)[:10].max() < 1e-3  # another seed, same PCA


def test_pca_refusals(device):
    ip = np.array([0, 2, 3], np.uint64)
    ix = np.array([0, 1, 1], np.uint32)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.upload_csr(ip, ix, np.array([1.0, 1e-20, 2.0], np.float32), 2)  # outside 1e-10..1e10
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.upload_csr(ip, ix, np.array([1.0, np.nan, 2.0], np.float32), 2)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.upload_csr(ip, np.array([1, 0, 1], np.uint32), np.ones(3, np.float32), 2)  # not increasing
    csr = device.upload_csr(ip, ix, np.array([1.0, 3.0, 2.0], np.float32), 2)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.pca(csr, 3, seed=0)  # above min(cells, genes)
    one = device.upload_csr(np.array([0, 1], np.uint64), np.array([0], np.uint32), np.ones(1, np.float32), 2)
    with pytest.raises(N.KytoError, match="KY-E215"):
        device.pca(one, 1, seed=0)  # fewer than 2 cells


def test_row_checks_at_lane_and_row_boundaries(device):
    """The upload's checks: a row may start below the previous row's last gene; a
    repeat or a decrease is refused wherever it falls in a long row (entries 63/64 and
    127/128 sit on either side of a 64-entry step), and so is a gene past the last."""
    g = 400
    long_row = np.arange(2, 2 * 151, 2, dtype=np.uint32)  # 150 increasing genes, 2 .. 300
    rows = [long_row, np.array([0, 1, 5], np.uint32), long_row]

    def upload(rs):
        ip = np.zeros(len(rs) + 1, np.uint64)
        ip[1:] = np.cumsum([len(r) for r in rs])
        ix = np.concatenate(rs)
        return device.upload_csr(ip, ix, np.ones(len(ix), np.float32), g)

    upload(rows).free()
    for at in (1, 63, 64, 65, 127, 128, 149):
        for kind in ("repeat", "decrease"):
            bad = long_row.copy()
            bad[at] = bad[at - 1] if kind == "repeat" else bad[at - 1] - 1
            with pytest.raises(N.KytoError, match="KY-E215"):
                upload([rows[1], bad, rows[1]])
    past = long_row.copy()
    past[-1] = g
    with pytest.raises(N.KytoError, match="KY-E215"):
        upload([rows[1], past])


def test_pca_with_empty_rows_and_constant_genes(device):
    rng = np.random.default_rng(3)
    n, g = 9000, 40
    x = (rng.random((n, g)) < 0.1) * rng.random((n, g)) * 3
    x[::7] = 0.0          # empty rows
    x[:, 5] = 0.0         # an empty gene
    x = x.astype(np.float32)
    ip = np.zeros(n + 1, np.uint64)
    ip[1:] = np.cumsum((x != 0).sum(axis=1))
    r, c = np.nonzero(x)
    csr = device.upload_csr(ip, c.astype(np.uint32), x[r, c], g)
    res = device.pca(csr, 10, seed=0, power_iterations=9)
    ref = oracle.pca(ip, c, x[r, c], g, 10, seed=0, q=9)
    assert (np.abs(res.variance - ref.variance) / ref.variance).max() < 1e-6
