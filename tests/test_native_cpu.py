"""The native library on a host without a GPU: it loads, exports every symbol the
header declares, its FP64 dense steps agree with the oracle's (LAPACK) ones, and the
plan API checks its arguments as kyto does. No device compute here."""
import os
import re

import numpy as np
import pytest

import oracle
import kyto_amd
from kyto_amd import KytoError
from kyto_amd import _native as N
from kyto_amd import device as D
from kyto_amd import synthetic as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "kyto_gpu.h")).read()
    return sorted(set(re.findall(r"\b(ky_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.gpu_lib()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(N.SIGNATURES) | {"ky_last_error", "ky_store_layer_read"} >= set(names)


def test_missing_device_is_e504():
    if D.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(N.KytoError) as e:
        D.Device(0)
    assert e.value.code == "KY-E504"


def test_extend_basis_against_oracle():
    rng = np.random.default_rng(0)
    basis = np.linalg.qr(rng.standard_normal((300, 40)))[0]
    block = rng.standard_normal((300, 30))
    block[:, 5] = basis[:, 3] * 2.0  # a direction the basis holds
    mine = D.dense_extend_basis(basis, block)
    ref = oracle.extend_basis(basis, block)
    assert mine.shape == ref.shape == (300, 29)
    # Same span, orthonormal, orthogonal to the basis.
    assert np.abs(mine.T @ mine - np.eye(29)).max() < 1e-12
    assert np.abs(basis.T @ mine).max() < 1e-12
    s = np.linalg.svd(mine.T @ ref, compute_uv=False)
    assert np.abs(s - 1.0).max() < 1e-10
    # Known answers (pca_dense.rs:319-334).
    b = np.array([[1.0], [0.0], [0.0]])
    blk = np.array([[2.0, 4.0], [0.0, 1.0], [0.0, 0.0]])
    added = D.dense_extend_basis(b, blk)
    assert added.shape[1] == 1 and abs(abs(added[1, 0]) - 1.0) < 1e-12
    assert D.dense_extend_basis(np.hstack([b, added]), blk).shape[1] == 0
    assert D.dense_extend_basis(b, np.zeros((3, 2))).shape[1] == 0
    assert D.dense_extend_basis(np.zeros((3, 0)), blk).shape[1] == 2


@pytest.mark.parametrize("spread", [1.0, 1e-8])
def test_extend_basis_well_and_ill_conditioned(spread):
    # A full-rank block: orthonormalised by Cholesky QR twice when well conditioned,
    # by Householder when a column is 1e-8 of the others (condition ~1e8, still above
    # the 1e-10 drop tolerance, so every column is kept).
    rng = np.random.default_rng(7)
    basis = np.linalg.qr(rng.standard_normal((500, 40)))[0]
    block = rng.standard_normal((500, 30))
    block[:, 7] *= spread
    mine = D.dense_extend_basis(basis, block)
    ref = oracle.extend_basis(basis, block)
    assert mine.shape == ref.shape == (500, 30)
    assert np.abs(mine.T @ mine - np.eye(30)).max() < 1e-12
    assert np.abs(basis.T @ mine).max() < 1e-12
    s = np.linalg.svd(mine.T @ ref, compute_uv=False)
    assert np.abs(s - 1.0).max() < 1e-10


def test_rayleigh_ritz_against_oracle():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((200, 200))
    c = a @ a.T
    basis = np.linalg.qr(rng.standard_normal((200, 90)))[0]
    ev, lo = D.dense_rayleigh_ritz(basis, c @ basis, 30)
    rev, rlo = oracle.rayleigh_ritz(basis, c @ basis, 30)
    assert np.abs(ev - rev).max() / rev.max() < 1e-12
    assert np.abs(lo - rlo).max() < 1e-8  # both sign-flipped
    # Known answers (pca_dense.rs:336-355).
    ev, lo = D.dense_rayleigh_ritz(np.eye(3), np.diag([1.0, 3.0, 2.0]), 2)
    assert abs(ev[0] - 3) < 1e-12 and abs(ev[1] - 2) < 1e-12
    assert abs(lo[1, 0] - 1) < 1e-12 and abs(lo[2, 1] - 1) < 1e-12
    ev, lo = D.dense_rayleigh_ritz(np.array([[0.0], [-1.0], [0.0]]), np.array([[0.0], [-5.0], [0.0]]), 3)
    assert list(ev) == [5.0, 0.0, 0.0]
    assert np.abs(lo.T @ lo - np.eye(3)).max() < 1e-12
    assert abs(lo[1, 0] - 1.0) < 1e-12


def test_rayleigh_ritz_full_space_equals_eigh():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((120, 120))
    c = a @ a.T
    ev, lo = D.dense_rayleigh_ritz(np.eye(120), c, 120)
    want = np.sort(np.linalg.eigvalsh(c))[::-1]
    assert np.abs(ev - want).max() / want.max() < 1e-12
    assert np.abs(lo.T @ c @ lo - np.diag(ev)).max() / want.max() < 1e-10


def test_plan_arguments_are_checked():
    layer = (np.array([0, 1, 2], np.uint64), np.array([0, 1], np.uint32), np.array([1.0, 2.0], np.float32), 3)
    a = kyto_amd.scan(layer)
    with pytest.raises(N.KytoError, match="KY-E215"):
        a.pca(n_components=2)  # seed required (R7)
    with pytest.raises(N.KytoError, match="KY-E215"):
        a.pca(seed=-1)
    with pytest.raises(N.KytoError, match="KY-E215"):
        a.pca(seed=0, n_components=0)
    p = a.pca(seed=0, n_components=2)
    with pytest.raises(N.KytoError, match="KY-E215"):
        p.neighbors(n_neighbors=1)
    with pytest.raises(N.KytoError, match="KY-E215"):
        p.neighbors(metric="manhattan")
    with pytest.raises(N.KytoError, match="KY-E109"):
        p.neighbors("nope")
    with pytest.raises(N.KytoError, match="KY-E215"):
        p.on("tpu")
    with pytest.raises(N.KytoError, match="KY-E501"):
        p.neighbors().on("cpu").collect()
    assert kyto_amd.parse_memory("32GB") == 32 * 10**9
    assert kyto_amd.parse_memory("512MiB") == 512 * 2**20
    with pytest.raises(N.KytoError):
        kyto_amd.parse_memory("lots")
    assert "NEIGHBORS" in p.neighbors().explain()


def test_synthetic_generator_is_deterministic_and_hits_density():
    a = S.counts(3000, 500, 8, 0.08, seed=3)
    b = S.counts(3000, 500, 8, 0.08, seed=3)
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.values, b.values)
    assert abs(a.density - 0.08) < 0.004
    c = S.counts(3000, 500, 8, 0.08, seed=4)
    assert not np.array_equal(a.values[:100], c.values[:100])
    # Rows: strictly increasing gene indices.
    for r in range(0, 3000, 97):
        s, e = int(a.indptr[r]), int(a.indptr[r + 1])
        assert np.all(np.diff(a.indices[s:e].astype(np.int64)) > 0)
    # The input layer is kyto's normalize_total + log1p.
    norm, log = oracle.normalize_log1p(a.indptr, a.values)
    S.normalize_log1p(a)
    assert np.array_equal(a.values.view(np.uint32), log.view(np.uint32))


def test_plan_layer_maps_are_checked():
    layer = (np.array([0, 1, 2], np.uint64), np.array([0, 1], np.uint32), np.array([1.0, 2.0], np.float32), 3)
    a = kyto_amd.scan(layer)
    with pytest.raises(kyto_amd.KytoError, match="KY-E501|501"):
        a.normalize_total()
    with pytest.raises(kyto_amd.KytoError, match="KY-E215|215"):
        a.normalize_total(-1.0)
    b = a.normalize_total(1e4).log1p()
    assert b.explain().splitlines()[:2] == ['LOG1P "normalized" -> "log1p" [device: auto]',
                                            'NORMALIZE_TOTAL "X" -> "normalized" target_sum=10000 [device: auto]']
    with pytest.raises(kyto_amd.KytoError, match="220"):
        b.log1p(layer="X")
    # pca reads the latest written layer.
    assert b.pca(seed=0).steps[-1].layer == "log1p"
    assert b.pca(seed=0, layer="X").steps[-1].layer == "X"
    with pytest.raises(kyto_amd.KytoError):
        b.pca(seed=0, layer="nope")


def _store_plan(n, g, nnz, k=50, nn=15, q=4, precision="bf16"):
    import kyto_amd
    ref = kyto_amd._StoreRef("/nowhere", "X", n, g, nnz, "float32", 4096, 1024)
    return kyto_amd.Atlas({"X": ref}, "X").pca(n_components=k, seed=0, power_iterations=q).neighbors(
        n_neighbors=nn, precision=precision)


def test_memory_limit_plans_rounds_and_refuses_with_e401():
    # memory_limit.rs:119-175: the search shrinks its rounds to fit; a plan that cannot
    # fit raises KY-E401 before anything runs (the store here is never opened).
    plan = _store_plan(1_000_000, 2000, 160_000_000)
    assert plan._memory_plan(None) == {}
    assert plan._memory_plan(64 << 30) == {1: 1_000_000}
    rounds = plan._memory_plan(4 << 30)
    assert 0 < rounds[1] < 1_000_000 and rounds[1] % 256 == 0
    with pytest.raises(KytoError) as e:
        plan._memory_plan(2 << 30)
    assert e.value.code == "KY-E401" and "the plan needs at least" in str(e.value)
    with pytest.raises(KytoError) as e:
        plan.collect(memory_limit="1GB")
    assert e.value.code == "KY-E401"
    # 4M x 3,000 (~840M values), k = 30: runs in 32 GB.
    big = _store_plan(4_000_000, 3000, 840_000_000, nn=30)
    assert big._memory_plan(32 * 10**9)[1] > 0
    assert big._total_rows() == 4_000_000 * (4 + 3) + 4_000_000


def test_collect_arguments_are_checked():
    plan = _store_plan(1000, 50, 5000, k=10)
    for bad in (0, True, 2.5):
        with pytest.raises(KytoError, match="KY-E215"):
            plan.collect(threads=bad)
    with pytest.raises(KytoError, match="KY-E215"):
        plan.collect(progress=5)
    with pytest.raises(KytoError, match="KY-E215"):
        plan.collect(memory_limit="lots")
