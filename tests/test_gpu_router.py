"""The router's hand-written kernels against the FP64 CPU restatement (oracle/router.py):
one step's gradients, Adam's closed form, and the learned IVF index end to end (its
recall against the exact graph, and what the trained router gains over probing the
nearest centroids)."""
import ctypes

import numpy as np
import pytest

import oracle
from oracle import router as R
from kyto_amd import _native as N
from tests.helpers import clustered

pytestmark = pytest.mark.gpu


def _grad(device, params, x, t_idx, t_w, d_in, hidden, n_out):
    g = np.empty_like(params)
    loss = ctypes.c_float(0)
    N.check(N.gpu_lib().ky_router_grad(device.handle, d_in, hidden, n_out, N.ptr(params, N.f32p), N.ptr(x, N.f32p),
                                       x.shape[0], N.ptr(t_idx, N.u32p), N.ptr(t_w, N.f32p), t_idx.shape[1],
                                       N.ptr(g, N.f32p), ctypes.byref(loss)))
    return g, loss.value


def test_router_gradients_match_fp64_oracle(device):
    rng = np.random.default_rng(0)
    d_in, hidden, n_out, B, tk = 20, 48, 70, 96, 5
    n_params = d_in * hidden + hidden + hidden * hidden + hidden + hidden * n_out + n_out
    params = (rng.standard_normal(n_params) * 0.2).astype(np.float32)
    x = rng.standard_normal((B, d_in)).astype(np.float32)
    t_idx = rng.integers(0, n_out, (B, tk)).astype(np.uint32)
    t_w = np.full((B, tk), 1.0 / tk, np.float32)
    g, loss = _grad(device, params, x, t_idx, t_w, d_in, hidden, n_out)
    rg, rloss = R.grad(params, x, t_idx, t_w, d_in, hidden, n_out)
    assert abs(loss - rloss) / abs(rloss) < 1e-4
    # Each parameter block within 1e-3 of its largest gradient.
    sizes = [d_in * hidden, hidden, hidden * hidden, hidden, hidden * n_out, n_out]
    at = 0
    for s in sizes:
        a, b = g[at:at + s], rg[at:at + s]
        assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max() + 1e-7, (at, s)
        at += s
    # A planted mistake in the oracle (targets ignored) is caught by the same bar.
    bad, _ = R.grad(params, x, t_idx, t_w * 0, d_in, hidden, n_out)
    assert np.abs(g[-n_out:] - bad[-n_out:]).max() > 1e-2 * np.abs(rg[-n_out:]).max()
    # Finite differences of the oracle's loss confirm its gradient direction.
    e = np.zeros(n_params)
    j = n_params - n_out - 3
    e[j] = 1e-3
    lp = R.grad(params.astype(np.float64) + e, x, t_idx, t_w, d_in, hidden, n_out)[1]
    lm = R.grad(params.astype(np.float64) - e, x, t_idx, t_w, d_in, hidden, n_out)[1]
    assert abs((lp - lm) / 2e-3 - rg[j]) < 1e-2 * max(abs(rg[j]), 1e-3)


def test_adam_closed_form(device):
    rng = np.random.default_rng(1)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    w0 = w.copy()
    N.check(N.gpu_lib().ky_adam_step(device.handle, N.ptr(w, N.f32p), N.ptr(g, N.f32p), N.ptr(m, N.f32p),
                                     N.ptr(v, N.f32p), n, ctypes.c_float(1e-3), 1))
    # Step 1 from zero moments: m_hat = g, v_hat = g^2, so w moves by -lr sign(g).
    assert np.abs((w - w0) + 1e-3 * np.sign(g)).max() < 1e-6
    # Two more steps against the FP64 restatement.
    rw, rm, rv = w0.astype(np.float64), np.zeros(n), np.zeros(n)
    rw, rm, rv = R.adam(rw, g.astype(np.float64), rm, rv, 1e-3, 1)
    for step in (2, 3):
        g2 = rng.standard_normal(n).astype(np.float32)
        N.check(N.gpu_lib().ky_adam_step(device.handle, N.ptr(w, N.f32p), N.ptr(g2, N.f32p), N.ptr(m, N.f32p),
                                         N.ptr(v, N.f32p), n, ctypes.c_float(1e-3), step))
        rw, rm, rv = R.adam(rw, g2.astype(np.float64), rm, rv, 1e-3, step)
    assert np.abs(w - rw).max() < 1e-6


def _recall(graph_rows, exact_rows):
    hits = sum(len(set(a.tolist()) & set(b.tolist())) for a, b in zip(graph_rows, exact_rows))
    return hits / exact_rows.size


def test_learned_ivf_recall_and_gain(device):
    pts = clustered(11, 24000, 50, 30, 0.0)
    m = device.upload_dense(pts)
    try:
        exact = device.neighbors(m, 15, precision="bf16").rows()
        index = device.ivf_build(m, n_lists=128, kmeans_iters=10, seed=0, hidden=128, train_queries=12000,
                                 batch=1024, epochs=20, lr=1e-3, label_k=14)
        try:
            assert index.info["final_loss"] > 0
            g_router = index.search(m, 15, n_probe=8, use_router=True, confidence=0.0)
            g_centroid = index.search(m, 15, n_probe=8, use_router=False)
            r_router = _recall(g_router.rows(), exact)
            r_centroid = _recall(g_centroid.rows(), exact)
            print(f"recall@15 router {r_router:.4f} centroid probing {r_centroid:.4f} "
                  f"label mass in top 8 {index.info['label_recall_top']:.4f}")
            assert r_router >= 0.95
            # Every returned neighbour carries its exact distance (kyto's key).
            q = 123
            row = g_router.rows()[q]
            d = g_router.distances.reshape(-1, 14)[q]
            for r_, dd in zip(row, d):
                assert dd == np.float32(np.sqrt(oracle.squared_distance(pts[q], pts[r_])))
            # Low-confidence rows go to the exact search: a threshold of 1.01 sends all.
            g_all = index.search(m, 15, n_probe=8, use_router=True, confidence=1.01)
            assert g_all.info["fallback"] == pts.shape[0]
            assert np.array_equal(g_all.rows(), exact)
        finally:
            index.free()
    finally:
        m.free()


def test_exact_fallback_of_a_row_list_above_a_device_round(device):
    # 250,000 rows all sent to the exact search as a row list (confidence above 1): 977
    # query blocks against 768 resident, so the listed rows' last round is split over
    # point segments; the graph must be the full search's, bit for bit.
    rng = np.random.default_rng(31)
    centres = rng.standard_normal((24, 50)).astype(np.float32) * 3
    pts = (centres[rng.integers(0, 24, 250_000)] + rng.standard_normal((250_000, 50)).astype(np.float32))
    m = device.upload_dense(pts.astype(np.float32))
    try:
        full = device.neighbors(m, 15, precision="bf16")
        index = device.ivf_build(m, n_lists=256, kmeans_iters=3, seed=0, hidden=64, train_queries=8192,
                                 batch=1024, epochs=1, lr=1e-3, label_k=14)
        try:
            g = index.search(m, 15, n_probe=8, use_router=True, confidence=1.01)
            assert g.info["fallback"] == 250_000
            assert np.array_equal(g.rows(), full.rows())
            assert np.array_equal(g.distances.view(np.uint32), full.distances.view(np.uint32))
        finally:
            index.free()
    finally:
        m.free()


def test_kmeans_layout_is_nearest_centroids_and_their_means(device):
    # 2,048 rows (every row in the k-means sample: at most 64 a list) around 40 far
    # apart centres: at the fixed point Lloyd's iterations reach, every row's list is its
    # nearest centroid (f64 distances, numpy brute force) up to the error of the BF16
    # scores 2 x.c - |c|^2 that choose it (bf16 operands: 2^-8 of 2 |x| |c|, doubled;
    # where a cluster holds two centroids that error decides between them), and every
    # non-empty list's centroid is the f64 mean of its rows, rounded to f32; the lists
    # hold exactly the rows assigned to them, ascending.
    rng = np.random.default_rng(17)
    centres = rng.standard_normal((40, 16)) * 50.0
    pts = (centres[rng.integers(0, 40, 2048)] + rng.standard_normal((2048, 16))).astype(np.float32)
    m = device.upload_dense(pts)
    try:
        index = device.ivf_build(m, n_lists=40, kmeans_iters=10, seed=0, hidden=32, train_queries=1024,
                                 batch=256, epochs=1, lr=1e-3, label_k=5)
        try:
            lay = index.layout(2048, 16, 40)
        finally:
            index.free()
    finally:
        m.free()
    c, assign = lay["centroids"].astype(np.float64), lay["assign"]
    d2 = ((pts.astype(np.float64)[:, None, :] - c[None, :, :]) ** 2).sum(axis=2)
    chosen = d2[np.arange(2048), assign]
    bound = 2.0 * 2.0 ** -8 * 2.0 * np.linalg.norm(pts, axis=1) * np.linalg.norm(c, axis=1).max() + 1e-3
    assert (chosen - d2.min(axis=1) <= 2.0 * bound).all()
    # A planted mistake (each row in its second nearest list) breaks the bound.
    second = np.argsort(d2, axis=1)[:, 1]
    assert not (d2[np.arange(2048), second] - d2.min(axis=1) <= 2.0 * bound).all()
    for l in range(40):
        members = np.flatnonzero(assign == l)
        if len(members):
            want = pts[members].astype(np.float64).mean(axis=0).astype(np.float32)
            assert np.abs(lay["centroids"][l] - want).max() <= 1e-6 * max(1.0, np.abs(want).max())
        rows = lay["list_rows"][lay["list_ptr"][l]:lay["list_ptr"][l + 1]]
        assert np.array_equal(rows, members)
    assert lay["list_ptr"][-1] == 2048 and lay["alpha"] > 0


def test_router_gains_over_centroid_probing(device):
    # 250,000 rows of a 60-component Gaussian mixture, 1,024 lists: the router's logits
    # start as the scaled nearest-centroid scores (the MLP's output layer at zero) and
    # training on the exact graph's list labels moves them; probing its top 8 lists
    # finds more of the exact neighbours than probing the 8 nearest centroids.
    rng = np.random.default_rng(29)
    centres = rng.standard_normal((60, 50)).astype(np.float32) * 2.5
    scales = rng.uniform(0.5, 2.0, 60).astype(np.float32)
    lab = rng.integers(0, 60, 250_000)
    pts = (centres[lab] + rng.standard_normal((250_000, 50)).astype(np.float32) * scales[lab, None]).astype(np.float32)
    m = device.upload_dense(pts)
    try:
        exact = device.neighbors(m, 15, precision="bf16").rows()
        index = device.ivf_build(m, n_lists=1024, kmeans_iters=10, seed=0, hidden=256, train_queries=200_000,
                                 batch=4096, epochs=4, lr=1e-3, label_k=14)
        try:
            r_router = _recall(index.search(m, 15, n_probe=8, use_router=True, confidence=0.0).rows(), exact)
            r_centroid = _recall(index.search(m, 15, n_probe=8, use_router=False).rows(), exact)
            alpha = index.layout(250_000, 50, 1024)["alpha"]
        finally:
            index.free()
    finally:
        m.free()
    print(f"recall@15 router {r_router:.4f} centroid probing {r_centroid:.4f} alpha {alpha:.4g} "
          f"label mass in top 8 {index.info['label_recall_top']:.4f}")
    assert r_router > r_centroid


@pytest.mark.parametrize("d, k", [(20, 15), (50, 31)])
def test_ivf_search_of_every_list_is_the_exact_graph(device, d, k):
    """With as many probes as lists every point is scored, so the list-by-list scan and
    the per-row merge must give the exact graph bit for bit (kyto's f64 key and (key,
    row) order); 8,000 rows in 8 lists of ~1,000 points are staged in LDS a chunk at a
    time."""
    pts = clustered(5, 8000, d, 12, 0.0)
    m = device.upload_dense(pts)
    try:
        exact = device.neighbors(m, k, precision="fp32")
        index = device.ivf_build(m, n_lists=8, kmeans_iters=3, seed=0, hidden=32, train_queries=2000,
                                 batch=512, epochs=1, lr=1e-3, label_k=k - 1)
        try:
            for use_router in (False, True):
                g = index.search(m, k, n_probe=8, use_router=use_router, confidence=0.0)
                assert g.info["fallback"] == 0
                assert np.array_equal(g.rows(), exact.rows())
                assert np.array_equal(g.distances.view(np.uint32), exact.distances.view(np.uint32))
        finally:
            index.free()
    finally:
        m.free()
