"""The plan's run control on the device (crates/kyto-exec/src/run_control.rs,
crates/kyto-plan/src/memory_limit.rs): a memory limit that forces the neighbours search
into rounds gives the same outputs; progress(done, total) moves forward and ends at
(total, total); a cancel stops a run at its next pass, block or round with KY-E503,
and the context runs again afterwards."""
import numpy as np
import pytest

import kyto_amd
from kyto_amd import KytoError
from kyto_amd import device as D
from kyto_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def layer():
    return S.log1p_layer(120_000, 800, 10, 0.08, seed=51)


def _plan(L):
    return (kyto_amd.scan((L.indptr, L.indices, L.values, 800)).pca(n_components=30, seed=0, power_iterations=4)
            .neighbors(n_neighbors=15).on(device="gpu"))


def test_memory_limit_rounds_give_the_same_outputs(layer):
    plan = _plan(layer)
    free = plan.collect()
    n = layer.n_cells
    pca = D.pca_bytes(n, 800, layer.nnz, 30, 4)
    smallest = D.neighbors_bytes(n, 30, 15, round_rows=256)
    limit = max(pca, smallest + 4 * n * 30 + 256) + (64 << 20)
    rounds = plan._memory_plan(limit)
    assert rounds[1] < n
    seen = []
    out = plan.collect(memory_limit=limit, threads=3, progress=lambda done, total: seen.append((done, total)))
    assert np.array_equal(out["X_pca"].view(np.uint32), free["X_pca"].view(np.uint32))
    assert np.array_equal(out["distances"].indices, free["distances"].indices)
    assert np.array_equal(out["distances"].distances.view(np.uint32), free["distances"].distances.view(np.uint32))
    assert np.array_equal(out["connectivities"].weights.view(np.uint32),
                          free["connectivities"].weights.view(np.uint32))
    total = plan._total_rows()
    assert seen and seen[-1] == (total, total)
    assert all(a[0] <= b[0] for a, b in zip(seen, seen[1:])) and all(t == total for _, t in seen)
    with pytest.raises(KytoError) as e:
        plan.collect(memory_limit=pca // 2)
    assert e.value.code == "KY-E401"


def test_cancel_stops_passes_blocks_and_rounds(device, layer):
    L = layer
    csr = device.upload_csr(L.indptr, L.indices, L.values, 800)
    try:
        device.run_begin(10)
        device.cancel()
        with pytest.raises(KytoError) as e:
            device.pca(csr, 30, seed=0, power_iterations=4)
        assert e.value.code == "KY-E503" and "cancelled after" in str(e.value)
        device.run_begin(10)
        res = device.pca(csr, 30, seed=0, power_iterations=4)  # runs again after run_begin
    finally:
        csr.free()
    device.cancel()
    with pytest.raises(KytoError) as e:
        device.neighbors(res.device_embedding, 15, precision="bf16")
    assert e.value.code == "KY-E503"
    device.run_begin(0)
    g = device.neighbors(res.device_embedding, 15, precision="bf16")
    assert g.info["certified"] + g.info["fallback"] == L.n_cells
    res.device_embedding.free()
    # A streamed layer: the reader cancels while the third block is read, the stream
    # stops before the fourth.
    reader = S.block_reader(40_000, 300, 4, 0.08, seed=3)
    calls = []

    def cancelling(first, count, indptr, indices, values, capacity):
        calls.append(first)
        if len(calls) == 3:
            device.cancel()
        return reader(first, count, indptr, indices, values, capacity)

    device.run_begin(40_000)
    with pytest.raises(KytoError) as e:
        device.stream_counts(40_000, 300, cancelling, block_rows=4096)
    assert e.value.code == "KY-E503"
    assert len(calls) in (3, 4)  # the third block's read (and its retry with a larger buffer)
    device.run_begin(0)


def test_ctrl_c_during_collect_cancels_the_run():
    L = S.log1p_layer(500_000, 1000, 12, 0.08, seed=52)
    plan = (kyto_amd.scan((L.indptr, L.indices, L.values, 1000)).pca(n_components=50, seed=0, power_iterations=4)
            .neighbors(n_neighbors=15).on(device="gpu"))

    def interrupt(done, total):
        if done > 0:
            raise KeyboardInterrupt

    with pytest.raises(KytoError) as e:
        plan.collect(progress=interrupt)
    assert e.value.code == "KY-E503"
