"""ky_neighbors_static, the entry point that allocates nothing: its graph equals
ky_neighbors' bit for bit on two calls in a row (the second with a smaller input,
fewer components used), at both precisions and both metrics; device memory is unchanged
by a call; an input beyond the fixed region is refused with KY-E215."""
import ctypes

import numpy as np
import pytest

from kyto_amd import _native as N
from tests.helpers import clustered

pytestmark = pytest.mark.gpu


def _static(device, m, n_neighbors, n_components, metric, precision):
    n = m.rows
    k = n_neighbors - 1
    idx = np.empty(n * k, np.uint32)
    dist = np.empty(n * k, np.float32)
    info = N.KnnInfo()
    N.check(N.gpu_lib().ky_neighbors_static(device.handle, m.handle, int(n_components or 0), int(n_neighbors),
                                            N.METRIC[metric], N.PRECISION[precision], N.ptr(idx, N.u32p),
                                            N.ptr(dist, N.f32p), ctypes.byref(info)))
    return idx, dist, info


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_static_entry_point_equals_ky_neighbors(device, precision, metric):
    # 40,000 rows: the BF16 search runs in rounds of 8,448 queries in the 64 MiB.
    rng = np.random.default_rng(41)
    centres = rng.standard_normal((11, 50)).astype(np.float32) * 3
    big = (centres[rng.integers(0, 11, 40_000)] + rng.standard_normal((40_000, 50))).astype(np.float32)
    small = clustered(42, 3000, 24, 5, 2.0)
    cases = [(big, 15, None), (small, 9, 16)]
    mats = [device.upload_dense(p) for p, _, _ in cases]
    try:
        for (pts, nn, nc), m in zip(cases, mats):
            want = device.neighbors(m, nn, n_components=nc, metric=metric, precision=precision)
            before = _free_bytes()
            idx, dist, info = _static(device, m, nn, nc, metric, precision)
            assert _free_bytes() == before
            assert np.array_equal(idx, want.indices)
            assert np.array_equal(dist.view(np.uint32), want.distances.view(np.uint32))
            assert info.scratch_bytes <= N.KY_STATIC_SCRATCH_BYTES
            assert info.certified + info.fallback == m.rows
            if precision == "bf16" and m.rows == 40_000:
                assert info.rounds > 1
    finally:
        for m in mats:
            m.free()


def test_static_entry_point_refuses_what_does_not_fit(device):
    pts = np.random.default_rng(43).standard_normal((200_000, 50)).astype(np.float32)
    m = device.upload_dense(pts)
    try:
        with pytest.raises(N.KytoError, match="KY-E215.*fixed region"):
            _static(device, m, 15, None, "euclidean", "bf16")
    finally:
        m.free()
