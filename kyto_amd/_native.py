"""ctypes binding of libkyto_gpu.so (include/kyto_gpu.h) and libkyto_synth.so.

The library is built in-tree (kyto_amd/build.py). Loading fails loudly when it is
missing: there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
# KYTO_AMD_SANITIZED=1: the host code's AddressSanitizer + UBSan build (kyto_amd/build.py
# sanitize), for CPU test runs under the sanitizer runtime (tests/sanitize.sh).
_SAN = "_san" if os.environ.get("KYTO_AMD_SANITIZED") == "1" else ""
GPU_LIB = os.path.join(HERE, f"libkyto_gpu{_SAN}.so")
SYNTH_LIB = os.path.join(HERE, f"libkyto_synth{_SAN}.so")

KY_OK = 0
KY_E215 = 215
KY_E401 = 401
KY_E501 = 501
KY_E503 = 503
KY_E504 = 504
KY_E505 = 505
# include/kyto_gpu.h: the fixed device region of ky_neighbors_static.
KY_STATIC_SCRATCH_BYTES = 64 << 20

KEY = {"squared_euclidean": 0, "cosine": 1}
METRIC = {"euclidean": 0, "cosine": 1}
PRECISION = {"fp32": 0, "bf16": 1}

u64p = POINTER(c_uint64)
u32p = POINTER(c_uint32)
f32p = POINTER(c_float)
f64p = POINTER(c_double)


class KytoError(RuntimeError):
    """A kyto error: `code` is "KY-Exxx" (errors.toml), `status` the integer."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.code = f"KY-E{status:03d}"
        super().__init__(f"{self.code}: {message}")


class PcaTiming(ctypes.Structure):
    _fields_ = [("upload_ms", c_double), ("pass_ms", c_double * 16), ("dense_ms", c_double),
                ("total_ms", c_double), ("n_passes", c_uint32), ("spmm_launches", c_uint32 * 2),
                ("spmm_ms", c_double * 2), ("spmm_bytes", c_double * 2)]


class KnnInfo(ctypes.Structure):
    _fields_ = [("candidates", c_uint64), ("certified", c_uint64), ("fallback", c_uint64),
                ("search_ms", c_double), ("rerank_ms", c_double), ("fallback_ms", c_double),
                ("total_ms", c_double), ("scan_kernel_ms", c_double), ("rounds", c_uint64),
                ("scratch_bytes", c_uint64), ("rescanned", c_uint64), ("rescan_ms", c_double)]


class StreamInfo(ctypes.Structure):
    _fields_ = [("blocks", c_uint64), ("nnz", c_uint64), ("read_ms", c_double), ("total_ms", c_double),
                ("bytes", c_double)]


# ky_block_reader: (user, first_row, n_rows, indptr, indices, values, capacity) -> nnz
BlockReader = ctypes.CFUNCTYPE(ctypes.c_int64, c_void_p, c_uint64, c_uint64, POINTER(c_uint64), POINTER(c_uint32),
                               POINTER(c_float), c_uint64)


class ConnInfo(ctypes.Structure):
    _fields_ = [("edges", c_uint64), ("zero_rho_rows", c_uint64), ("upload_ms", c_double), ("scales_ms", c_double),
                ("union_ms", c_double), ("total_ms", c_double)]


class IvfParams(ctypes.Structure):
    _fields_ = [("n_lists", c_uint32), ("kmeans_iters", c_uint32), ("seed", c_uint64),
                ("hidden", c_uint32), ("train_queries", c_uint32), ("batch", c_uint32),
                ("epochs", c_uint32), ("lr", c_float), ("label_k", c_uint32)]


class IvfInfo(ctypes.Structure):
    _fields_ = [("kmeans_ms", c_double), ("labels_ms", c_double), ("train_ms", c_double),
                ("final_loss", c_float), ("label_recall_top", c_float)]


class IvfSearchInfo(ctypes.Structure):
    _fields_ = [("fallback", c_uint64), ("route_ms", c_double), ("search_ms", c_double),
                ("fallback_ms", c_double), ("graph_ms", c_double), ("total_ms", c_double)]


# name -> (restype, argtypes); every function returns an int status.
SIGNATURES = {
    "ky_device_count": [POINTER(c_int)],
    "ky_open": [c_int, POINTER(c_void_p)],
    "ky_close": [c_void_p],
    "ky_device_info": [c_void_p, c_char_p, c_size_t, u64p],
    "ky_run_begin": [c_void_p, c_uint64],
    "ky_cancel": [c_void_p],
    "ky_progress": [c_void_p, u64p, u64p],
    "ky_set_threads": [c_uint32],
    "ky_trim": [c_void_p],
    "ky_pca_bytes": [c_uint64, c_uint32, c_uint64, c_uint32, c_uint32, u64p],
    "ky_connectivities_bytes": [c_uint64, c_uint32, u64p],
    "ky_csr_upload": [c_void_p, u64p, u32p, f32p, c_uint64, c_uint32, POINTER(c_void_p)],
    "ky_csr_free": [c_void_p],
    "ky_csr_shape": [c_void_p, u64p, u32p, u64p],
    "ky_dense_upload": [c_void_p, f32p, c_uint64, c_uint32, POINTER(c_void_p)],
    "ky_dense_download": [c_void_p, f32p],
    "ky_dense_shape": [c_void_p, u64p, u32p],
    "ky_dense_free": [c_void_p],
    "ky_pca_sketch": [c_void_p, c_void_p, c_uint64, c_uint64, c_uint32, f64p, f64p, f64p],
    "ky_pca_product": [c_void_p, c_void_p, f64p, c_uint32, f64p, f64p, f64p],
    "ky_pca_project": [c_void_p, c_void_p, f64p, c_uint32, f64p, f64p, f32p, POINTER(c_void_p), f64p, u64p],
    "ky_pca": [c_void_p, c_void_p, c_uint32, c_uint64, c_uint32, f32p, POINTER(c_void_p), f32p, f64p, f64p,
               f64p, u32p, POINTER(PcaTiming)],
    "ky_csr_upload_counts": [c_void_p, u64p, u32p, f32p, c_uint64, c_uint32, c_double, c_int, POINTER(c_void_p)],
    "ky_csr_values": [c_void_p, f32p],
    "ky_csr_stream": [c_void_p, c_uint64, c_uint32, c_uint64, BlockReader, c_void_p, c_uint64, c_double, c_int,
                      POINTER(c_void_p), POINTER(StreamInfo)],
    "ky_neighbors": [c_void_p, c_void_p, c_uint32, c_uint32, c_int, c_int, u32p, f32p, POINTER(KnnInfo)],
    "ky_neighbors_rounds": [c_void_p, c_void_p, c_uint32, c_uint32, c_int, c_int, c_uint64, u32p, f32p,
                            POINTER(KnnInfo)],
    "ky_neighbors_static": [c_void_p, c_void_p, c_uint32, c_uint32, c_int, c_int, u32p, f32p, POINTER(KnnInfo)],
    "ky_neighbors_bytes": [c_void_p, c_uint64, c_uint32, c_uint32, c_uint32, c_int, c_int, c_uint64, u64p],
    "ky_neighbors_plan": [c_void_p, c_uint64, c_uint32, c_uint32, c_uint32, c_int, c_int, c_uint64, u64p, u64p],
    "ky_nearest_keys": [c_void_p, f32p, c_uint64, c_void_p, c_int, f32p, f32p, u32p, c_uint32, f32p, u32p],
    "ky_connectivities": [c_void_p, c_uint64, c_uint32, u32p, f32p, POINTER(c_void_p), f32p, f32p,
                          POINTER(ConnInfo)],
    "ky_graph_shape": [c_void_p, u64p, u64p],
    "ky_graph_download": [c_void_p, u64p, u32p, f32p],
    "ky_graph_free": [c_void_p],
    "ky_ivf_build": [c_void_p, c_void_p, POINTER(IvfParams), POINTER(c_void_p), POINTER(IvfInfo)],
    "ky_ivf_free": [c_void_p],
    "ky_ivf_layout": [c_void_p, f32p, u32p, u32p, u32p, POINTER(c_float)],
    "ky_ivf_search": [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_int, c_float, u32p, f32p,
                      POINTER(IvfSearchInfo)],
    "ky_router_grad": [c_void_p, c_uint32, c_uint32, c_uint32, f32p, f32p, c_uint32, u32p, f32p, c_uint32, f32p,
                       POINTER(c_float)],
    "ky_adam_step": [c_void_p, f32p, f32p, f32p, f32p, c_uint64, c_float, c_uint32],
    "ky_dense_extend_basis": [f64p, c_uint32, c_uint32, f64p, c_uint32, f64p, u32p],
    "ky_dense_rayleigh_ritz": [f64p, f64p, c_uint32, c_uint32, c_uint32, f64p, f64p],
    "ky_store_layer_open": [c_char_p, c_char_p, POINTER(c_void_p)],
    "ky_store_layer_shape": [c_void_p, u64p, u32p, u64p, u64p, u64p, POINTER(c_int)],
    "ky_store_layer_close": [c_void_p],
    "ky_store_layer_rows": [c_void_p, c_uint64, c_uint64, u64p, u32p, f32p, c_uint64, u64p],
    "ky_csr_stream_layer": [c_void_p, c_void_p, c_uint64, c_double, c_int, POINTER(c_void_p), POINTER(StreamInfo)],
    "ky_store_write_layer": [c_char_p, c_char_p, c_uint64, c_uint64, u64p, u32p, c_void_p, c_int, c_uint64, c_uint64,
                             c_uint64, c_uint64, c_int],
    "ky_store_write_embedding": [c_char_p, c_char_p, c_int, f32p, c_uint64, c_uint32, c_uint64, c_int],
    "ky_store_read_embedding": [c_char_p, c_char_p, u64p, u32p, f32p, c_uint64],
    "ky_store_write_graph": [c_char_p, c_char_p, c_uint64, u64p, u32p, f32p, c_uint64, c_uint64, c_int],
    "ky_store_read_graph": [c_char_p, c_char_p, u64p, u64p, u64p, u32p, f32p],
}

_gpu = None
_synth = None


def gpu_lib():
    """libkyto_gpu.so, loaded once. Raises if it was not built."""
    global _gpu
    if _gpu is None:
        if not os.path.exists(GPU_LIB):
            raise ImportError(f"{GPU_LIB} is missing: build it with `python -m kyto_amd.build` "
                              "(there is no CPU fallback)")
        lib = ctypes.CDLL(GPU_LIB)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = c_int
            fn.argtypes = args
        lib.ky_last_error.restype = c_char_p
        lib.ky_last_error.argtypes = []
        # A ky_block_reader: returns the block's number of values, -1 on failure.
        lib.ky_store_layer_read.restype = ctypes.c_int64
        lib.ky_store_layer_read.argtypes = [c_void_p, c_uint64, c_uint64, u64p, u32p, f32p, c_uint64]
        _gpu = lib
    return _gpu


def check(status: int) -> None:
    if status != KY_OK:
        raise KytoError(status, gpu_lib().ky_last_error().decode(errors="replace"))


class SynthSpec(ctypes.Structure):
    _fields_ = [("n_cells", c_uint64), ("n_genes", c_uint32), ("n_types", c_uint32), ("seed", c_uint64),
                ("mu_log", c_double), ("sigma_log", c_double), ("size_sigma", c_double), ("phi", c_double),
                ("marker_share", c_double), ("mean_scale", c_double)]


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_LIB):
            raise ImportError(f"{SYNTH_LIB} is missing: build it with `python -m kyto_amd.build`")
        lib = ctypes.CDLL(SYNTH_LIB)
        lib.ks_expected_density.restype = c_double
        lib.ks_expected_density.argtypes = [POINTER(SynthSpec), c_uint64]
        lib.ks_count.restype = c_uint64
        lib.ks_count.argtypes = [POINTER(SynthSpec), u64p]
        lib.ks_fill.restype = None
        lib.ks_fill.argtypes = [POINTER(SynthSpec), u64p, u32p, f32p]
        lib.ks_block.restype = c_uint64
        lib.ks_block.argtypes = [POINTER(SynthSpec), c_uint64, c_uint64, c_void_p, c_void_p, c_void_p, c_uint64]
        lib.ks_cell_types.restype = None
        lib.ks_cell_types.argtypes = [POINTER(SynthSpec), u32p]
        lib.ks_normalize_log1p.restype = None
        lib.ks_normalize_log1p.argtypes = [u64p, c_uint64, c_double, f32p]
        _synth = lib
    return _synth


def ptr(a, t):
    return a.ctypes.data_as(t)
