"""ctypes wrapper for the compiled reference harness in oracle/_ref/.

Test infrastructure only (golden-vector generation, oracle validation and the
CPU baseline leg of bench.py).  The library is built from the reference
sources by oracle/Makefile; see ref_harness.cpp for the forwarded API.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict = {}


def lib_path(portable: bool = False) -> str:
    name = "libglassref_portable.so" if portable else "libglassref.so"
    return os.path.join(HERE, "_ref", name)


def available(portable: bool = False) -> bool:
    return os.path.exists(lib_path(portable))


def load(portable: bool = False):
    key = bool(portable)
    if key in _LIBS:
        return _LIBS[key]
    lib = C.CDLL(lib_path(portable))
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    lib.ref_set_num_threads.argtypes = [C.c_int]
    lib.ref_set_num_threads.restype = C.c_int
    lib.ref_build_graph.argtypes = [C.c_char_p, f32p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_int,
                                    C.c_char_p]
    lib.ref_build_graph.restype = C.c_double
    lib.ref_searcher_new.argtypes = [C.c_char_p, f32p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]
    lib.ref_searcher_new.restype = C.c_void_p
    lib.ref_searcher_new_factor.argtypes = [C.c_char_p, f32p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p,
                                            C.c_float]
    lib.ref_searcher_new_factor.restype = C.c_void_p
    lib.ref_searcher_free.argtypes = [C.c_void_p]
    lib.ref_set_ef.argtypes = [C.c_void_p, C.c_int]
    lib.ref_optimize.argtypes = [C.c_void_p, C.c_int]
    lib.ref_search.argtypes = [C.c_void_p, f32p, C.c_int, i32p, f32p]
    lib.ref_search_batch.argtypes = [C.c_void_p, f32p, C.c_int, C.c_int, i32p, f32p]
    lib.ref_dist_cmps.argtypes = [C.c_void_p]
    lib.ref_dist_cmps.restype = C.c_double
    lib.ref_quant_probe.argtypes = [C.c_char_p, C.c_char_p, f32p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]
    lib.ref_quant_probe.restype = C.c_int
    _LIBS[key] = lib
    return lib


def set_num_threads(n: int, portable: bool = False) -> int:
    return load(portable).ref_set_num_threads(int(n))


def build_graph(index_type: str, data: np.ndarray, path: str, metric: str = "L2", quant: str = "FP32", R: int = 32,
                L: int = 200, portable: bool = False) -> float:
    data = np.ascontiguousarray(data, dtype=np.float32)
    t = load(portable).ref_build_graph(index_type.encode(), data, data.shape[0], data.shape[1], metric.encode(),
                                       quant.encode(), R, L, path.encode())
    if t < 0:
        raise RuntimeError("reference build_graph failed")
    return t


def graph_convert(src: str, fmt: str, dst: str, portable: bool = False) -> int:
    """glass::Graph::load(src, fmt) + save(dst); returns N."""
    lib = load(portable)
    lib.ref_graph_convert.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    lib.ref_graph_convert.restype = C.c_int
    return lib.ref_graph_convert(src.encode(), fmt.encode(), dst.encode())


class RefSearcher:
    """Mirror of glass.Searcher (python/bindings.cc:90-154) over the harness."""

    def __init__(self, graph_path: str, data: np.ndarray, metric: str, quantizer: str, refine_quant: str = "",
                 portable: bool = False, refine_factor: float | None = None):
        self.lib = load(portable)
        data = np.ascontiguousarray(data, dtype=np.float32)
        self.dim = data.shape[1]
        if refine_factor is None:
            self.h = self.lib.ref_searcher_new(graph_path.encode(), data, data.shape[0], data.shape[1],
                                               metric.encode(), quantizer.encode(), refine_quant.encode())
        else:
            self.h = self.lib.ref_searcher_new_factor(graph_path.encode(), data, data.shape[0], data.shape[1],
                                                      metric.encode(), quantizer.encode(), refine_quant.encode(),
                                                      float(refine_factor))
        if not self.h:
            raise RuntimeError("reference create_searcher returned null")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_searcher_free(self.h)
            self.h = None

    def set_ef(self, ef: int):
        self.lib.ref_set_ef(self.h, int(ef))

    def optimize(self, num_threads: int = 0):
        self.lib.ref_optimize(self.h, int(num_threads))

    def search(self, q: np.ndarray, k: int):
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(-1)
        ids = np.zeros(k, np.int32)
        dis = np.zeros(k, np.float32)
        self.lib.ref_search(self.h, q, int(k), ids, dis)
        return ids, dis

    def batch_search(self, qs: np.ndarray, k: int):
        qs = np.ascontiguousarray(qs, dtype=np.float32)
        nq = qs.shape[0]
        ids = np.zeros((nq, k), np.int32)
        dis = np.zeros((nq, k), np.float32)
        self.lib.ref_search_batch(self.h, qs, nq, int(k), ids, dis)
        return ids, dis

    def get_last_search_dist_cmps(self) -> float:
        return self.lib.ref_dist_cmps(self.h)


def quant_probe(quant: str, metric: str, data: np.ndarray, queries: np.ndarray | None = None,
                ids: np.ndarray | None = None, want_codes: bool = True, portable: bool = False):
    """Train `quant` on data with the reference quantizer; return
    (codes[n, code_size] uint8, params float32, dists[nq, nids] float64,
    code_size, dim_align)."""
    lib = load(portable)
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, d = data.shape
    cs = C.c_int(0)
    da = C.c_int(0)
    # first pass: sizes only
    lib.ref_quant_probe(quant.encode(), metric.encode(), data[: max(2, min(n, 16))], min(n, max(2, min(n, 16))), d,
                        None, 0, None, 0, None, None, None, C.byref(cs), C.byref(da))
    code_size, dim_align = cs.value, da.value
    codes = np.zeros((n, code_size), np.uint8) if want_codes else None
    params = np.zeros(2 * dim_align + 2, np.float32)
    nq = 0 if queries is None else queries.shape[0]
    nids = 0 if ids is None else len(ids)
    qs = None if queries is None else np.ascontiguousarray(queries, dtype=np.float32)
    idv = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
    dist = np.zeros((nq, nids), np.float64) if nq and nids else None
    rc = lib.ref_quant_probe(quant.encode(), metric.encode(), data, n, d,
                             None if qs is None else qs.ctypes.data_as(C.c_void_p), nq,
                             None if idv is None else idv.ctypes.data_as(C.c_void_p), nids,
                             None if codes is None else codes.ctypes.data_as(C.c_void_p),
                             params.ctypes.data_as(C.c_void_p),
                             None if dist is None else dist.ctypes.data_as(C.c_void_p), C.byref(cs), C.byref(da))
    if rc != 0:
        raise ValueError(f"unknown quantizer {quant}")
    return codes, params, dist, code_size, dim_align


def pq_probe(metric: str, data: np.ndarray, queries: np.ndarray, ids: np.ndarray):
    """Train the reference ProductQuant (glass/quant/product_quant.hpp, 4 dims
    per sub-quantizer, 256 centroids) on data; return (codebook[nsq, 256, 4],
    codes[n, nsq], lut[nq, nsq, 256], dists[nq, nids] float64)."""
    lib = load(False)
    data = np.ascontiguousarray(data, dtype=np.float32)
    qs = np.ascontiguousarray(queries, dtype=np.float32)
    idv = np.ascontiguousarray(ids, dtype=np.int32)
    n, d = data.shape
    nsq = d // 4
    cb = np.zeros((nsq, 256, 4), np.float32)
    codes = np.zeros((n, nsq), np.uint8)
    lut = np.zeros((qs.shape[0], nsq, 256), np.float32)
    dist = np.zeros((qs.shape[0], len(idv)), np.float64)
    vp = C.c_void_p
    lib.ref_pq_probe.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp, vp]
    rc = lib.ref_pq_probe(metric.encode(), data.ctypes.data, n, d, qs.ctypes.data, qs.shape[0], idv.ctypes.data,
                          len(idv), cb.ctypes.data, codes.ctypes.data, lut.ctypes.data, dist.ctypes.data)
    if rc != 0:
        raise ValueError("dim must be a multiple of 4")
    return cb, codes, lut, dist


def clustering_probe(data: np.ndarray, n_cluster: int, epochs: int, init: str, pred: np.ndarray,
                     portable: bool = False):
    """Fit the reference glass::Clustering; return (centroids[n_cluster, d],
    predict(pred) ids, pred after the in-place transform)."""
    lib = load(portable)
    data = np.ascontiguousarray(data, dtype=np.float32)
    pr = np.ascontiguousarray(pred, dtype=np.float32).copy()
    n, d = data.shape
    cent = np.zeros((n_cluster, d), np.float32)
    ids = np.zeros(pr.shape[0], np.int32)
    vp = C.c_void_p
    lib.ref_clustering_probe.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, vp, vp, C.c_int, vp]
    lib.ref_clustering_probe(data.ctypes.data, n, d, n_cluster, epochs, init.encode(), cent.ctypes.data,
                             pr.ctypes.data, pr.shape[0], ids.ctypes.data)
    return cent, ids, pr

