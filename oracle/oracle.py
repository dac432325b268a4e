"""ctypes front end of the oracle (oracle/glass_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg use it as the checker of the HIP path.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KINDS = {"FP32": 0, "FP16": 1, "BF16": 2, "E5M2": 3, "SQ8U": 4, "SQ8": 5, "SQ6": 6, "SQ4U": 7, "SQ4": 9, "SQ2U": 11, "PQ8": 13}
METRICS = {"L2": 0, "IP": 1}


class _Quant(C.Structure):
    _fields_ = [("kind", C.c_int32), ("metric", C.c_int32), ("dim", C.c_int32), ("dim_align", C.c_int32),
                ("code_size", C.c_int32), ("nbits", C.c_int32), ("align", C.c_int32), ("min", C.c_float),
                ("dif", C.c_float), ("max", C.c_float), ("mins", C.POINTER(C.c_float)),
                ("difs", C.POINTER(C.c_float)), ("codebook", C.POINTER(C.c_float))]


class _Graph(C.Structure):
    _fields_ = [("N", C.c_int32), ("K", C.c_int32), ("adj", C.POINTER(C.c_int32)), ("nep", C.c_int32),
                ("eps", C.POINTER(C.c_int32)), ("has_init", C.c_int32), ("init_K", C.c_int32), ("ep", C.c_int32),
                ("levels", C.POINTER(C.c_int32)), ("offsets", C.POINTER(C.c_int64)),
                ("lists", C.POINTER(C.c_int32))]


class _Counters(C.Structure):
    _fields_ = [("dist_cmps", C.c_int64), ("expansions", C.c_int64), ("upper_evals", C.c_int64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.orq_quant_init.argtypes = [C.POINTER(_Quant), C.c_int32, C.c_int32, C.c_int32]
    L.orq_quant_free.argtypes = [C.POINTER(_Quant)]
    L.orq_quant_train.argtypes = [C.POINTER(_Quant), vp, C.c_int64]
    L.orq_encode_rows.argtypes = [C.POINTER(_Quant), vp, C.c_int64, vp]
    L.orq_query_size.argtypes = [C.POINTER(_Quant)]
    L.orq_query_size.restype = C.c_int64
    L.orq_prepare_query.argtypes = [C.POINTER(_Quant), vp, vp]
    L.orq_distance.argtypes = [C.POINTER(_Quant), vp, vp]
    L.orq_distance.restype = C.c_double
    L.orq_distance_ctx.argtypes = [C.POINTER(_Quant), vp, vp, C.c_int32]
    L.orq_distance_ctx.restype = C.c_double
    L.orq_graph_load.argtypes = [C.POINTER(_Graph), C.c_char_p]
    L.orq_graph_from_arrays.argtypes = [C.POINTER(_Graph), C.c_int32, C.c_int32, vp, C.c_int32, vp]
    L.orq_graph_free.argtypes = [C.POINTER(_Graph)]
    L.orq_search.argtypes = [C.POINTER(_Graph), C.POINTER(_Quant), vp, vp, C.c_int32, C.c_int32, C.c_int32, vp, vp,
                             C.POINTER(_Counters)]
    L.orq_search_refine.argtypes = [C.POINTER(_Graph), C.POINTER(_Quant), vp, C.POINTER(_Quant), vp, C.c_float, vp,
                                    C.c_int32, C.c_int32, C.c_int32, vp, vp, C.POINTER(_Counters)]
    L.orq_refine_factor.argtypes = [C.c_int32]
    L.orq_refine_factor.restype = C.c_float
    L.orq_mt19937_outputs.argtypes = [C.c_int32, vp]
    L.orq_kmeans_fit4.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp]
    L.orq_pq_lut_sum.argtypes = [vp, vp, C.c_int32]
    L.orq_clustering_fit.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]
    L.orq_clustering_predict.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_int32, vp]
    L.orq_clustering_transform.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_int32]
    L.orq_pq_lut_sum.restype = C.c_float
    L.orq_search_many.argtypes = [C.POINTER(_Graph), C.POINTER(_Quant), vp, C.POINTER(_Quant), vp, C.c_float, vp,
                                  C.c_int64, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, C.c_int32]
    L.orq_pool_run.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp]
    L.orq_f32_to_f16.argtypes = [C.c_float]
    L.orq_f32_to_f16.restype = C.c_uint16
    L.orq_f16_to_f32.argtypes = [C.c_uint16]
    L.orq_f16_to_f32.restype = C.c_float
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Quant:
    """Calibrated quantizer + codes (GraphSearcher::SetData, graph_searcher.hpp:51-60)."""

    def __init__(self, kind: str, metric: str, dim: int):
        self.L = lib()
        self.s = _Quant()
        self.kind, self.metric = kind, metric
        if self.L.orq_quant_init(C.byref(self.s), KINDS[kind], METRICS[metric], int(dim)) != 0:
            raise ValueError(kind)
        self.codes = None

    def __del__(self):
        if getattr(self, "L", None) is not None:
            self.L.orq_quant_free(C.byref(self.s))

    @property
    def code_size(self):
        return self.s.code_size

    @property
    def dim_align(self):
        return self.s.dim_align

    def params(self) -> np.ndarray:
        """Same layout as refwrap.quant_probe's params."""
        da = self.s.dim_align
        out = np.zeros(2 * da + 2, np.float32)
        if self.kind in ("SQ8U", "SQ4U", "SQ2U"):
            if self.kind == "SQ8U" and self.metric == "IP":
                out[0] = self.s.max
            else:
                out[0], out[1] = self.s.min, self.s.dif
        elif self.kind in ("SQ8", "SQ6", "SQ4"):
            out[:da] = np.ctypeslib.as_array(self.s.mins, (da,))
            out[da:2 * da] = np.ctypeslib.as_array(self.s.difs, (da,))
        return out

    def codebook(self) -> np.ndarray:
        """PQ8 centroids, [nsq, 256, 4]."""
        nsq = self.s.code_size
        return np.ctypeslib.as_array(self.s.codebook, (nsq * 256 * 4,)).reshape(nsq, 256, 4).copy()

    def lut(self, x: np.ndarray) -> np.ndarray:
        """PQ8 query lookup table (ProductQuant::get_lut), [nsq, 256]."""
        nsq = self.s.code_size
        qb = np.zeros(nsq * 256, np.float32)
        xx = np.ascontiguousarray(x, np.float32)
        self.L.orq_prepare_query(C.byref(self.s), _p(xx), _p(qb))
        return qb.reshape(nsq, 256)

    def train(self, data: np.ndarray):
        data = np.ascontiguousarray(data, np.float32)
        self.L.orq_quant_train(C.byref(self.s), _p(data), data.shape[0])

    def encode(self, data: np.ndarray) -> np.ndarray:
        data = np.ascontiguousarray(data, np.float32)
        codes = np.zeros((data.shape[0], self.s.code_size), np.uint8)
        self.L.orq_encode_rows(C.byref(self.s), _p(data), data.shape[0], _p(codes))
        return codes

    def fit(self, data: np.ndarray) -> "Quant":
        self.train(data)
        self.codes = self.encode(data)
        return self

    def prepare_query(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        buf = np.zeros(self.L.orq_query_size(C.byref(self.s)), np.uint8)
        self.L.orq_prepare_query(C.byref(self.s), _p(x), _p(buf))
        return buf

    CTX = {"computer": 0, "search": 1, "batch": 2}

    def distances(self, x: np.ndarray, ids, ctx: str = "computer") -> np.ndarray:
        """Distances as the reference's computer evaluates them; ctx names the
        calling context whose float contraction is followed ("computer",
        "search" = Search's inlined loop, "batch" = SearchBatch's)."""
        qb = self.prepare_query(x)
        out = np.empty(len(ids), np.float64)
        c = self.CTX[ctx]
        for j, i in enumerate(ids):
            out[j] = self.L.orq_distance_ctx(C.byref(self.s), _p(qb), _p(self.codes[int(i)]), c)
        return out


class Graph:
    def __init__(self, path: str | None = None, adj: np.ndarray | None = None, eps=None):
        self.L = lib()
        self.s = _Graph()
        if path is not None:
            rc = self.L.orq_graph_load(C.byref(self.s), path.encode())
            if rc != 0:
                raise IOError(f"cannot load {path} ({rc})")
        else:
            adj = np.ascontiguousarray(adj, np.int32)
            eps = np.ascontiguousarray(np.asarray(eps, np.int32))
            self.L.orq_graph_from_arrays(C.byref(self.s), adj.shape[0], adj.shape[1], _p(adj), len(eps), _p(eps))

    def __del__(self):
        if getattr(self, "L", None) is not None:
            self.L.orq_graph_free(C.byref(self.s))

    @property
    def N(self):
        return self.s.N

    @property
    def K(self):
        return self.s.K


class Searcher:
    """Oracle counterpart of glass.Searcher (python/bindings.cc:90-154)."""

    def __init__(self, graph: Graph, data: np.ndarray, metric: str, quantizer: str, refine_quant: str = "",
                 refine_factor: float | None = None):
        self.L = lib()
        self.graph = graph
        self.dim = data.shape[1]
        self.q = Quant(quantizer, metric, self.dim).fit(data)
        self.rq = Quant(refine_quant, metric, self.dim).fit(data) if refine_quant else None
        self.factor = self.L.orq_refine_factor(KINDS[quantizer]) if refine_quant else 1.0
        if refine_quant and refine_factor:
            self.factor = float(refine_factor)
        self.ef = 32
        self.last_dist_cmps = 0.0

    def set_ef(self, ef: int):
        self.ef = int(ef)

    def _run(self, qs, k, ef, batch_mode, nthreads=0):
        qs = np.ascontiguousarray(qs, np.float32).reshape(-1, self.dim)
        nq = qs.shape[0]
        ids = np.zeros((nq, k), np.int32)
        dis = np.zeros((nq, k), np.float32)
        dc = np.zeros(nq, np.int64)
        ex = np.zeros(nq, np.int64)
        rq = C.byref(self.rq.s) if self.rq else None
        rcodes = _p(self.rq.codes) if self.rq else None
        self.L.orq_search_many(C.byref(self.graph.s), C.byref(self.q.s), _p(self.q.codes), rq, rcodes,
                               self.factor, _p(qs), nq, int(k), int(ef), int(batch_mode), _p(ids), _p(dis), _p(dc),
                               _p(ex), int(nthreads))
        return ids, dis, dc, ex

    def search(self, q, k: int):
        ids, dis, dc, _ = self._run(q, k, self.ef, 0)
        return ids[0], dis[0]

    def batch_search(self, qs, k: int, num_threads: int = 0, ef=None):
        ids, dis, dc, ex = self._run(qs, k, self.ef if ef is None else ef, 1, num_threads)
        self.last_dist_cmps = float(dc.mean()) if len(dc) else 0.0
        self.last_counts = (dc, ex)
        return ids, dis

    def get_last_search_dist_cmps(self) -> float:
        return self.last_dist_cmps


def pool_run(cap: int, ef: int, ops, dists):
    """Drives one oracle LinearPool: ops[i] >= 0 inserts (ops[i], dists[i]),
    ops[i] < 0 pops when has_next() holds.  Returns (popped per op, final ids
    with the checked bit, final dists, cur)."""
    L = lib()
    ops = np.ascontiguousarray(ops, np.int32)
    dists = np.ascontiguousarray(dists, np.float64)
    popped = np.zeros(len(ops), np.int32)
    ids = np.zeros(cap, np.int32)
    dd = np.zeros(cap, np.float64)
    size = C.c_int32()
    cur = C.c_int32()
    if L.orq_pool_run(cap, ef, len(ops), _p(ops), _p(dists), _p(popped), _p(ids), _p(dd), C.byref(size),
                      C.byref(cur)) != 0:
        raise ValueError("bad pool arguments")
    return popped, ids[:size.value], dd[:size.value], cur.value


def f32_to_f16(x: float) -> int:
    return int(lib().orq_f32_to_f16(float(x)))


def f16_to_f32(h: int) -> float:
    return float(lib().orq_f16_to_f32(int(h)))


def mt19937_outputs(count: int) -> np.ndarray:
    """The first `count` outputs of a default-seeded std::mt19937 (seed 5489)."""
    out = np.zeros(count, np.uint32)
    lib().orq_mt19937_outputs(count, _p(out))
    return out


def kmeans_fit4(rows: np.ndarray, k: int = 256, epochs: int = 10) -> np.ndarray:
    """glass::Clustering(k, epochs).fit over rows of 4 floats; [k, 4] centroids."""
    r = np.ascontiguousarray(rows, np.float32)
    out = np.zeros((k, 4), np.float32)
    lib().orq_kmeans_fit4(_p(r), r.shape[0], k, epochs, _p(out))
    return out


def pq_lut_sum(lut: np.ndarray, code: np.ndarray) -> float:
    """ProductQuant::dist_func over one code row, in the reference's order."""
    lt = np.ascontiguousarray(lut, np.float32)
    c = np.ascontiguousarray(code, np.uint8)
    return float(lib().orq_pq_lut_sum(_p(lt), _p(c), c.size))


class Clustering:
    """glass::Clustering as the Python binding drives it (python/bindings.cc:155-203)."""

    def __init__(self, n_cluster: int, epochs: int = 10, init: str = "random"):
        if init not in ("random", "kmeans++"):
            raise ValueError(init)
        self.k, self.epochs, self.kpp = n_cluster, epochs, init == "kmeans++"
        self.centroids = None

    def fit(self, data):
        x = np.ascontiguousarray(data, np.float32)
        self.centroids = np.zeros((self.k, x.shape[1]), np.float32)
        lib().orq_clustering_fit(_p(x), x.shape[0], x.shape[1], self.k, self.epochs, int(self.kpp),
                                 _p(self.centroids))
        return self

    def predict(self, data) -> np.ndarray:
        x = np.ascontiguousarray(data, np.float32)
        ids = np.zeros(x.shape[0], np.int32)
        lib().orq_clustering_predict(_p(self.centroids), self.k, x.shape[1], _p(x), x.shape[0], _p(ids))
        return ids

    def transform(self, data: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(data, np.float32).copy()
        lib().orq_clustering_transform(_p(self.centroids), self.k, x.shape[1], _p(x), x.shape[0])
        return x

