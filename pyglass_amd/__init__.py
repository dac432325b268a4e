"""pyglass_amd: MI355X-native query-time graph search with pyglass's API.

`Graph` and `Searcher` mirror glass.Graph and glass.Searcher
(python/bindings.cc:40-58,90-154,242-269 of the reference): same
constructors, method names, argument meaning and return shapes.  Every call
goes through the C ABI of libglass_hip.so (include/glass_hip.h), whose HIP
kernels run the search on the GPU.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import GlassError

__all__ = ["Graph", "Searcher", "Server", "Index", "build_graph", "Clustering", "GlassError", "set_num_threads",
           "device_count", "version", "quant_fp32", "quant_bf16", "quant_fp16", "quant_e5m2", "quant_sq8", "quant_sq8u"]

_NUM_THREADS = 0


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _as_f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def set_num_threads(num_threads: int):
    """glass.set_num_threads (bindings.cc:38).  Host threads are only used by
    graph preparation; the search runs on the GPU."""
    global _NUM_THREADS
    _NUM_THREADS = int(num_threads)


def device_count() -> int:
    return int(_lib.load().glass_hip_device_count())


def version() -> str:
    return _lib.load().glass_hip_version().decode()


def device_info() -> dict:
    """Name/arch, compute units, LDS per CU (as the runtime reports it), HBM bytes, clock."""
    buf = C.create_string_buffer(256)
    cu, clk = C.c_int32(), C.c_int32()
    lds, hbm = C.c_int64(), C.c_int64()
    _lib.check(_lib.load().glass_hip_device_info(buf, 256, C.byref(cu), C.byref(lds), C.byref(hbm), C.byref(clk)))
    return {"name": buf.value.decode(errors="replace"), "cus": cu.value, "lds_per_cu": lds.value,
            "hbm_bytes": hbm.value, "clock_khz": clk.value}


class Graph:
    """glass.Graph: fixed-degree adjacency (N x K int32, -1 padded), entry
    points and optional HNSW upper layers (glass/graph.hpp:26-246)."""

    def __init__(self, filename: str | None = None, format: str = "glass"):
        self._L = _lib.load()
        self._h = C.c_void_p()
        if filename is not None:
            self.load(filename, format)

    def _free(self):
        if self._h:
            self._L.glass_hip_graph_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

    def load(self, filename: str, format: str = "glass"):
        self._free()
        h = C.c_void_p()
        _lib.check(self._L.glass_hip_graph_load(str(filename).encode(), format.encode(), C.byref(h)))
        self._h = h

    def save(self, filename: str):
        if not self._h:
            raise GlassError("empty graph")
        _lib.check(self._L.glass_hip_graph_save(self._h, str(filename).encode()))

    @classmethod
    def from_arrays(cls, adj: np.ndarray, eps=(), upper=None) -> "Graph":
        """adj: (N, K) int32, -1 padded.  upper: optional (up_K, ep, levels[N],
        lists) for HNSW (lists concatenated per node, levels[i]*up_K each)."""
        g = cls()
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        eps = np.ascontiguousarray(np.asarray(eps, dtype=np.int32).reshape(-1))
        h = C.c_void_p()
        _lib.check(g._L.glass_hip_graph_create(adj.shape[0], adj.shape[1], _p(adj), len(eps), _p(eps), C.byref(h)))
        g._h = h
        if upper is not None:
            up_k, ep, levels, lists = upper
            levels = np.ascontiguousarray(levels, dtype=np.int32)
            lists = np.ascontiguousarray(np.asarray(lists, dtype=np.int32).reshape(-1))
            _lib.check(g._L.glass_hip_graph_set_upper(h, int(up_k), int(ep), _p(levels), _p(lists)))
        return g

    def info(self):
        n, k, nep, up = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self._L.glass_hip_graph_info(self._h, C.byref(n), C.byref(k), C.byref(nep), C.byref(up)))
        return {"N": n.value, "K": k.value, "nep": nep.value, "has_upper": bool(up.value)}

    def arrays(self):
        inf = self.info()
        adj = np.empty((inf["N"], inf["K"]), np.int32)
        eps = np.empty(inf["nep"], np.int32)
        _lib.check(self._L.glass_hip_graph_copy(self._h, _p(adj), _p(eps)))
        return adj, eps

    def _take(self):
        h, self._h = self._h, C.c_void_p()
        return h


class Index:
    """glass.Index(index_type, metric, quant="FP32", R=32, L=200)
    (python/bindings.cc:60-88): CPU graph construction, "HNSW", "NSG" or
    "SSG" (NSG/SSG ignore the metric and build on L2, as create_nsg does).
    Distances are fp32 on the raw rows whatever `quant` says."""

    def __init__(self, index_type: str, metric: str, quant: str = "FP32", R: int = 32, L: int = 200):
        self.index_type, self.metric, self.quant, self.R, self.L = index_type, metric, quant, int(R), int(L)
        self.construction_time = 0.0

    def build(self, data) -> Graph:
        x = _as_f32(data)
        L = _lib.load()
        h = C.c_void_p()
        secs = C.c_double()
        _lib.check(L.glass_hip_build_graph(self.index_type.encode(), _p(x), x.shape[0], x.shape[1],
                                           self.metric.encode(), self.R, self.L or 200, _NUM_THREADS, C.byref(h),
                                           C.byref(secs)))
        self.construction_time = secs.value
        g = Graph()
        g._h = h
        return g

    def get_construction_time(self) -> float:
        return self.construction_time


def build_graph(index_type: str, input, graph_file: str, metric: str, quant: str = "FP32", R: int = 32,
                L: int = 200) -> float:
    """glass.build_graph (bindings.cc:231-237): build and save; returns seconds."""
    idx = Index(index_type, metric, quant, R, L)
    idx.build(input).save(graph_file)
    return idx.get_construction_time()


class Searcher:
    """glass.Searcher(graph, data, metric, quantizer, refine_quant="").

    The graph moves into the searcher (the Python Graph is left empty, as in
    bindings.cc:94-98); `data` is trained on, encoded on the device and may be
    freed by the caller afterwards."""

    def __init__(self, graph: Graph, data, metric: str, quantizer: str, refine_quant: str = ""):
        self._L = _lib.load()
        self._h = C.c_void_p()
        if not graph._h:
            raise GlassError("empty graph")
        h = C.c_void_p()
        _lib.check(self._L.glass_hip_searcher_create(graph._h, metric.encode(), quantizer.encode(),
                                                    refine_quant.encode(), C.byref(h)))
        graph._free()  # contents moved; release the emptied shell
        self._h = h
        x = _as_f32(data)
        if x.ndim == 1:
            x = x.reshape(1, -1)
        self.dim = x.shape[1]
        _lib.check(self._L.glass_hip_searcher_set_data(self._h, _p(x), x.shape[0], x.shape[1]))

    def __del__(self):
        try:
            if self._h:
                self._L.glass_hip_searcher_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def set_ef(self, ef: int):
        _lib.check(self._L.glass_hip_searcher_set_ef(self._h, int(ef)))

    def get_ef(self) -> int:
        return int(self._L.glass_hip_searcher_get_ef(self._h))

    def _query(self, q) -> np.ndarray:
        q = _as_f32(q)
        if q.shape[-1] != self.dim:
            raise GlassError(f"query dimension {q.shape[-1]} != data dimension {self.dim}")
        return q

    def search(self, query, k: int):
        """One query, one launch (GraphSearcher::Search semantics)."""
        q = self._query(query).reshape(-1)
        ids = np.empty(int(k), np.int32)
        dis = np.empty(int(k), np.float32)
        _lib.check(self._L.glass_hip_searcher_search(self._h, _p(q), int(k), _p(ids), _p(dis)))
        return ids, dis

    def batch_search(self, query, k: int, num_threads: int = 0):
        """All queries in one persistent launch (SearchBatch semantics)."""
        q = self._query(query)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        nq = q.shape[0]
        ids = np.empty((nq, int(k)), np.int32)
        dis = np.empty((nq, int(k)), np.float32)
        _lib.check(self._L.glass_hip_searcher_search_batch(self._h, _p(q), nq, int(k), _p(ids), _p(dis)))
        return ids, dis

    def batch_search_efk(self, query, ef, k, with_counters: bool = False):
        """Per-query ef and k in one launch.  Returns lists of id / distance
        arrays (and an (nq, 3) counter array: distance evaluations,
        expansions, upper-layer evaluations)."""
        q = self._query(query)
        nq = q.shape[0]
        ef = np.ascontiguousarray(np.broadcast_to(np.asarray(ef, np.int32), (nq,)))
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (nq,)))
        total = int(kk.sum())
        ids = np.empty(total, np.int32)
        dis = np.empty(total, np.float32)
        cnt = np.zeros((nq, 3), np.int32)
        _lib.check(self._L.glass_hip_searcher_search_batch_efk(self._h, _p(q), nq, _p(ef), _p(kk), _p(ids), _p(dis),
                                                               _p(cnt)))
        off = np.concatenate([[0], np.cumsum(kk)])
        I = [ids[off[i]:off[i + 1]] for i in range(nq)]
        D = [dis[off[i]:off[i + 1]] for i in range(nq)]
        return (I, D, cnt) if with_counters else (I, D)

    def search_batch_dev(self, q_ptr: int, nq: int, k: int, ids_ptr: int, dists_ptr: int = 0,
                         counters_ptr: int = 0, stream: int = 0, mode: int = 1):
        """Device-resident batch: all pointers are device addresses (e.g.
        torch tensor .data_ptr()); enqueued on `stream` without a host sync."""
        _lib.check(self._L.glass_hip_searcher_search_batch_dev(
            self._h, C.c_void_p(q_ptr), int(nq), int(k), int(mode), C.c_void_p(ids_ptr),
            C.c_void_p(dists_ptr or None), C.c_void_p(counters_ptr or None), C.c_void_p(stream or None)))

    def search_batch_dev_noalloc(self, q_ptr: int, nq: int, k: int, ids_ptr: int, dists_ptr: int = 0,
                                 counters_ptr: int = 0, stream: int = 0, mode: int = 1):
        """search_batch_dev without any device allocation: the library's fixed
        per-device scratch array holds the work counter and overflow bitsets.
        Calls that overlap on one GPU race (they share that array)."""
        _lib.check(self._L.glass_hip_searcher_search_batch_dev_noalloc(
            self._h, C.c_void_p(q_ptr), int(nq), int(k), int(mode), C.c_void_p(ids_ptr),
            C.c_void_p(dists_ptr or None), C.c_void_p(counters_ptr or None), C.c_void_p(stream or None)))

    def search_batch_efk_dev(self, q_ptr: int, nq: int, ef_ptr: int, k_ptr: int, off_ptr: int, ids_ptr: int,
                             max_ef: int, max_k: int, dists_ptr: int = 0, counters_ptr: int = 0, stream: int = 0):
        """Device-resident batch with per-query ef / k (int32 device arrays)
        and result offsets (int64 device array); max_ef / max_k bound every
        entry (the launch's LDS is sized for them)."""
        _lib.check(self._L.glass_hip_searcher_search_batch_efk_dev(
            self._h, C.c_void_p(q_ptr), int(nq), C.c_void_p(ef_ptr), C.c_void_p(k_ptr), C.c_void_p(off_ptr),
            int(max_ef), int(max_k), C.c_void_p(ids_ptr), C.c_void_p(dists_ptr or None),
            C.c_void_p(counters_ptr or None), C.c_void_p(stream or None)))

    def optimize(self, num_threads: int = 0):
        """Autotune the launch configuration (GraphSearcher::Optimize)."""
        _lib.check(self._L.glass_hip_searcher_optimize(self._h, int(num_threads)))

    def get_last_search_dist_cmps(self) -> float:
        return float(self._L.glass_hip_searcher_last_avg_dist_cmps(self._h))

    def set_refine_factor(self, factor: float):
        """Re-rank depth multiplier (reorder_k = int(k * factor)); 0 = the
        reference's get_refine_factor(quantizer)."""
        _lib.check(self._L.glass_hip_searcher_set_refine_factor(self._h, float(factor)))

    # ---- launch shape and introspection ---------------------------------
    def set_launch(self, team: int = 0, waves_per_block: int = 0, blocks: int = 0):
        """Force lanes-per-candidate team / waves per workgroup (0 = tuned) and
        cap the persistent grid (0 = fill the device)."""
        _lib.check(self._L.glass_hip_searcher_set_launch(self._h, int(team), int(waves_per_block), int(blocks)))

    def set_graph_capture(self, on: bool = True):
        """search_batch_dev replays a captured hipGraph when its arguments repeat."""
        _lib.check(self._L.glass_hip_searcher_set_graph_capture(self._h, int(bool(on))))

    def set_query_waves(self, waves: int = 0):
        """Waves per query: 1, or 2 / 4 waves of a workgroup sharing each
        query's distance work; 0 lets optimize() choose for small batches."""
        _lib.check(self._L.glass_hip_searcher_set_query_waves(self._h, int(waves)))

    def get_query_waves(self):
        """(waves per query of the last launch, tuned choice for small batches)."""
        a, b = C.c_int32(), C.c_int32()
        _lib.check(self._L.glass_hip_searcher_get_query_waves(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_gather(self, mode: int = -1):
        """Integer codes: 1 gathers every neighbour's code row before the
        visited test, 0 only the unvisited ones, -1 lets optimize() choose."""
        _lib.check(self._L.glass_hip_searcher_set_gather(self._h, int(mode)))

    def set_visited_capacity(self, entries: int = 0):
        """LDS visited-table entries per query (0 = tuned / sized from ef)."""
        _lib.check(self._L.glass_hip_searcher_set_visited_capacity(self._h, int(entries)))

    def kernel_info(self):
        """VGPRs, LDS bytes per workgroup and resident workgroups per CU of the last launch."""
        v, l, b = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self._L.glass_hip_searcher_kernel_info(self._h, C.byref(v), C.byref(l), C.byref(b)))
        return {"vgprs": v.value, "lds_per_block": l.value, "blocks_per_cu": b.value}

    def get_launch(self):
        """Shape of the last launch (team, waves per workgroup, workgroups,
        visited entries); before the first one, the shape at the current ef."""
        t, w, b, v = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self._L.glass_hip_searcher_get_launch(self._h, C.byref(t), C.byref(w), C.byref(b), C.byref(v)))
        return {"team": t.value, "waves_per_block": w.value, "blocks": b.value, "vis_entries": v.value}

    def info(self):
        cs, rb, deg, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        f = C.c_float()
        _lib.check(self._L.glass_hip_searcher_info(self._h, C.byref(cs), C.byref(rb), C.byref(deg), C.byref(n),
                                                   C.byref(f)))
        return {"code_size": cs.value, "rerank_row_bytes": rb.value, "degree": deg.value, "n": n.value,
                "refine_factor": f.value}

    def codes(self) -> np.ndarray:
        inf = self.info()
        out = np.empty((inf["n"], inf["code_size"]), np.uint8)
        _lib.check(self._L.glass_hip_searcher_get_codes(self._h, _p(out)))
        return out

    def calibration(self) -> np.ndarray:
        n = C.c_int32()
        _lib.check(self._L.glass_hip_searcher_get_calibration(self._h, None, C.byref(n)))
        out = np.empty(max(1, n.value), np.float32)
        _lib.check(self._L.glass_hip_searcher_get_calibration(self._h, _p(out), C.byref(n)))
        return out[: n.value]

    def sync(self):
        _lib.check(self._L.glass_hip_searcher_sync(self._h))


class Server:
    """Served form over a Searcher: requests queue from any thread and a
    dispatcher thread batches whatever is queued into one launch as soon as a
    staging slot is free (glass_hip_server_* in include/glass_hip.h); a
    request larger than max_batch is split over several batches, and the
    batches in flight are admitted by free device memory (up to
    max_inflight).  Each query is answered with SearchBatch semantics at its
    request's ef and k."""

    def __init__(self, searcher: Searcher, max_batch: int = 8192, max_k: int = 100, max_inflight: int = 2):
        self._L = _lib.load()
        self.searcher = searcher  # keeps the searcher alive while serving
        self.dim = searcher.dim
        h = C.c_void_p()
        _lib.check(self._L.glass_hip_server_create(searcher._h, int(max_batch), int(max_k), int(max_inflight),
                                                   C.byref(h)))
        self._h = h
        self._k = {}

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            self._L.glass_hip_server_free(self._h)
            self._h = None

    def submit(self, queries, k: int, ef: int) -> int:
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, self.dim)
        t = C.c_int64()
        _lib.check(self._L.glass_hip_server_submit(self._h, _p(q), q.shape[0], int(ef), int(k), C.byref(t)))
        self._k[t.value] = (q.shape[0], int(k))
        return t.value

    def poll(self, ticket: int):
        """(ids[nq, k], dists[nq, k]) when the request has completed, else None."""
        nq, k = self._k[ticket]
        ids = np.empty((nq, k), np.int32)
        dis = np.empty((nq, k), np.float32)
        done = C.c_int32()
        _lib.check(self._L.glass_hip_server_poll(self._h, int(ticket), _p(ids), _p(dis), C.byref(done)))
        if not done.value:
            return None
        del self._k[ticket]
        return ids, dis

    def wait(self, ticket: int):
        nq, k = self._k.pop(ticket)
        ids = np.empty((nq, k), np.int32)
        dis = np.empty((nq, k), np.float32)
        _lib.check(self._L.glass_hip_server_wait(self._h, int(ticket), _p(ids), _p(dis)))
        return ids, dis

    def cancel(self, ticket: int):
        _lib.check(self._L.glass_hip_server_cancel(self._h, int(ticket)))
        self._k.pop(ticket, None)

    def replay(self, queries, nq, ef, k, arrival_ns):
        """Offered-load driver: request i (nq[i] consecutive rows of `queries`)
        arrives arrival_ns[i] after the call starts.  Returns (latency_ns per
        request, ids, dists) with request i's results after those of < i."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, self.dim)
        nreq = len(arrival_ns)
        nq = np.ascontiguousarray(np.broadcast_to(np.asarray(nq, np.int32), (nreq,)))
        ef = np.ascontiguousarray(np.broadcast_to(np.asarray(ef, np.int32), (nreq,)))
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (nreq,)))
        arr = np.ascontiguousarray(arrival_ns, np.int64)
        if int(nq.sum()) > q.shape[0]:
            raise GlassError("requests name more query rows than given")
        total = int((nq.astype(np.int64) * k).sum())
        lat = np.zeros(nreq, np.int64)
        ids = np.empty(total, np.int32)
        dis = np.empty(total, np.float32)
        _lib.check(self._L.glass_hip_server_replay(self._h, _p(q), nreq, _p(nq), _p(ef), _p(k), _p(arr), _p(lat),
                                                   _p(ids), _p(dis)))
        return lat, ids, dis

    def info(self):
        """Batches in flight the device memory admitted, max_batch, max_k."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self._L.glass_hip_server_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"max_inflight": a.value, "max_batch": b.value, "max_k": c.value}

    def stats(self):
        b, q, r = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.check(self._L.glass_hip_server_stats(self._h, C.byref(b), C.byref(q), C.byref(r)))
        return {"batches": b.value, "queries": q.value, "requests": r.value}


class Clustering:
    """glass.Clustering (python/bindings.cc:155-203; glass/algorithms/clustering.hpp):
    k-means with `init` "random" (the binding's default) or "kmeans++"."""

    def __init__(self, n_cluster: int, epochs: int = 10, init: str = "random"):
        self._L = _lib.load()
        h = C.c_void_p()
        _lib.check(self._L.glass_hip_clustering_create(int(n_cluster), int(epochs), init.encode(), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.glass_hip_clustering_free(h)
            self._h = None

    def fit(self, data):
        x = _as_f32(data)
        _lib.check(self._L.glass_hip_clustering_fit(self._h, _p(x), x.shape[0], x.shape[1]))

    def predict(self, data) -> np.ndarray:
        x = _as_f32(data)
        ids = np.empty(x.shape[0], np.int32)
        _lib.check(self._L.glass_hip_clustering_predict(self._h, _p(x), x.shape[0], x.shape[1], _p(ids)))
        return ids

    def transform(self, data, inplace: bool = False):
        """With inplace=True subtracts each row's nearest centroid from `data`
        (a float32 C-contiguous array is changed in place); returns None, as
        the binding does (it does nothing without inplace)."""
        if not inplace:
            return None
        x = data if isinstance(data, np.ndarray) and data.dtype == np.float32 and data.flags.c_contiguous \
            else _as_f32(data)
        _lib.check(self._L.glass_hip_clustering_transform(self._h, _p(x), x.shape[0], x.shape[1]))
        return None

    def centroids(self) -> np.ndarray:
        """The fitted centroids, [n_cluster, dim] (not in the binding)."""
        k, d = C.c_int32(), C.c_int32()
        _lib.check(self._L.glass_hip_clustering_centroids(self._h, None, C.byref(k), C.byref(d)))
        out = np.zeros((k.value, d.value), np.float32)
        _lib.check(self._L.glass_hip_clustering_centroids(self._h, _p(out), C.byref(k), C.byref(d)))
        return out


class _QuantTrainer:
    """quant_<name>(dim).train(data) (python/bindings.cc:206-229): trains the
    quantizer on data with the L2 metric.  `params()` (not in the binding)
    returns what training found, in the layout of Searcher.calibration()."""

    _quant = ""

    def __init__(self, dim: int):
        self.dim = int(dim)
        if self.dim <= 0:
            raise GlassError(f"dim must be positive (got {dim})")
        self._params = np.zeros(0, np.float32)

    def train(self, data):
        L = _lib.load()
        x = _as_f32(data).reshape(-1, self.dim)
        n = C.c_int32()
        _lib.check(L.glass_hip_quant_train(self._quant.encode(), b"L2", None, 0, self.dim, None, C.byref(n)))
        out = np.zeros(max(1, n.value), np.float32)
        _lib.check(L.glass_hip_quant_train(self._quant.encode(), b"L2", _p(x), x.shape[0], self.dim, _p(out),
                                           C.byref(n)))
        self._params = out[: n.value]

    def params(self) -> np.ndarray:
        return self._params


class quant_fp32(_QuantTrainer):
    _quant = "FP32"


class quant_bf16(_QuantTrainer):
    _quant = "BF16"


class quant_fp16(_QuantTrainer):
    _quant = "FP16"


class quant_e5m2(_QuantTrainer):
    _quant = "E5M2"


class quant_sq8(_QuantTrainer):
    _quant = "SQ8"


class quant_sq8u(_QuantTrainer):
    _quant = "SQ8U"

