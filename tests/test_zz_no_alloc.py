"""glass_hip_searcher_search_batch_dev_noalloc against the allocating
glass_hip_searcher_search_batch_dev it mirrors, through the C ABI: ids,
distances and counters equal on two calls in a row (the second with fewer
queries), with a small visited table so that queries overflow into the
scratch array's bitsets, for SQ4U + FP32 and for PQ8 (whose lookup tables
also live in that array); and the argument checks."""

import ctypes as C
import os

import numpy as np
import pytest

import pyglass_amd as pg
from pyglass_amd import _lib, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


def _run(L, name, s, q, nq, k, out):
    ids, dis, cnt = out
    rc = getattr(L, name)(s._h, C.c_void_p(q.data_ptr()), nq, k, 1, C.c_void_p(ids.data_ptr()),
                          C.c_void_p(dis.data_ptr()), C.c_void_p(cnt.data_ptr()), None)
    assert rc == 0, _lib.load().glass_hip_last_error().decode()


@pytest.mark.parametrize("quant,refine", [("SQ4U", "FP32"), ("PQ8", "")])
def test_noalloc_matches_the_allocating_entry_point(golden_dir, quant, refine):
    import torch

    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, "c1_hnsw_l2.glass")), X, "L2", quant, refine)
    s.set_ef(96)
    s.set_visited_capacity(256)  # most queries move on to their wave's scratch bitset
    L = _lib.load()
    dev = torch.device("cuda:0")
    q = torch.from_numpy(Q).to(dev)
    k = 10
    for nq in (1000, 240):
        print(f"no-alloc {quant}+{refine or 'none'} nq={nq}", flush=True)
        a = (torch.full((nq, k), -7, dtype=torch.int32, device=dev), torch.zeros((nq, k), device=dev),
             torch.zeros((nq, 3), dtype=torch.int32, device=dev))
        b = (torch.full((nq, k), -9, dtype=torch.int32, device=dev), torch.zeros((nq, k), device=dev),
             torch.zeros((nq, 3), dtype=torch.int32, device=dev))
        _run(L, "glass_hip_searcher_search_batch_dev", s, q, nq, k, a)
        torch.cuda.synchronize()
        _run(L, "glass_hip_searcher_search_batch_dev_noalloc", s, q, nq, k, b)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
        assert (b[0].cpu().numpy() >= 0).all()


def test_noalloc_argument_checks(golden_dir):
    import torch

    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, "c1_hnsw_l2.glass")), X, "L2", "SQ8U")
    L = _lib.load()
    dev = torch.device("cuda:0")
    q = torch.zeros((4, 128), device=dev)
    ids = torch.zeros((4, 10), dtype=torch.int32, device=dev)
    bad_k = L.glass_hip_searcher_search_batch_dev_noalloc(s._h, C.c_void_p(q.data_ptr()), 4, 0, 1,
                                                          C.c_void_p(ids.data_ptr()), None, None, None)
    assert bad_k == -1
    s.set_ef(40000)  # a pool larger than a CU's LDS
    big = L.glass_hip_searcher_search_batch_dev_noalloc(s._h, C.c_void_p(q.data_ptr()), 4, 10, 1,
                                                        C.c_void_p(ids.data_ptr()), None, None, None)
    assert big == -1 and b"LDS" in L.glass_hip_last_error()
