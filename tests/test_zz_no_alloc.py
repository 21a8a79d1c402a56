"""hm_index_search_device_noalloc (the entry point that allocates nothing)
against hm_index_search_device in the exact form: ids, distances and
counts bit-identical, on two calls in a row, the second with fewer
queries, a smaller k and another metric; and its size limit."""
import numpy as np
import pytest

import hora_amd
from hora_amd import BruteForceIndex, HoraError, Metric
from hora_amd.datasets import bf_normal, bf_sparse_u8

pytestmark = pytest.mark.gpu


def _run(idx, q, k, noalloc):
    import torch

    nq = q.shape[0]
    dq = torch.from_numpy(q).cuda()
    di = torch.full((nq, max(k, 1)), -7, dtype=torch.int64, device="cuda")
    dd = torch.full((nq, max(k, 1)), -7.0, dtype=torch.float32, device="cuda")
    dc = torch.full((nq,), -7, dtype=torch.int32, device="cuda")
    fn = idx.search_device_noalloc if noalloc else idx.search_device
    fn(dq.data_ptr(), nq, k, di.data_ptr(), dd.data_ptr(), dc.data_ptr(),
       torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return (di.cpu().numpy().view(np.uint64)[:, :k], dd.cpu().numpy()[:, :k],
            dc.cpu().numpy().view(np.uint32))


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y)


def test_noalloc_matches_search_device_on_two_calls(orc):
    # 0-255 integer rows (many tied distances) at d=784: the 32 MiB scratch
    # holds about 130 queries of 60,000-float distance rows, so the first
    # call runs in several chunks
    base = bf_sparse_u8(60_000, 784, seed=201)
    q = bf_sparse_u8(400, 784, seed=202)
    for mt, calls in ((Metric.Euclidean, [(400, 10), (37, 3)]),
                      (Metric.CosineSimilarity, [(300, 50), (5, 1)]),
                      (Metric.Manhattan, [(200, 10), (20, 100)])):
        idx = BruteForceIndex(784)
        idx.add_batch(base)
        idx.build(mt)
        idx.set_search_mode(hora_amd.HM_SEARCH_FP8_RERANK, 0)  # the entry ignores it
        for nq, k in calls:
            got = _run(idx, q[:nq], k, True)
            idx.set_search_mode(hora_amd.HM_SEARCH_EXACT, 0)
            want = _run(idx, q[:nq], k, False)
            idx.set_search_mode(hora_amd.HM_SEARCH_FP8_RERANK, 0)
            _same(got, want)
        ri, rd, rc = orc.bf_search(base, q[:5], calls[-1][1], int(mt))
        assert np.array_equal(got[0][:5], ri) and np.array_equal(got[2][:5], rc)


def test_noalloc_small_cases_and_errors():
    base = bf_normal(500, 20, seed=203)
    idx = BruteForceIndex(20)
    idx.add_batch(base)
    idx.build(Metric.DotProduct)
    q = bf_normal(9, 20, seed=204)
    _same(_run(idx, q, 600, True), _run(idx, q, 600, False))  # k > n
    gi, gd, gc = _run(idx, q, 0, True)                        # k == 0
    assert (gc == 0).all()
    bad = q.copy()
    bad[4, 2] = np.inf
    with pytest.raises(HoraError) as e:
        _run(idx, bad, 5, True)
    assert e.value.status == 3
    _same(_run(idx, q, 5, True), _run(idx, q, 5, False))     # the error word was cleared


def test_noalloc_rejects_an_index_beyond_its_scratch():
    big = BruteForceIndex(1)
    big.add_batch(np.arange(8_500_000, dtype=np.float32).reshape(-1, 1))  # 34 MB per distance row
    big.build(Metric.Euclidean)
    with pytest.raises(HoraError) as e:
        _run(big, np.zeros((2, 1), np.float32), 3, True)
    assert e.value.status == 6  # HM_ERR_CAPACITY
    assert "scratch" in e.value.message
