"""Ragged query batches (uneven sizes, per-batch k) through the captured
hipGraph path, checked against the oracle; the autotuner's choice keeps
results exact."""
import numpy as np
import pytest

import hora_amd
from hora_amd import BruteForceIndex, Metric
from hora_amd.datasets import bf_normal, unit_normal

pytestmark = pytest.mark.gpu


def _ragged(idx, q, sizes, ks):
    import torch

    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    tot = int(sum(s * k for s, k in zip(sizes, ks)))
    dq = torch.from_numpy(q).cuda()
    di = torch.empty(max(tot, 1), dtype=torch.int64, device="cuda")
    dd = torch.empty(max(tot, 1), dtype=torch.float32, device="cuda")
    dc = torch.empty(q.shape[0], dtype=torch.int32, device="cuda")
    idx.search_ragged_device(dq.data_ptr(), off, ks, di.data_ptr(), dd.data_ptr(),
                             dc.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return (di.cpu().numpy().view(np.uint64)[:tot], dd.cpu().numpy()[:tot],
            dc.cpu().numpy().view(np.uint32), off)


def _check(orc, base, q, sizes, ks, mt, gi, gd, gc, off):
    pos = 0
    for b, (s, k) in enumerate(zip(sizes, ks)):
        qb = q[off[b]:off[b + 1]]
        ri, rd, rc = orc.bf_search(base, qb, k, int(mt))
        assert np.array_equal(gc[off[b]:off[b + 1]], rc), f"batch {b} counts"
        for r in range(s):
            assert np.array_equal(gi[pos:pos + k], ri[r]), f"batch {b} row {r}"
            assert np.array_equal(gd[pos:pos + k], rd[r])
            pos += k


@pytest.mark.parametrize("mode", [hora_amd.HM_SEARCH_BF16_RERANK, hora_amd.HM_SEARCH_EXACT])
def test_ragged_batches_match_oracle(orc, mode):
    rng = np.random.default_rng(9)
    base = unit_normal(30_000, 96, seed=51)
    idx = BruteForceIndex(96)
    idx.add_batch(base)
    idx.build(Metric.CosineSimilarity)
    idx.set_search_mode(mode, 0)
    for call in range(2):  # the second call replays the captured graph
        sizes = rng.integers(1, 700, 9).tolist()
        ks = rng.choice([1, 10, 100], 9).tolist()
        q = unit_normal(int(sum(sizes)), 96, seed=60 + call)
        gi, gd, gc, off = _ragged(idx, q, sizes, ks)
        _check(orc, base, q, sizes, ks, Metric.CosineSimilarity, gi, gd, gc, off)


def test_ragged_chunks_above_graph_capacity(orc):
    base = bf_normal(8_000, 40, seed=71)
    idx = BruteForceIndex(40)
    idx.add_batch(base)
    idx.build(Metric.Euclidean)
    idx.set_search_mode(hora_amd.HM_SEARCH_BF16_RERANK, 128)
    sizes, ks = [4096, 1, 3000], [10, 100, 1]  # 7097 queries: two graph chunks
    q = bf_normal(sum(sizes), 40, seed=72)
    gi, gd, gc, off = _ragged(idx, q, sizes, ks)
    _check(orc, base, q, sizes, ks, Metric.Euclidean, gi, gd, gc, off)


def test_autotune_keeps_results_exact(orc):
    base = unit_normal(20_000, 64, seed=81)
    q = unit_normal(256, 64, seed=82)
    idx = BruteForceIndex(64)
    idx.add_batch(base)
    idx.build(Metric.DotProduct)
    idx.set_search_mode(hora_amd.HM_SEARCH_BF16_RERANK, 0)
    sizes, ks = [100, 156], [10, 10]
    gi0, gd0, gc0, off0 = _ragged(idx, q, sizes, ks)  # captures a graph before tuning
    s = idx.autotune(q, 10)
    assert 64 <= s <= 4096
    t = idx.tuning()
    assert t["shortlist"] == s and t["score_group_kib"] in (512, 1024, 2048, 3072)
    gi, gd, gc = idx.search_batch(q, 10)
    ri, rd, rc = orc.bf_search(base, q, 10, orc.DOT)
    assert np.array_equal(gi, ri) and np.array_equal(gd, rd)
    # the ragged graphs are rebuilt for the tuned tile order, same results
    gi1, gd1, gc1, off1 = _ragged(idx, q, sizes, ks)
    _check(orc, base, q, sizes, ks, Metric.DotProduct, gi1, gd1, gc1, off1)


def _tied_base(n, d, seed):
    # small integers in few dims: many equal distances
    rng = np.random.default_rng(seed)
    return rng.integers(0, 4, (n, d)).astype(np.float32)


@pytest.mark.parametrize("mode,mt", [(hora_amd.HM_SEARCH_EXACT, Metric.Manhattan),
                                     (hora_amd.HM_SEARCH_BF16_RERANK, Metric.Euclidean)])
def test_ragged_tied_rows_use_their_own_k(orc, mode, mt):
    # a tied row's result depends on k (the heap's capacity), so batches
    # with k 1, 10 and 100 over the same tied data must each match the
    # oracle run with that batch's k
    base = _tied_base(6_000, 6, seed=91)
    q = _tied_base(300, 6, seed=92)
    idx = BruteForceIndex(6)
    idx.add_batch(base)
    idx.build(mt)
    idx.set_search_mode(mode, 0)
    sizes, ks = [100, 100, 100], [100, 1, 10]
    gi, gd, gc, off = _ragged(idx, q, sizes, ks)
    _check(orc, base, q, sizes, ks, mt, gi, gd, gc, off)


@pytest.mark.parametrize("mode", [hora_amd.HM_SEARCH_EXACT, hora_amd.HM_SEARCH_BF16_RERANK])
def test_ragged_errors_are_statuses(mode):
    from hora_amd import HoraError

    base = bf_normal(2_000, 16, seed=93)
    idx = BruteForceIndex(16)
    idx.add_batch(base)
    idx.build(Metric.Euclidean)
    idx.set_search_mode(mode, 0)
    q = bf_normal(40, 16, seed=94)
    q[27, 5] = np.nan
    with pytest.raises(HoraError) as e:
        _ragged(idx, q, [20, 20], [10, 1])
    assert e.value.status == 3
    with pytest.raises(HoraError) as e:  # the wrapper checks the offsets' length
        idx.search_ragged_device(1, [0, 20], [10, 1], 1, 1, 1)
    assert e.value.status == 1
    with pytest.raises(HoraError) as e:  # batch_offsets[0] must be 0
        idx.search_ragged_device(1, [5, 20], [10], 1, 1, 1)
    assert e.value.status == 1
    cos = BruteForceIndex(16)
    zb = base.copy()
    zb[11] = 0.0  # a zero-norm base row: the Cosine distance is NaN
    cos.add_batch(zb)
    cos.build(Metric.CosineSimilarity)
    cos.set_search_mode(mode, 0)
    with pytest.raises(HoraError) as e:
        _ragged(cos, bf_normal(30, 16, seed=95), [10, 20], [5, 3])
    assert e.value.status == 8
