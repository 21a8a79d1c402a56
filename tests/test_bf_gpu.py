"""BruteForceIndex on the GPU vs the oracle: ids bit-exact, distances
bit-equal (the exact path reproduces the reference's f32 arithmetic), ties
resolved exactly as the reference's std-BinaryHeap does."""
import numpy as np
import pytest

import hora_amd
from hora_amd import BruteForceIndex, HoraError, Metric
from hora_amd.datasets import bf_normal, bf_sparse_u8

pytestmark = pytest.mark.gpu

METRICS = [Metric.Euclidean, Metric.Manhattan, Metric.DotProduct, Metric.CosineSimilarity,
           Metric.Angular]


def _index(base, mt):
    idx = BruteForceIndex(base.shape[1])
    idx.add_batch(base)
    idx.build(mt)
    return idx


def _check(orc, base, q, k, mt, idx=None):
    idx = idx or _index(base, mt)
    gi, gd, gc = idx.search_batch(q, k)
    ri, rd, rc = orc.bf_search(base, q, k, int(mt))
    assert np.array_equal(gc, rc)
    for r in range(q.shape[0]):
        m = int(rc[r])
        assert gi[r, :m].tolist() == ri[r, :m].tolist(), f"row {r} ids"
        # bit-equal up to the sign of a zero
        assert np.array_equal(gd[r, :m], rd[r, :m]), f"row {r} distances"
        assert (gi[r, m:] == np.iinfo(np.uint64).max).all()
    return gi, gd


def test_config1_euclidean_full(orc):
    base = bf_normal(10_000, 64, seed=101)
    q = bf_normal(1_000, 64, seed=102)
    _check(orc, base, q, 10, Metric.Euclidean)


@pytest.mark.parametrize("mt", METRICS, ids=lambda m: m.name)
def test_sparse_u8_784_all_metrics(orc, mt):
    base = bf_sparse_u8(6_000, 784, seed=201)
    q = bf_sparse_u8(200, 784, seed=202)
    _check(orc, base, q, 10, mt)


@pytest.mark.parametrize("mt", [Metric.Euclidean, Metric.Manhattan, Metric.DotProduct,
                                Metric.CosineSimilarity], ids=lambda m: m.name)
def test_sparse_u8_full_base(orc, mt):
    base = bf_sparse_u8(60_000, 784, seed=301)
    q = bf_sparse_u8(128, 784, seed=302)
    _check(orc, base, q, 10, mt)


@pytest.mark.parametrize("k", [1, 5, 100, 300])
def test_ties_replay_the_std_heap(orc, k):
    # every base row appears three times: each query's distances come in
    # equal triples, so every row needs the heap replay
    rng = np.random.default_rng(7)
    uniq = rng.integers(0, 4, (400, 6)).astype(np.float32)
    base = np.repeat(uniq, 3, axis=0)
    rng.shuffle(base)
    q = rng.integers(0, 4, (64, 6)).astype(np.float32)
    for mt in (Metric.Euclidean, Metric.Manhattan, Metric.DotProduct):
        _check(orc, base, q, k, mt)


@pytest.mark.parametrize("k", [63, 64, 200])
def test_ties_replay_long_rows(orc, k):
    # 30,000 rows in equal triples: many groups of the row to visit; k = 63
    # runs the group-minimum replay (k + 1 <= 64), 64 and 200 the streamed
    # LDS replay
    rng = np.random.default_rng(11)
    uniq = rng.integers(0, 6, (10_000, 6)).astype(np.float32)
    base = np.repeat(uniq, 3, axis=0)
    rng.shuffle(base)
    q = rng.integers(0, 6, (48, 6)).astype(np.float32)
    for mt in (Metric.Euclidean, Metric.Manhattan):
        _check(orc, base, q, k, mt)


def test_ties_replay_rows_past_4m(orc):
    # 4.3 million rows: more than 16,384 groups of 256 values, so each group
    # minimum covers two blocks of the row.  Continuous values, except two
    # copies of each query's nearest row far apart in the row: every query is
    # tied at its head, and the replay visits few groups
    rng = np.random.default_rng(13)
    base = rng.standard_normal((4_300_000, 4), dtype=np.float32)
    q = rng.standard_normal((3, 4), dtype=np.float32)
    for r in range(3):
        near = q[r] + np.float32(1e-3)
        base[1000 + r] = near
        base[4_200_000 + r] = near
    gi, _ = _check(orc, base, q, 6, Metric.Euclidean)
    assert [sorted(gi[r, :2].tolist()) for r in range(3)] == \
        [[1000 + r, 4_200_000 + r] for r in range(3)]


@pytest.mark.parametrize("mt", [Metric.Angular, Metric.CosineSimilarity], ids=lambda m: m.name)
@pytest.mark.parametrize("k", [10, 40, 63, 100, 200, 600])  # register heap to 63, LDS heap above
def test_ties_inside_the_heap(orc, mt, k):
    # every direction four times (x, 2x, 4x, 8x: the same cosine to the
    # last bit), so a full heap holds equal distances and its layout is not
    # stable: runs of elements above the root take the round-trip map F^j
    # (until stable, or by its period), and the result must still be the
    # reference's, in the exact form and in the FP8 form's rerun
    half = bf_sparse_u8(3_000, 784, seed=481)
    base = np.concatenate([half, 2 * half, 4 * half, 8 * half]).astype(np.float32)
    q = bf_sparse_u8(20, 784, seed=482)
    _, rd1, _ = orc.bf_search(base, q, k + 1, int(mt))
    assert np.all(np.any(rd1[:, 1:] == rd1[:, :-1], axis=1))  # every row tied
    idx = _index(base, mt)
    _check(orc, base, q, k, mt, idx)
    idx.set_search_mode(hora_amd.HM_SEARCH_FP8_RERANK, 0)
    _check(orc, base, q, k, mt, idx)
    assert idx.search_stats()[0] == q.shape[0]


def test_hand_traced_tie_case(orc):
    # distances [1,1,1], k=2: the reference returns ids [2, 1]
    base = np.array([[1.0], [1.0], [1.0]], np.float32)
    idx = _index(base, Metric.Euclidean)
    gi, gd, gc = idx.search_batch(np.zeros((1, 1), np.float32), 2)
    assert gi[0].tolist() == [2, 1] and gd[0].tolist() == [1.0, 1.0]


@pytest.mark.parametrize("d", [1, 10, 17, 100])
def test_unpadded_dimensions(orc, d):
    base = bf_normal(777, d, seed=d)
    q = bf_normal(33, d, seed=d + 1)
    for mt in METRICS:
        _check(orc, base, q, 7, mt)


def test_edge_sizes(orc):
    base = bf_normal(5, 8, seed=1)
    q = bf_normal(3, 8, seed=2)
    _check(orc, base, q, 10, Metric.Euclidean)   # k > n
    _check(orc, base[:1], q, 3, Metric.Euclidean)  # n == 1
    idx = _index(base, Metric.Euclidean)
    gi, gd, gc = idx.search_batch(q, 0)            # k == 0
    assert gc.tolist() == [0, 0, 0]
    empty = BruteForceIndex(8)
    empty.build(Metric.Euclidean)
    gi, gd, gc = empty.search_batch(q, 4)
    assert gc.tolist() == [0, 0, 0] and np.isinf(gd).all()


def test_search_api_and_ids(orc):
    base = bf_normal(300, 16, seed=3)
    idx = BruteForceIndex(16)
    for i, v in enumerate(base):
        idx.add(v, f"id{i}")
    idx.build(Metric.Euclidean)
    ri, rd, _ = orc.bf_search(base, base[7:8], 4, orc.EUCLIDEAN)
    assert idx.search(base[7], 4) == [f"id{int(i)}" for i in ri[0]]
    nodes = idx.search_nodes(base[7], 4)
    assert nodes[0][0].idx() == "id7" and nodes[0][1] == 0.0
    assert np.array_equal(nodes[0][0].vectors(), base[7])
    assert idx.nodes_size() == 300 and idx.dimension() == 16 and idx.name() == "BruteForceIndex"
    with pytest.raises(HoraError):
        idx.search(base[7][:5], 4)


def test_errors_are_statuses():
    base = bf_normal(50, 8, seed=4)
    idx = BruteForceIndex(8)
    idx.add_batch(base)
    with pytest.raises(HoraError) as e:  # search before build: Metric::Unknown
        idx.search_batch(base[:1], 3)
    assert e.value.status == 4 and "unknown method" in e.value.message
    idx.build(Metric.Euclidean)
    bad = base[:2].copy()
    bad[1, 3] = np.nan
    with pytest.raises(HoraError) as e:
        idx.search_batch(bad, 3)
    assert e.value.status == 3
    with pytest.raises(HoraError) as e:
        idx.add_batch(bad)
    assert e.value.status == 3 and idx.nodes_size() == 50
    cos = BruteForceIndex(8)
    cos.add_batch(base)
    cos.build(Metric.CosineSimilarity)
    with pytest.raises(HoraError) as e:  # zero query: 0/0 NaN distance
        cos.search_batch(np.zeros((1, 8), np.float32), 3)
    assert e.value.status == 8


def test_search_device_matches_host_entry(orc):
    import torch

    base = bf_sparse_u8(3000, 784, seed=9)
    q = bf_sparse_u8(100, 784, seed=10)
    idx = _index(base, Metric.Euclidean)
    hi, hd, hc = idx.search_batch(q, 10)
    dq = torch.from_numpy(q).cuda()
    di = torch.empty((100, 10), dtype=torch.int64, device="cuda")
    dd = torch.empty((100, 10), dtype=torch.float32, device="cuda")
    dc = torch.empty((100,), dtype=torch.int32, device="cuda")
    idx.search_device(dq.data_ptr(), 100, 10, di.data_ptr(), dd.data_ptr(), dc.data_ptr(),
                      torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(di.cpu().numpy().view(np.uint64), hi)
    assert np.array_equal(dd.cpu().numpy(), hd)
    assert np.array_equal(dc.cpu().numpy().view(np.uint32), hc)


def test_add_after_build_and_clear(orc):
    base = bf_normal(400, 12, seed=5)
    q = bf_normal(10, 12, seed=6)
    idx = BruteForceIndex(12)
    idx.add_batch(base[:100])
    idx.build(Metric.CosineSimilarity)
    idx.search_batch(q, 5)
    idx.add_batch(base[100:])  # derived norms must be refreshed
    _check(orc, base, q, 5, Metric.CosineSimilarity, idx=idx)
    idx.clear()
    assert idx.nodes_size() == 0


@pytest.mark.parametrize("mode", [hora_amd.HM_SEARCH_EXACT, hora_amd.HM_SEARCH_BF16_RERANK])
def test_k_beyond_select_limit_on_small_index(orc, mode):
    # bruteforce_idx.rs:47-73 returns min(k, n) results for any k
    base = bf_normal(100, 8, seed=96)
    q = bf_normal(3, 8, seed=97)
    idx = _index(base, Metric.Euclidean)
    idx.set_search_mode(mode, 0)
    gi, gd, gc = idx.search_batch(q, 9000)
    ri, rd, rc = orc.bf_search(base, q, 9000, orc.EUCLIDEAN)
    assert np.array_equal(gc, rc) and (gc == 100).all()
    assert np.array_equal(gi[:, :100], ri[:, :100]) and np.array_equal(gd[:, :100], rd[:, :100])
    assert (gi[:, 100:] == np.iinfo(np.uint64).max).all() and np.isinf(gd[:, 100:]).all()
