"""BruteForceIndex shortlist forms (BF16 / FP8 MFMA shortlist + exact FP32
rerank) vs the oracle.

BF16 is certified: every returned row set is provably the reference's, so
ids and distances must be bit-exact.  FP8 is not certified: its recall@10 is
measured against the oracle (tolerance written below) and every distance it
returns must still be the exact reference value for that pair."""
import numpy as np
import pytest

import hora_amd
from hora_amd import BruteForceIndex, Metric
from hora_amd.datasets import bf_normal, bf_sparse_u8, unit_normal

pytestmark = pytest.mark.gpu

# Angular runs the same score kernels as Cosine: its instantiations are the
# Cosine machine code with the products' signs flipped.
GEMM_METRICS = [Metric.Euclidean, Metric.DotProduct, Metric.CosineSimilarity, Metric.Angular]


def _index(base, mt, mode, shortlist=0):
    idx = BruteForceIndex(base.shape[1])
    idx.add_batch(base)
    idx.build(mt)
    idx.set_search_mode(mode, shortlist)
    return idx


def _exact_check(orc, idx, base, q, k, mt):
    gi, gd, gc = idx.search_batch(q, k)
    ri, rd, rc = orc.bf_search(base, q, k, int(mt))
    assert np.array_equal(gc, rc)
    for r in range(q.shape[0]):
        m = int(rc[r])  # the reference returns min(k, n) results; the rest is padding
        assert np.array_equal(gi[r, :m], ri[r, :m]), f"row {r} ids"
        assert np.array_equal(gd[r, :m], rd[r, :m]), f"row {r} distances"
        assert (gi[r, m:] == np.iinfo(np.uint64).max).all()
        assert np.isinf(gd[r, m:]).all()
    return idx.search_stats()


@pytest.mark.parametrize("mt", GEMM_METRICS, ids=lambda m: Metric(m).name)
def test_bf16_exact_on_u8_data(orc, mt):
    # integers 0..255 are exact in bf16: certification is tight
    base = bf_sparse_u8(20_000, 784, seed=401)
    q = bf_sparse_u8(256, 784, seed=402)
    idx = _index(base, mt, hora_amd.HM_SEARCH_BF16_RERANK)
    fallback, uncert = _exact_check(orc, idx, base, q, 10, mt)
    assert uncert <= 256 // 20, (fallback, uncert)


@pytest.mark.parametrize("mt", GEMM_METRICS, ids=lambda m: Metric(m).name)
def test_bf16_exact_on_float_data(orc, mt):
    base = bf_normal(10_000, 64, seed=411)
    q = bf_normal(300, 64, seed=412)
    idx = _index(base, mt, hora_amd.HM_SEARCH_BF16_RERANK, shortlist=256)
    _exact_check(orc, idx, base, q, 10, mt)


def test_bf16_small_shortlist_falls_back_exactly(orc):
    # a shortlist of k+1 on rounded data rarely certifies: the exact path
    # takes those rows and the answer is still the reference's
    base = bf_normal(5_000, 48, seed=421)
    q = bf_normal(128, 48, seed=422)
    idx = _index(base, Metric.Euclidean, hora_amd.HM_SEARCH_BF16_RERANK, shortlist=11)
    fallback, uncert = _exact_check(orc, idx, base, q, 10, Metric.Euclidean)
    assert fallback >= uncert > 0


def test_bf16_ties_and_manhattan(orc):
    rng = np.random.default_rng(3)
    uniq = rng.integers(0, 4, (500, 16)).astype(np.float32)
    base = np.repeat(uniq, 2, axis=0)
    q = rng.integers(0, 4, (64, 16)).astype(np.float32)
    for mt in (Metric.Euclidean, Metric.DotProduct, Metric.Manhattan):
        idx = _index(base, mt, hora_amd.HM_SEARCH_BF16_RERANK)
        _exact_check(orc, idx, base, q, 5, mt)


def test_bf16_large_k_and_small_n(orc):
    base = bf_normal(300, 33, seed=431)
    q = bf_normal(40, 33, seed=432)
    for k in (1, 100, 299, 400):
        idx = _index(base, Metric.Euclidean, hora_amd.HM_SEARCH_BF16_RERANK)
        _exact_check(orc, idx, base, q, k, Metric.Euclidean)


@pytest.mark.parametrize("mt", [Metric.DotProduct, Metric.CosineSimilarity, Metric.Angular],
                         ids=lambda m: m.name)
def test_fp8_recall_and_exact_distances(orc, mt):
    base = unit_normal(20_000, 256, seed=441)
    q = unit_normal(200, 256, seed=442)
    idx = _index(base, mt, hora_amd.HM_SEARCH_FP8_RERANK, shortlist=512)
    gi, gd, gc = idx.search_batch(q, 10)
    ri, rd, rc = orc.bf_search(base, q, 10, int(mt))
    recall = np.mean([len(set(gi[r].tolist()) & set(ri[r].tolist())) / 10 for r in range(200)])
    assert recall >= 0.95, recall  # FP8 shortlist of 512, exact rerank
    # each returned distance is the exact reference value of that pair
    rows = orc.distance_rows(base, q, int(mt))
    for r in range(200):
        assert np.array_equal(gd[r], rows[r][gi[r].astype(np.int64)])


def test_nan_distances_fail_like_the_exact_form():
    # Cosine with a zero-norm row or query is NaN for that pair, which the
    # reference cannot order (neighbor.rs:32): the shortlist forms report it
    # exactly as the exact form does
    from hora_amd import HoraError
    base = bf_normal(6_000, 32, seed=451)
    q = bf_normal(4, 32, seed=452)
    for mode in (hora_amd.HM_SEARCH_BF16_RERANK, hora_amd.HM_SEARCH_FP8_RERANK):
        idx = _index(base, Metric.CosineSimilarity, mode)
        zq = q.copy()
        zq[2] = 0.0
        with pytest.raises(HoraError) as e:
            idx.search_batch(zq, 5)
        assert e.value.status == 8
        zb = base.copy()
        zb[1234] = 0.0
        idx = _index(zb, Metric.CosineSimilarity, mode)
        with pytest.raises(HoraError) as e:
            idx.search_batch(q, 5)
        assert e.value.status == 8


@pytest.mark.parametrize("mt", GEMM_METRICS, ids=lambda m: Metric(m).name)
@pytest.mark.parametrize("mode", [hora_amd.HM_SEARCH_BF16_RERANK, hora_amd.HM_SEARCH_FP8_RERANK],
                         ids=["bf16", "fp8"])
def test_filtered_scan_on_a_large_batch(orc, mt, mode):
    # 2,500 queries (a partial last query tile), 6,000 rows (a partial last
    # base tile), the filter form
    base = bf_sparse_u8(6_000, 784, seed=411)
    q = bf_sparse_u8(2_500, 784, seed=412)
    idx = _index(base, mt, mode)
    gi, gd, gc = idx.search_batch(q, 10)
    ri, rd, rc = orc.bf_search(base, q, 10, int(mt))
    assert np.array_equal(gc, rc)
    if mode == hora_amd.HM_SEARCH_BF16_RERANK:  # certified: bit-exact
        assert np.array_equal(gi, ri) and np.array_equal(gd, rd)
    else:
        hits = np.mean([len(set(gi[r].tolist()) & set(ri[r].tolist())) / 10
                        for r in range(q.shape[0])])
        assert hits >= 0.95, hits
        for r in range(0, q.shape[0], 97):  # every returned distance is the exact value
            for j in range(10):
                e = orc.metric(q[r], base[int(gi[r, j])], int(mt))
                assert np.float32(e).view(np.uint32) == gd[r, j].view(np.uint32)


@pytest.mark.parametrize("shortlist", [64, 512])
@pytest.mark.parametrize("spoil", [None, 0.5, -0.0, 256.0])
def test_u8_copy_of_integer_base_keeps_results_exact(orc, shortlist, spoil):
    # 0..255 integer rows: the rerank gathers a u8 copy of the base (slab path
    # at 64 candidates, direct loads at 512).  One value that a u8 cannot
    # hold bit for bit (0.5, -0.0, 256) switches the copy off for the whole
    # base.  Either way the ids and distances are the reference's.
    base = bf_sparse_u8(12_000, 784, seed=431)
    if spoil is not None:
        base[7, 100] = np.float32(spoil)
    q = bf_sparse_u8(200, 784, seed=432)
    idx = _index(base, Metric.Euclidean, hora_amd.HM_SEARCH_BF16_RERANK, shortlist=shortlist)
    _exact_check(orc, idx, base, q, 10, Metric.Euclidean)
    idx.set_search_mode(hora_amd.HM_SEARCH_FP8_RERANK, shortlist)
    gi, gd, gc = idx.search_batch(q, 10)
    ref = orc.distance_rows(base, q, orc.EUCLIDEAN)
    for r in range(q.shape[0]):  # every returned distance is the reference's value
        ids = gi[r, :gc[r]].astype(np.int64)
        assert np.array_equal(gd[r, :gc[r]].view(np.uint32), ref[r, ids].view(np.uint32))


@pytest.mark.parametrize("mt", [Metric.Euclidean, Metric.DotProduct, Metric.CosineSimilarity],
                         ids=lambda m: Metric(m).name)
@pytest.mark.parametrize("mode", [hora_amd.HM_SEARCH_BF16_RERANK, hora_amd.HM_SEARCH_FP8_RERANK],
                         ids=["bf16", "fp8"])
@pytest.mark.parametrize("nq", [6, 30, 40])
def test_tied_rows_rerun_from_the_u8_copy(orc, mt, mode, nq):
    # every integer row twice: each query's top-(k+1) holds equal distances,
    # so every row is rerun -- on the device for up to 32 rows (exact rows
    # from the u8 copy of the base, eight per group, then the heap replay),
    # on the exact path past that -- and must be the reference's
    half = bf_sparse_u8(3_000, 784, seed=471)
    base = np.concatenate([half, half])
    q = bf_sparse_u8(nq, 784, seed=472)
    idx = _index(base, mt, mode)
    rerun, _ = _exact_check(orc, idx, base, q, 10, mt)
    assert rerun == nq
