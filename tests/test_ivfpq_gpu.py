"""IVFPQIndex on the GPU vs the oracle.

f32 accumulation (the reference's point-order sums): training is bit-exact
end to end given the seed (coarse centroids, inverted lists, per-list
codebooks and codes), and search results are bit-exact.
fp64 accumulation: parity is conditional on the trained state -- the oracle
re-derives the list assignments, the codes and the search results from the
exported centroids and codebooks, and they must be identical."""
import numpy as np
import pytest

from hora_amd import HoraError, IVFPQIndex, IVFPQParams, Metric
from hora_amd.datasets import gaussian_mixture
from ivf_rerank_oracle import rerank_oracle as _rerank_oracle

pytestmark = pytest.mark.gpu


def _build(x, mt, nlist, n_sub, sub_bits, epochs, seed, fp64=False, nprobe=4):
    p = (IVFPQParams().n_sub(n_sub).sub_bits(sub_bits).n_kmeans_center(nlist)
         .search_n_center(nprobe).train_epoch(epochs).seed(seed)
         .kmeans_accumulate(1 if fp64 else 0))
    idx = IVFPQIndex(x.shape[1], p)
    idx.add_batch(x)
    idx.build(mt)
    return idx


def _lists(ex):
    off = ex["list_offsets"].astype(np.int64)
    return [ex["list_ids"][off[l]:off[l + 1]].astype(np.int64) for l in range(len(off) - 1)]


def _same_valid(gi, gd, ri, rd, counts):
    # entries past a row's count are unspecified fill in the oracle
    for r, c in enumerate(counts):
        assert np.array_equal(gi[r, :c], ri[r, :c]), f"ids of row {r}"
        assert np.array_equal(gd[r, :c], rd[r, :c]), f"distances of row {r}"


def _check_conditional(orc, idx, x, mt, q, nprobe, k):
    ex = idx.export()
    cen, cb, codes = ex["centers"], ex["codebooks"], ex["codes"]
    n_sub = cb.shape[1]
    # list assignments from the exported centroids (kmeans.rs:127-148)
    assign = orc.kmeans_assign(x, cen, mt=int(mt))
    lists = _lists(ex)
    for l, mem in enumerate(lists):
        assert np.array_equal(np.sort(mem), mem)
        assert (assign[mem] == l).all()
    assert sum(len(m) for m in lists) == x.shape[0]
    # codes from the exported codebooks (pq_idx.rs:143-144)
    off = ex["list_offsets"].astype(np.int64)
    for l, mem in enumerate(lists):
        if len(mem) == 0:
            continue
        ref = orc.pq_encode(x[mem], n_sub, cb[l], cen[l])
        assert np.array_equal(ref, codes[off[l]:off[l + 1]]), f"codes of list {l}"
    # search
    idx.set_search_n_center(nprobe)
    gi, gd, gc = idx.search_batch(q, k)
    ri, rd, rc, _ = orc.ivfpq_search(q, int(mt), cen, cb, ex["list_offsets"], ex["list_ids"],
                                     codes, nprobe, k)
    assert np.array_equal(gc, rc)
    _same_valid(gi, gd, ri, rd, rc)
    return ex


def test_f32_training_is_bit_exact_end_to_end(orc):
    x, _, _ = gaussian_mixture(3000, 16, 12, seed=5, sigma=0.3)
    nlist, n_sub, ks, epochs, seed = 16, 4, 16, 6, 77
    idx = _build(x, Metric.Euclidean, nlist, n_sub, 4, epochs, seed)
    ex = idx.export()
    cen_ref, assign_ref = orc.kmeans_train(x, nlist, epochs, seed)
    assert np.array_equal(ex["centers"], cen_ref)
    lists = _lists(ex)
    for l, mem in enumerate(lists):
        assert np.array_equal(mem, np.nonzero(assign_ref == l)[0])
    off = ex["list_offsets"].astype(np.int64)
    for l, mem in enumerate(lists):
        cb_ref, codes_ref = orc.pq_train(x[mem], n_sub, ks, cen_ref[l], epochs, seed,
                                         1 + l * n_sub, ex["codebooks"].shape[3])
        assert np.array_equal(ex["codebooks"][l], cb_ref), f"codebook {l}"
        assert np.array_equal(ex["codes"][off[l]:off[l + 1]], codes_ref), f"codes {l}"
    q, _, _ = gaussian_mixture(100, 16, 12, seed=6, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 4, 10)


def test_fp64_training_conditional_parity(orc):
    x, _, _ = gaussian_mixture(5000, 32, 24, seed=8, sigma=0.3)
    idx = _build(x, Metric.Euclidean, 32, 8, 8, 5, 9, fp64=True)
    q, _, _ = gaussian_mixture(128, 32, 24, seed=10, sigma=0.3)
    for nprobe, k in ((1, 10), (8, 10), (32, 50)):
        _check_conditional(orc, idx, x, Metric.Euclidean, q, nprobe, k)


def test_reference_default_params_uneven_subspaces(orc):
    # IVFPQParams::default: n_sub=25, sub_bits=4 -> d=128 gives 3 subs of 6
    # dims and 22 of 5 (pq_idx.rs:75-88)
    x, _, _ = gaussian_mixture(2000, 128, 8, seed=11, sigma=0.3)
    p = IVFPQParams().n_kmeans_center(8).train_epoch(3).seed(3)
    idx = IVFPQIndex(128, p)
    idx.add_batch(x)
    idx.build(Metric.Euclidean)
    info = idx.info()
    assert (info["n_sub"], info["n_sub_center"], info["sub_dim_max"]) == (25, 16, 6)
    q, _, _ = gaussian_mixture(50, 128, 8, seed=12, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 8, 10)


@pytest.mark.parametrize("mt", [Metric.DotProduct, Metric.Manhattan, Metric.CosineSimilarity],
                         ids=lambda m: m.name)
def test_coarse_metric_variants(orc, mt):
    x, _, _ = gaussian_mixture(3000, 24, 10, seed=13, sigma=0.3)
    idx = _build(x, mt, 12, 6, 5, 4, 21)
    q, _, _ = gaussian_mixture(64, 24, 10, seed=14, sigma=0.3)
    _check_conditional(orc, idx, x, mt, q, 3, 10)


def test_large_codebook_scan(orc):
    # 8 x 256 codewords of 32 dims (256 KiB per list) do not fit in LDS: the
    # per-(query, probe) scan runs instead of the list-major one
    x, _, _ = gaussian_mixture(3000, 256, 8, seed=17, sigma=0.3)
    idx = _build(x, Metric.Euclidean, 8, 8, 8, 3, 23)
    assert idx.info()["sub_dim_max"] == 32
    q, _, _ = gaussian_mixture(40, 256, 8, seed=18, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 3, 10)


@pytest.mark.parametrize("form", ["lists", "rows"])
@pytest.mark.parametrize("n", [12_000, 40_000], ids=["registers", "chunked"])
def test_long_lists_select(orc, n, form):
    # lists form: 16 x 256 codewords of 8 dims leave room for a 1,024-key
    # sorting pool, so lists of ~3,000 take the histogram select over
    # register-held keys and lists of ~10,000 the chunked sort; rows form:
    # rows of ~6,000 and ~20,000 distances through the radix row select
    x, _, _ = gaussian_mixture(n, 128, 4, seed=19, sigma=0.3)
    idx = _build(x, Metric.Euclidean, 4, 16, 8, 2, 29)
    idx.set_scan_form(form)
    q, _, _ = gaussian_mixture(32, 128, 4, seed=20, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 2, 10)


@pytest.mark.parametrize("form", ["lists", "rows"])
def test_long_lists_all_tied(orc, form):
    # 3 distinct points repeated 1,500 times in one list: its ADC keys take
    # 3 values (lists form: the threshold bin holds 1,500 > 1,024 keys, so
    # the select falls back to the chunked sort), and every query goes
    # through the heap replay
    base, _, _ = gaussian_mixture(3, 128, 3, seed=21, sigma=0.3)
    x = np.repeat(base, 1500, axis=0)
    idx = _build(x, Metric.Euclidean, 1, 16, 8, 2, 31, nprobe=1)
    idx.set_scan_form(form)
    q, _, _ = gaussian_mixture(8, 128, 3, seed=22, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 1, 10)
    assert idx.search_stats()[0] == 8


@pytest.mark.parametrize("form", ["per_probe", "lists", "rows"])
def test_ties_replay(orc, form):
    # every row twice: equal codes give equal ADC distances in every list
    x, _, _ = gaussian_mixture(800, 12, 6, seed=15, sigma=0.3)
    x = np.repeat(x, 2, axis=0)
    idx = _build(x, Metric.Euclidean, 8, 3, 4, 4, 5)
    idx.set_scan_form(form)
    q, _, _ = gaussian_mixture(40, 12, 6, seed=16, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 3, 7)
    assert idx.search_stats()[0] > 0  # rows went through the replay


@pytest.mark.parametrize("form", ["lists", "rows"])
def test_deep_list_heaps_with_ties(orc, form):
    # every row twice in lists of ~3,000 items and k=100: the per-list heap
    # replays run through levels 7-11 (subtree pushes and their events,
    # pop tasks two to a lane) for every tied query
    x, _, _ = gaussian_mixture(3_000, 32, 2, seed=61, sigma=0.3)
    x = np.repeat(x, 2, axis=0)
    idx = _build(x, Metric.Euclidean, 2, 4, 8, 3, 63, nprobe=2)
    idx.set_scan_form(form)
    q, _, _ = gaussian_mixture(24, 32, 2, seed=62, sigma=0.3)
    _check_conditional(orc, idx, x, Metric.Euclidean, q, 2, 100)
    assert idx.search_stats()[0] > 0


@pytest.mark.parametrize("shape", [(64, 16, 8, 24), (32, 4, 4, 40), (48, 6, 5, 12),
                                   (96, 3, 8, 6)],
                         ids=["16x256x4", "4x16x8", "6x32x8", "3x256x32"])
def test_scan_forms_agree(orc, shape):
    # every scan form that fits gives the same results, tie flags included;
    # k larger than some probed lists' total, and a k of 1
    d, n_sub, bits, nlist = shape
    x, cen, _ = gaussian_mixture(4000, d, nlist, seed=41, sigma=0.3)
    idx = _build(x, Metric.Euclidean, nlist, n_sub, bits, 3, 43, nprobe=3)
    q, _, _ = gaussian_mixture(70, d, nlist, seed=42, sigma=0.3, centres=cen)
    ex = _check_conditional(orc, idx, x, Metric.Euclidean, q, 3, 10)
    for k, nprobe in ((10, 3), (1, 1), (400, 2)):
        idx.set_search_n_center(nprobe)
        got = {}
        for form in ("per_probe", "lists", "rows"):
            idx.set_scan_form(form)
            try:
                got[form] = idx.search_batch(q, k) + (idx.search_stats()[0],)
            except HoraError:
                assert form != "per_probe"  # the codebook does not fit that form
        assert "rows" in got or n_sub * (1 << bits) > 2048
        ref = got["per_probe"]
        for form, g in got.items():
            for a, b in zip(ref, g):
                assert np.array_equal(a, b), (form, k, nprobe)
        ri, rd, rc, _ = orc.ivfpq_search(q, orc.EUCLIDEAN, ex["centers"], ex["codebooks"],
                                         ex["list_offsets"], ex["list_ids"], ex["codes"],
                                         nprobe, k)
        assert np.array_equal(ref[2], rc)
        _same_valid(ref[0], ref[1], ri, rd, rc)
        assert (ref[2] > 0).mean() > 0.5  # most queries found neighbours
    idx.set_scan_form("auto")


def test_errors():
    x, _, _ = gaussian_mixture(300, 8, 4, seed=1)
    p = IVFPQParams().n_sub(2).sub_bits(4).n_kmeans_center(4).search_n_center(2).train_epoch(2)
    idx = IVFPQIndex(8, p)
    idx.add_batch(x)
    with pytest.raises(HoraError) as e:
        idx.search_batch(x[:2], 3)
    assert e.value.status == 5
    idx.build(Metric.Euclidean)
    idx.set_search_n_center(5)  # more than n_kmeans_center
    with pytest.raises(HoraError):
        idx.search_batch(x[:2], 3)
    empty = IVFPQIndex(8, p)
    with pytest.raises(HoraError) as e:
        empty.build(Metric.Euclidean)
    assert e.value.status == 9
    with pytest.raises(HoraError) as e:
        IVFPQIndex(8, IVFPQParams().n_sub(2).sub_bits(9))
    assert e.value.status == 10


@pytest.mark.parametrize("mt,tied", [(Metric.Euclidean, False), (Metric.Euclidean, True),
                                     (Metric.CosineSimilarity, False), (Metric.Manhattan, True),
                                     (Metric.DotProduct, False)])
def test_exact_rerank_matches_oracle(orc, mt, tied):
    x, cen, _ = gaussian_mixture(6000, 24, 32, seed=61, sigma=0.4)
    if tied:  # duplicated rows: equal ADC and equal exact distances
        x = np.round(np.repeat(x[:3000], 2, axis=0) * 2.0).astype(np.float32)
    idx = _build(x, mt, 32, 6, 8, 4, 62, fp64=True, nprobe=4)
    q, _, _ = gaussian_mixture(60, 24, 32, seed=63, sigma=0.4, centres=cen)
    if tied:
        q = np.round(q * 2.0).astype(np.float32)
    ex = idx.export()
    for depth, k, nprobe in ((50, 10, 4), (10, 10, 2), (200, 25, 6), (5, 1, 1)):
        idx.set_search_n_center(nprobe)
        idx.set_rerank(depth)
        gi, gd, gc = idx.search_batch(q, k)
        ri, rd, rc = _rerank_oracle(orc, x, q, mt, ex, nprobe, k, depth)
        assert np.array_equal(gc, rc), (depth, k, nprobe)
        for r in range(q.shape[0]):
            assert np.array_equal(gi[r, :gc[r]], ri[r]), (depth, k, nprobe, r)
            assert np.array_equal(gd[r, :gc[r]].view(np.uint32), rd[r].view(np.uint32))
    idx.set_rerank(0)  # back to the reference's ADC result
    ai, ad, ac = idx.search_batch(q, 10)
    ri, rd, rc, _ = orc.ivfpq_search(q, int(mt), ex["centers"], ex["codebooks"],
                                     ex["list_offsets"], ex["list_ids"], ex["codes"], 1, 10)
    assert np.array_equal(ac, rc)
    _same_valid(ai, ad, ri, rd, rc)


@pytest.mark.parametrize("form", ["per_probe", "lists", "rows"])
def test_ties_among_empty_lists_need_no_replay(orc, form):
    # 4 tight clusters far from the origin, 32 lists and one epoch: the
    # init has two distinct centres, so 30 lists end empty with an all-zero
    # centroid (kmeans.rs:115-122) and the queries see runs of equal
    # coarse distances; those runs only hold empty lists, which cannot
    # change the result (the tie check skips them)
    x, cen, _ = gaussian_mixture(3000, 16, 4, seed=71, sigma=0.05)
    x = (x + 3.0).astype(np.float32)
    idx = _build(x, Metric.Euclidean, 32, 4, 4, 1, 72, fp64=True)
    idx.set_scan_form(form)
    ex = idx.export()
    sizes = np.diff(ex["list_offsets"].astype(np.int64))
    zero = int((np.abs(ex["centers"]).sum(1) == 0).sum())
    assert (sizes == 0).sum() >= 20 and zero >= 6
    q, _, _ = gaussian_mixture(50, 16, 4, seed=73, sigma=0.05, centres=cen)
    q = (q + 3.0).astype(np.float32)
    for nprobe in (2, 5, 16):
        _check_conditional(orc, idx, x, Metric.Euclidean, q, nprobe, 10)
    cd = np.sort(orc.distance_rows(ex["centers"], q, orc.EUCLIDEAN), axis=1)
    assert (cd[:, 15] == cd[:, 16]).all()  # the 16th and 17th are zero centroids
    idx.set_rerank(30)
    idx.set_search_n_center(16)
    gi, gd, gc = idx.search_batch(q, 10)
    ri, rd, rc = _rerank_oracle(orc, x, q, Metric.Euclidean, ex, 16, 10, 30)
    assert np.array_equal(gc, rc)
    for r in range(q.shape[0]):
        assert np.array_equal(gi[r, :gc[r]], ri[r])
