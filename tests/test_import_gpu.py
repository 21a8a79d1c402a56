"""Trained state in through the ABI (hm_ivfpq_import, hm_pq_import), and the
trait defaults rebuild / idx_info (ann_index.rs:66-72, 113-115).

An index that imports another's exported state searches exactly like it:
ids, distances and counts bit-identical, and its own export returns the
same arrays.  Malformed state is refused with a status."""
import numpy as np
import pytest

from hora_amd import BruteForceIndex, HoraError, IVFPQIndex, IVFPQParams, Metric, PQIndex, PQParams
from hora_amd.datasets import gaussian_mixture

pytestmark = pytest.mark.gpu


def _ivf(x, fp64=False):
    p = (IVFPQParams().n_sub(8).sub_bits(8).n_kmeans_center(16).search_n_center(4)
         .train_epoch(4).seed(5).kmeans_accumulate(1 if fp64 else 0))
    ix = IVFPQIndex(x.shape[1], p)
    ix.add_batch(x)
    return ix


def _same_search(a, b, q, k):
    ai, ad, ac = a.search_batch(q, k)
    bi, bd, bc = b.search_batch(q, k)
    assert np.array_equal(ac, bc)
    assert np.array_equal(ai, bi)
    assert np.array_equal(ad.view(np.uint32), bd.view(np.uint32))


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("mt", [Metric.Euclidean, Metric.DotProduct])
def test_ivfpq_import_searches_like_the_exporter(fp64, mt):
    x, _, _ = gaussian_mixture(6000, 32, 24, seed=41, sigma=0.3)
    q, _, _ = gaussian_mixture(200, 32, 24, seed=42, sigma=0.3)
    a = _ivf(x, fp64)
    a.build(mt)
    st = a.export()
    b = _ivf(x, fp64)
    assert b.info()["trained"] == 0  # built() is always true (pq_idx.rs:454-456)
    b.import_trained(st, mt)
    assert b.info()["trained"] == 1
    st2 = b.export()
    for key in st:
        assert np.array_equal(st[key], st2[key]), key
    for nprobe, k in ((1, 10), (4, 10), (16, 50)):
        a.set_search_n_center(nprobe)
        b.set_search_n_center(nprobe)
        _same_search(a, b, q, k)


def test_ivfpq_import_refuses_bad_state():
    x, _, _ = gaussian_mixture(3000, 32, 8, seed=43, sigma=0.3)
    a = _ivf(x)
    a.build(Metric.Euclidean)
    st = a.export()
    b = _ivf(x[:2000])  # fewer rows than the lists hold
    with pytest.raises(HoraError, match="cover the rows"):
        b.import_trained(st, Metric.Euclidean)
    c = _ivf(x)
    bad = dict(st)
    bad["list_ids"] = st["list_ids"].copy()
    bad["list_ids"][0] = bad["list_ids"][1]  # not a permutation
    with pytest.raises(HoraError, match="permutation"):
        c.import_trained(bad, Metric.Euclidean)
    bad = dict(st)
    bad["codebooks"] = st["codebooks"][:-1]
    with pytest.raises(HoraError, match="does not match"):
        c.import_trained(bad, Metric.Euclidean)
    assert c.info()["trained"] == 0


@pytest.mark.parametrize("mt", [Metric.Euclidean, Metric.CosineSimilarity])
def test_pq_import_searches_like_the_exporter(mt):
    x, _, _ = gaussian_mixture(5000, 32, 16, seed=44, sigma=1.0)
    q, _, _ = gaussian_mixture(150, 32, 16, seed=45, sigma=1.0)
    p = PQParams().n_sub(8).sub_bits(6).train_epoch(5).seed(3)
    a = PQIndex(32, p)
    a.add_batch(x)
    a.build(mt)
    st = a.export()
    b = PQIndex(32, p)
    b.add_batch(x)
    b.import_trained(st, mt)
    assert b.info()["trained"] == 1
    st2 = b.export()
    assert np.array_equal(st["codebook"], st2["codebook"])
    assert np.array_equal(st["codes"], st2["codes"])
    for k in (1, 10, 64):
        _same_search(a, b, q, k)
    bad = {"codebook": st["codebook"], "codes": np.full_like(st["codes"], 255)}
    c = PQIndex(32, p)
    c.add_batch(x)
    with pytest.raises(HoraError, match="code out of range"):
        c.import_trained(bad, mt)


def test_rebuild_and_idx_info_are_the_trait_defaults():
    x, _, _ = gaussian_mixture(500, 16, 4, seed=46, sigma=0.3)
    bf = BruteForceIndex(16)
    pq = PQIndex(16, PQParams().n_sub(4).sub_bits(4).train_epoch(2))
    ivf = IVFPQIndex(16, IVFPQParams().n_sub(4).sub_bits(4).n_kmeans_center(4).train_epoch(2))
    for ix in (bf, pq, ivf):
        ix.add_batch(x)
        ix.build(Metric.Euclidean)
        with pytest.raises(HoraError, match="not implement"):
            ix.rebuild(Metric.Euclidean)
        assert ix.built()  # the reference's built() is always true
        assert ix.idx_info() == "not implement"
