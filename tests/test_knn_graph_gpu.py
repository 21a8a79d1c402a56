"""All-pairs kNN graph (naive_build_knn_graph, knn.rs:14-41) on the GPU vs
the oracle: ids and distances bit-exact, rows largest distance first, the
node itself skipped (not pushed), ties resolved as the std heap does."""
import numpy as np
import pytest

from hora_amd import BruteForceIndex, HoraError, Metric, naive_build_knn_graph
from hora_amd.datasets import bf_normal, bf_sparse_u8

pytestmark = pytest.mark.gpu


def _check(orc, x, k, mt):
    ix = BruteForceIndex(x.shape[1])
    ix.add_batch(x)
    ix.build(mt)
    gi, gd, gc = ix.build_knn_graph(k)
    ri, rd, rc = orc.knn_graph(x, k, int(mt))
    assert np.array_equal(gc, rc)
    for r in range(x.shape[0]):
        m = int(rc[r])
        assert gi[r, :m].tolist() == ri[r, :m].tolist(), f"row {r} ids"
        assert np.array_equal(gd[r, :m], rd[r, :m]), f"row {r} distances"
        assert (gi[r, m:] == np.iinfo(np.uint64).max).all()
    return ix


@pytest.mark.parametrize("mt", [Metric.Euclidean, Metric.Manhattan, Metric.DotProduct,
                                Metric.CosineSimilarity, Metric.Angular],
                         ids=lambda m: m.name)
def test_graph_metrics(orc, mt):
    x = bf_normal(1500, 24, seed=91, sigma=1.0)
    _check(orc, x, 10, mt)


@pytest.mark.parametrize("k", [1, 10, 70])
def test_graph_ties_replay(orc, k):
    # integer points in a small box: most rows have equal distances among
    # their nearest, and go through the heap replay with the self column
    # removed
    rng = np.random.default_rng(92)
    x = rng.integers(0, 4, (900, 3)).astype(np.float32)
    ix = _check(orc, x, k, Metric.Manhattan)
    assert ix.search_stats()[0] > 0


def test_graph_sparse_784(orc):
    x = bf_sparse_u8(2500, 784, seed=93)
    _check(orc, x, 10, Metric.Euclidean)


def test_graph_edges(orc):
    x = bf_normal(5, 4, seed=94, sigma=1.0)
    _check(orc, x, 10, Metric.Euclidean)  # k > n - 1
    _check(orc, x[:1], 3, Metric.Euclidean)  # a single node: no neighbours
    _check(orc, x[:2], 3, Metric.Euclidean)
    ix = BruteForceIndex(4)
    ix.add_batch(x)
    with pytest.raises(HoraError) as e:  # Metric::Unknown before build
        ix.build_knn_graph(3)
    assert e.value.status == 4
    ix.build(Metric.Euclidean)
    gi, gd, gc = ix.build_knn_graph(0)
    assert gc.tolist() == [0] * 5
    g = naive_build_knn_graph(x, Metric.Euclidean, 2)
    ri, rd, _ = orc.knn_graph(x, 2, orc.EUCLIDEAN)
    assert [[i for i, _ in row] for row in g] == ri.tolist()
    z = np.zeros((4, 4), np.float32)
    z[1:] = x[1:4]
    cos = BruteForceIndex(4)
    cos.add_batch(z)
    cos.build(Metric.CosineSimilarity)
    with pytest.raises(HoraError) as e:  # a zero row: NaN distances
        cos.build_knn_graph(2)
    assert e.value.status == 8
