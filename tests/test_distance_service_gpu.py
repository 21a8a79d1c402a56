"""Batched distance service of the graph searches (hnsw_idx.rs:243-254,
ssg_idx.rs:402-456) vs the oracle: metric(node, query), bit-exact."""
import numpy as np
import pytest

from hora_amd import BruteForceIndex, HoraError, Metric
from hora_amd.datasets import bf_normal, bf_sparse_u8

pytestmark = pytest.mark.gpu

METRICS = [Metric.Euclidean, Metric.Manhattan, Metric.DotProduct, Metric.CosineSimilarity,
           Metric.Angular]


@pytest.mark.parametrize("mt", METRICS, ids=lambda m: m.name)
def test_ragged_lists_bit_exact(orc, mt):
    x = bf_normal(3000, 37, seed=95, sigma=1.0)
    q = bf_normal(50, 37, seed=96, sigma=1.0)
    rng = np.random.default_rng(97)
    lists = [rng.integers(0, 3000, int(rng.integers(0, 70))) for _ in range(50)]
    lists[3] = np.array([5, 5, 5])  # repeated node
    ix = BruteForceIndex(37)
    ix.add_batch(x)
    ix.build(mt)
    got = ix.distances(q, lists)
    full = orc.distance_rows(x, q, int(mt))  # metric(query, node): symmetric in f32
    for r, ids in enumerate(lists):
        assert np.array_equal(got[r], full[r, ids.astype(np.int64)]), f"query {r}"
    # the operand order of the graph searches: node first
    for r in (0, 7):
        for j, i in enumerate(lists[r][:5]):
            assert got[r][j] == np.float32(orc.metric(x[i], q[r], int(mt)))


def test_sparse_784_and_errors(orc):
    x = bf_sparse_u8(2000, 784, seed=98)
    q = bf_sparse_u8(10, 784, seed=99)
    ix = BruteForceIndex(784)
    ix.add_batch(x)
    with pytest.raises(HoraError) as e:  # before build: Metric::Unknown
        ix.distances(q[:1], [[1]])
    assert e.value.status == 4
    ix.build(Metric.Euclidean)
    lists = [np.arange(i, 2000, 97) for i in range(10)]
    got = ix.distances(q, lists)
    full = orc.distance_rows(x, q, orc.EUCLIDEAN)
    for r in range(10):
        assert np.array_equal(got[r], full[r, lists[r]])
    with pytest.raises(HoraError) as e:
        ix.distances(q[:1], [[2000]])
    assert e.value.status == 1 and "out of range" in e.value.message
    assert [len(v) for v in ix.distances(q[:2], [[], []])] == [0, 0]
