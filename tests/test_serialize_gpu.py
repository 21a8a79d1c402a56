"""SerializableIndex dump / load (ann_index.rs:141-158) on the GPU indexes.

The dumped bytes must equal the oracle's restatement of the reference's
bincode layout (oracle/bincode.py, itself pinned by hand-assembled bytes in
test_oracle.py) built from the index's own exported state; loading gives an
index whose state and search results equal the original's; files written
by the restatement load and search like the oracle."""
import numpy as np
import pytest

from hora_amd import (BruteForceIndex, HoraError, IVFPQIndex, IVFPQParams, Metric, PQIndex,
                      PQParams)
from hora_amd.datasets import bf_normal, gaussian_mixture
from oracle import bincode as bc

pytestmark = pytest.mark.gpu


def _ids(n):
    return [bc.option_usize(i) for i in range(n)]


def _same_search(a, b, q, k):
    x, y = a.search_batch(q, k), b.search_batch(q, k)
    for u, v in zip(x, y):
        assert np.array_equal(u, v)


def test_bruteforce_bytes_and_round_trip(tmp_path):
    x = bf_normal(300, 12, seed=81)
    idx = BruteForceIndex(12)
    idx.add_batch(x)
    idx.build(Metric.Manhattan)
    path = tmp_path / "bf.idx"
    idx.dump(path)
    assert path.read_bytes() == bc.bruteforce(x, _ids(300), int(Metric.Manhattan), 12)
    back = BruteForceIndex.load(path)
    assert back.nodes_size() == 300 and back.dimension() == 12
    assert back._ids == list(range(300))
    _same_search(idx, back, bf_normal(20, 12, seed=82), 7)  # mt restored: no build needed


def test_bruteforce_string_ids_and_unbuilt(tmp_path):
    x = bf_normal(40, 5, seed=83)
    idx = BruteForceIndex(5)
    names = [f"item-{i}" for i in range(40)]
    idx.add_batch(x, names)
    path = tmp_path / "bf_s.idx"
    idx.dump(path, idx_type="String")
    enc = [b"\x01" + len(s).to_bytes(8, "little") + s.encode() for s in names]
    assert path.read_bytes() == bc.bruteforce(x, enc, int(Metric.Unknown), 5)
    back = BruteForceIndex.load(path, idx_type="String")
    assert back._ids == names
    with pytest.raises(HoraError) as e:  # mt Unknown: search fails as before the dump
        back.search_batch(x[:1], 3)
    assert e.value.status == 4
    back.build(Metric.Euclidean)
    assert back.search(x[3], 1) == ["item-3"]


def test_pq_bytes_and_round_trip(orc, tmp_path):
    x = bf_normal(1500, 20, seed=84, sigma=1.0)
    idx = PQIndex(20, PQParams().n_sub(3).sub_bits(4).train_epoch(4).seed(5))
    idx.add_batch(x)
    path0 = tmp_path / "pq0.idx"
    idx.dump(path0)  # before build: untrained, mt Euclidean (pq_idx.rs:71)
    assert path0.read_bytes() == bc.pqindex(20, 3, 4, 4, None, None, x, _ids(1500), 3)
    idx.build(Metric.DotProduct)
    ex = idx.export()
    path = tmp_path / "pq.idx"
    idx.dump(path)
    assert path.read_bytes() == bc.pqindex(20, 3, 4, 4, ex["codebook"], ex["codes"], x,
                                           _ids(1500), int(Metric.DotProduct))
    back = PQIndex.load(path)
    ex2 = back.export()
    assert np.array_equal(ex["codebook"], ex2["codebook"])
    assert np.array_equal(ex["codes"], ex2["codes"])
    _same_search(idx, back, bf_normal(30, 20, seed=85, sigma=1.0), 9)
    un = PQIndex.load(path0)
    assert un.info()["trained"] == 0


def test_pq_file_from_the_restatement_searches_like_the_oracle(orc, tmp_path):
    x = bf_normal(800, 16, seed=86, sigma=1.0)
    cb, codes = orc.pqindex_train(x, 4, 16, orc.EUCLIDEAN, 3, 9)
    path = tmp_path / "pq_orc.idx"
    path.write_bytes(bc.pqindex(16, 4, 4, 3, cb, codes, x, _ids(800), orc.EUCLIDEAN))
    idx = PQIndex.load(path)
    q = bf_normal(25, 16, seed=87, sigma=1.0)
    gi, gd, gc = idx.search_batch(q, 6)
    ri, rd, rc = orc.pqindex_search(q, orc.EUCLIDEAN, cb, codes, 6)
    assert np.array_equal(gi, ri) and np.array_equal(gd, rd) and np.array_equal(gc, rc)


def test_ivfpq_bytes_and_round_trip(orc, tmp_path):
    x, _, _ = gaussian_mixture(2500, 16, 8, seed=88)
    p = IVFPQParams().n_sub(4).sub_bits(4).n_kmeans_center(8).search_n_center(3) \
        .train_epoch(3).seed(2)
    idx = IVFPQIndex(16, p)
    idx.add_batch(x)
    path0 = tmp_path / "ivf0.idx"
    idx.dump(path0)
    assert path0.read_bytes() == bc.ivfpq(16, 4, 4, 3, 3, 8, None, None, None, None, None, x,
                                          _ids(2500), int(Metric.Unknown))
    idx.build(Metric.Euclidean)
    ex = idx.export()
    path = tmp_path / "ivf.idx"
    idx.dump(path)
    assert path.read_bytes() == bc.ivfpq(16, 4, 4, 3, 3, 8, ex["centers"], ex["list_offsets"],
                                         ex["list_ids"], ex["codebooks"], ex["codes"], x,
                                         _ids(2500), int(Metric.Euclidean))
    back = IVFPQIndex.load(path)
    ex2 = back.export()
    for key in ("centers", "list_offsets", "list_ids", "codebooks", "codes"):
        assert np.array_equal(ex[key], ex2[key]), key
    assert back.info()["nprobe"] == 3
    _same_search(idx, back, gaussian_mixture(40, 16, 8, seed=89)[0], 10)


def test_malformed_files_are_statuses(tmp_path):
    x = bf_normal(10, 4, seed=90)
    idx = BruteForceIndex(4)
    idx.add_batch(x)
    idx.build(Metric.Euclidean)
    path = tmp_path / "bf.idx"
    idx.dump(path)
    raw = path.read_bytes()
    bad = tmp_path / "bad.idx"
    bad.write_bytes(raw[:-3])  # truncated
    with pytest.raises(HoraError) as e:
        BruteForceIndex.load(bad)
    assert e.value.status == 1 and "end of file" in e.value.message
    bad.write_bytes(raw + b"\x00")  # trailing bytes
    with pytest.raises(HoraError):
        BruteForceIndex.load(bad)
    mt_at = len(raw) - 12
    bad.write_bytes(raw[:mt_at] + (9).to_bytes(4, "little") + raw[mt_at + 4:])
    with pytest.raises(HoraError) as e:  # no Metric variant 9
        BruteForceIndex.load(bad)
    assert "Metric" in e.value.message
    with pytest.raises(HoraError):
        BruteForceIndex.load(tmp_path / "missing.idx")
