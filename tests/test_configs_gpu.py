"""The three large workloads at their full shapes, checked against the
oracle on query subsets (the oracle finishes those in seconds) and through
size-independent properties on the rest:

  PQIVFIndex 1,000,000 x 128 Gaussian mixture (1,024 centres, sigma 0.3),
      nlist 1,024, 16 x 256 PQ, k=100, nprobe 8 and 64;
  BruteForceIndex 1,000,000 x 768 unit-normalised rows, DotProduct and
      Cosine, ragged batches of 1-4,096 queries with k in {1, 10, 100},
      BF16 and FP8 shortlists of 4,096 with the exact rerank;
  PQIVFIndex 5,000,000 x 96 anisotropic mixture (4,096 centres, per-cluster
      sigma in 0.1-0.5), nlist 4,096, 12 x 256 PQ, the probe MLP
      96 -> 512 -> 4,096 picking nprobe per query."""
import time

import numpy as np
import pytest

import hora_amd
from hora_amd import BruteForceIndex, IVFPQIndex, IVFPQParams, Metric
from hora_amd.datasets import gaussian_mixture

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[configs {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _recall(got, truth):
    k = truth.shape[1]
    return float(np.mean([len(set(got[i].tolist()) & set(truth[i].tolist())) / k
                          for i in range(len(truth))]))


def _list_of_point(ex, n):
    off = ex["list_offsets"].astype(np.int64)
    lst = np.empty(n, np.int64)
    lst[ex["list_ids"].astype(np.int64)] = np.repeat(np.arange(len(off) - 1), np.diff(off))
    return lst


def _ivfpq_parity(orc, idx, ex, q, nprobe, k, given=None):
    gi, gd, gc = idx.search_batch(q, k)
    kw = {}
    np_ = nprobe
    if given is not None:
        probes, counts = given
        np_ = probes.shape[1]
        kw = {"given_probes": probes, "given_counts": counts}
    ri, rd, rc, _ = orc.ivfpq_search(q, orc.EUCLIDEAN, ex["centers"], ex["codebooks"],
                                     ex["list_offsets"], ex["list_ids"], ex["codes"], np_, k, **kw)
    assert np.array_equal(gc, rc)
    for r in range(q.shape[0]):
        c = int(rc[r])
        assert np.array_equal(gi[r, :c], ri[r, :c]), f"ids of query {r}"
        assert np.array_equal(gd[r, :c].view(np.uint32), rd[r, :c].view(np.uint32))
    return gi, gd, gc


def test_ivfpq_1m_x_128_nlist_1024(orc):
    N, D, NL, K = 1_000_000, 128, 1024, 100
    x, cen, _ = gaussian_mixture(N, D, NL, seed=303, sigma=0.3)
    q, _, _ = gaussian_mixture(2_000, D, NL, seed=304, sigma=0.3, centres=cen)
    p = (IVFPQParams().n_sub(16).sub_bits(8).n_kmeans_center(NL).search_n_center(8)
         .train_epoch(10).seed(7).kmeans_accumulate(1))
    idx = IVFPQIndex(D, p)
    idx.add_batch(x)
    t0 = time.perf_counter()
    idx.build(Metric.Euclidean)
    _log(f"IVFPQ 1M x 128 build (10 epochs): {time.perf_counter() - t0:.2f}s")
    ex = idx.export()
    # list assignment of every 50th point: the strict-'<' argmin against
    # the exported centroids (kmeans.rs:127-148, pq_idx.rs:366-371)
    samp = np.arange(0, N, 50)
    lst = _list_of_point(ex, N)
    assert np.array_equal(lst[samp], orc.kmeans_assign(x[samp], ex["centers"]).astype(np.int64))
    # codes of every 64th list from the exported codebooks (pq_idx.rs:143-144)
    off = ex["list_offsets"].astype(np.int64)
    for l in range(0, NL, 64):
        mem = ex["list_ids"][off[l]:off[l + 1]].astype(np.int64)
        if len(mem):
            ref = orc.pq_encode(x[mem], 16, ex["codebooks"][l], ex["centers"][l])
            assert np.array_equal(ref, ex["codes"][off[l]:off[l + 1]]), f"codes of list {l}"
    _log("assignments and codes match the oracle")
    # search: 200 queries against the oracle at nprobe 8 and 64, k=100
    for nprobe in (8, 64):
        idx.set_search_n_center(nprobe)
        gi, gd, gc = _ivfpq_parity(orc, idx, ex, q[:200], nprobe, K)
        # all 2,000 queries: the first 200 rows are the same, every row is
        # full and ascending
        ai, ad, ac = idx.search_batch(q, K)
        assert np.array_equal(ai[:200], gi) and np.array_equal(ad[:200], gd)
        assert (ac == K).all() and (np.diff(ad, axis=1) >= 0).all()
        _log(f"nprobe {nprobe}: 200 queries bit-identical to the oracle, 2,000 well formed")
    # the exact rerank of the ADC top-200 reaches recall@10 >= 0.95 at nprobe 8
    idx.set_search_n_center(8)
    idx.set_rerank(200)
    ri, rd, rc = idx.search_batch(q[:200], 10)
    truth, tdist, _ = orc.bf_search(x, q[:200], 10, orc.EUCLIDEAN)
    rec = _recall(ri, truth)
    _log(f"exact rerank 200, nprobe 8: recall@10 {rec:.4f}")
    assert rec >= 0.95
    hit = ri == truth  # where the ids agree, the distances are the exact ones
    assert np.array_equal(rd[hit].view(np.uint32), tdist[hit].view(np.uint32))


@pytest.fixture(scope="module")
def unit_base_768():
    import torch

    N, D = 1_000_000, 768
    g = torch.Generator(device="cuda")
    g.manual_seed(404)
    xb = torch.randn((N, D), generator=g, device="cuda", dtype=torch.float32)
    xb /= torch.linalg.norm(xb, dim=1, keepdim=True)
    xq = torch.randn((8_000, D), generator=g, device="cuda", dtype=torch.float32)
    xq /= torch.linalg.norm(xq, dim=1, keepdim=True)
    torch.cuda.synchronize()
    return xb, xb.cpu().numpy(), xq, xq.cpu().numpy()


# batches of 1-4,096 queries (two graph chunks), k in {1, 10, 100}
_SIZES = [1, 4096, 37, 2500, 900]
_KS = [10, 1, 100, 100, 10]


def _ragged_run(idx, xq, sizes, ks):
    import torch

    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    tot = int(sum(s * k for s, k in zip(sizes, ks)))
    di = torch.empty(tot, dtype=torch.int64, device="cuda")
    dd = torch.empty(tot, dtype=torch.float32, device="cuda")
    dc = torch.empty(int(off[-1]), dtype=torch.int32, device="cuda")
    idx.search_ragged_device(xq.data_ptr(), off, ks, di.data_ptr(), dd.data_ptr(),
                             dc.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return (di.cpu().numpy().view(np.uint64), dd.cpu().numpy(),
            dc.cpu().numpy().view(np.uint32), off)


def _checked_rows(sizes, ks, per_k=32):
    """(query, k, output offset) of up to per_k queries per k value."""
    rows, taken, pos = [], {}, 0
    for b, (s, k) in enumerate(zip(sizes, ks)):
        q0 = int(np.sum(sizes[:b]))
        step = max(1, s // per_k)
        for i in range(0, s, step):
            if taken.get(k, 0) < per_k:
                rows.append((q0 + i, k, pos + i * k))
                taken[k] = taken.get(k, 0) + 1
        pos += s * k
    return rows


@pytest.mark.parametrize("mt", [Metric.DotProduct, Metric.CosineSimilarity])
def test_bruteforce_1m_x_768_ragged_bf16_fp8(orc, unit_base_768, mt):
    xb, base, xq, qh = unit_base_768
    sizes, ks = _SIZES, _KS
    Q = int(np.sum(sizes))
    rows = _checked_rows(sizes, ks)
    truth = {}
    for k in sorted(set(ks)):
        sel = [r for r in rows if r[1] == k]
        ri, rd, rc = orc.bf_search(base, qh[[r[0] for r in sel]], k, int(mt))
        for j, r in enumerate(sel):
            truth[r[0]] = (ri[j], rd[j], rc[j])
    _log(f"{mt.name}: oracle over {len(rows)} of {Q} queries done")
    for mode, name in ((hora_amd.HM_SEARCH_BF16_RERANK, "bf16"),
                       (hora_amd.HM_SEARCH_FP8_RERANK, "fp8")):
        idx = BruteForceIndex(768)
        idx.add_device(xb.data_ptr(), xb.shape[0])
        idx.build(mt)
        idx.set_search_mode(mode, 4096)
        gi, gd, gc, off = _ragged_run(idx, xq, sizes, ks)
        assert (gc == np.repeat(ks, sizes)).all()
        hits = tot = 0
        for q, k, o in rows:
            ri, rd, rc = truth[q]
            if name == "bf16":  # certified: ids and distances bit-exact
                assert np.array_equal(gi[o:o + k], ri), (name, q, k)
                assert np.array_equal(gd[o:o + k].view(np.uint32), rd.view(np.uint32))
            else:  # measured: recall, and every returned distance is the exact one
                hits += len(set(gi[o:o + k].tolist()) & set(ri.tolist()))
                tot += k
                for j in range(k):
                    e = orc.metric(qh[q], base[int(gi[o + j])], int(mt))
                    assert np.float32(e).view(np.uint32) == gd[o + j].view(np.uint32)
        if name == "fp8":
            _log(f"{mt.name}/fp8 shortlist 4096: recall {hits / tot:.4f} on {len(rows)} queries")
            assert hits / tot >= 0.95
        else:
            _log(f"{mt.name}/bf16 shortlist 4096: {len(rows)} queries bit-exact")
        pos = 0  # every query's results ascending
        for s_, k in zip(sizes, ks):
            blk = gd[pos:pos + s_ * k].reshape(s_, k)
            assert (np.diff(blk, axis=1) >= 0).all()
            pos += s_ * k
        del idx


def test_learned_probe_5m_x_96_nlist_4096(orc):
    N, D, NL, K = 5_000_000, 96, 4096, 10
    t0 = time.perf_counter()
    x, cen, _ = gaussian_mixture(N, D, NL, seed=505, aniso=(0.1, 0.5))
    tr, _, _ = gaussian_mixture(50_000, D, NL, seed=506, aniso=(0.1, 0.5), centres=cen)
    te, _, _ = gaussian_mixture(1_000, D, NL, seed=507, aniso=(0.1, 0.5), centres=cen)
    _log(f"5M x 96 mixture generated in {time.perf_counter() - t0:.1f}s")
    p = (IVFPQParams().n_sub(12).sub_bits(8).n_kmeans_center(NL).search_n_center(8)
         .train_epoch(6).seed(9).kmeans_accumulate(1))
    idx = IVFPQIndex(D, p)
    idx.add_batch(x)
    t0 = time.perf_counter()
    idx.build(Metric.Euclidean)
    _log(f"IVFPQ 5M x 96 build (6 epochs): {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    loss = idx.train_probe(tr, label_k=K, hidden=512, epochs=2, batch=1024, lr=1e-3, seed=1)
    _log(f"probe MLP 96-512-4096 trained on 50,000 queries in {time.perf_counter() - t0:.1f}s, "
         f"loss {loss:.3f}")
    assert np.isfinite(loss)
    ex = idx.export()
    sub = te[:200]
    # fixed nprobe: the reference's search
    _ivfpq_parity(orc, idx, ex, sub, 8, K)
    # learned nprobe: the reference's per-list scan and merge over the
    # network's lists, in the network's order
    idx.set_probe_policy(1, mass=0.9, pmin=1, pmax=32)
    gi, gd, gc = idx.search_batch(sub, K)
    probes, counts = idx.last_probes(len(sub), 32)
    assert 1 <= counts.min() and counts.max() <= 32
    _ivfpq_parity(orc, idx, ex, sub, None, K, given=(probes, counts))
    _log(f"learned form: 200 queries bit-identical to the oracle over the chosen lists "
         f"(mean nprobe {idx.mean_nprobe():.2f})")
