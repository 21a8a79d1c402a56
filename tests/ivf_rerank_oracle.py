"""The exact-rerank form of IVFPQ search restated from its definition, on an
exported index (test helper; the checks it relies on -- the LUT, the ADC
fold and the metric -- are pinned in test_oracle.py)."""
import numpy as np


def _adc_fold(lut, codes):
    # sequential f32 sum over the sub-spaces from +0.0 (pq_idx.rs:183-186)
    d = np.zeros(codes.shape[0], np.float32)
    for j in range(codes.shape[1]):
        d = (d + lut[j][codes[:, j]]).astype(np.float32)
    return d


def rerank_oracle(orc, x, q, mt, ex, nprobe, k, depth, given=None):
    """The exact-rerank form, restated from its definition: the nprobe
    lists nearest by (coarse distance with the build metric, list index) --
    or the given lists --, the R = max(depth, k) smallest ADC distances of
    their items by (distance, row index) (the reference's per-list LUT and
    fold, pq_idx.rs:170-188), each candidate's exact distance by the oracle
    metric, sorted by (exact distance, row index), the first k."""
    R = max(depth, k)
    cen, cb, codes = ex["centers"], ex["codebooks"], ex["codes"]
    off = ex["list_offsets"].astype(np.int64)
    lid = ex["list_ids"].astype(np.int64)
    n_sub = cb.shape[1]
    cd = orc.distance_rows(cen, q, int(mt))
    ids, dist, cnt = [], [], []
    for r in range(q.shape[0]):
        probes = (given[0][r, :given[1][r]] if given is not None
                  else np.argsort(cd[r], kind="stable")[:nprobe])
        adc, who = [], []
        for l in probes:
            lut = orc.pq_lut(q[r], cen[l], n_sub, cb[l])
            adc.append(_adc_fold(lut, codes[off[l]:off[l + 1]].astype(np.int64)))
            who.append(lid[off[l]:off[l + 1]])
        adc, who = np.concatenate(adc), np.concatenate(who)
        cand = who[np.lexsort((who, adc))[:R]]
        ed = np.array([orc.metric(q[r], x[c], int(mt)) for c in cand], np.float32)
        order = np.lexsort((cand, ed))[:k]
        ids.append(cand[order])
        dist.append(ed[order])
        cnt.append(len(order))
    return ids, dist, np.array(cnt, np.uint32)
