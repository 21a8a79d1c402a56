"""Python mirror of hora's ANNIndex surface over the C ABI.

Names, argument meaning and error strings follow the reference:
  trait ANNIndex            src/core/ann_index.rs:25-121
  enum Metric               src/core/metrics.rs:5-13
  Node                      src/core/node.rs:93-167
  BruteForceIndex / Params  src/index/bruteforce_idx.rs, bruteforce_params.rs
  IVFPQIndex / IVFPQParams  src/index/pq_idx.rs:267-523, pq_params.rs:43-91
  PQIndex / PQParams        src/index/pq_idx.rs:21-265, pq_params.rs:7-41
Where the reference returns Err(&'static str) or panics, this mirror raises
HoraError carrying the status and the reference's message.

Everything runs through libhora_mi355x.so on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from enum import IntEnum
from typing import Any, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import HmParams, HoraError, check, lib

HM_INDEX_BRUTE_FORCE = 1
HM_INDEX_IVFPQ = 2
HM_INDEX_PQ = 3

HM_SEARCH_EXACT = 0
HM_SEARCH_BF16_RERANK = 1
HM_SEARCH_FP8_RERANK = 2


class Metric(IntEnum):
    """metrics.rs:5-13, same variant order."""
    Unknown = 0
    Manhattan = 1
    DotProduct = 2
    Euclidean = 3
    CosineSimilarity = 4
    Angular = 5


class Node:
    """node.rs:93-167: an owned f32 vector plus an optional id.  Non-finite
    elements are rejected (node.rs:158-166 panics)."""

    def __init__(self, vectors, idx: Any = None):
        v = np.array(vectors, dtype=np.float32, copy=True).reshape(-1)
        if not np.all(np.isfinite(v)):
            raise HoraError(3, "invalid float element")
        self._vectors = v
        self._idx = idx

    @classmethod
    def new(cls, vectors) -> "Node":
        return cls(vectors)

    @classmethod
    def new_with_idx(cls, vectors, idx) -> "Node":
        return cls(vectors, idx)

    def vectors(self) -> np.ndarray:
        return self._vectors

    def idx(self) -> Any:
        return self._idx

    def len(self) -> int:
        return int(self._vectors.size)

    def is_empty(self) -> bool:
        return self._vectors.size == 0

    def __repr__(self) -> str:
        return f"Node(key: {self._idx!r}, vectors: {self._vectors!r})"


class BruteForceParams:
    """bruteforce_params.rs:5-14 (no fields)."""


class IVFPQParams:
    """pq_params.rs:43-91 builder; defaults of IVFPQParams::default().
    seed / kmeans_accumulate are this build's additions (seeded split RNG,
    f32-point-order or fp64 centroid sums)."""

    def __init__(self):
        p = HmParams()
        check(lib().hm_params_default(HM_INDEX_IVFPQ, C.byref(p)))
        self._p = p

    def n_sub(self, v: int) -> "IVFPQParams":
        self._p.n_sub = v
        return self

    def sub_bits(self, v: int) -> "IVFPQParams":
        self._p.sub_bits = v
        return self

    def n_kmeans_center(self, v: int) -> "IVFPQParams":
        self._p.n_kmeans_center = v
        return self

    def search_n_center(self, v: int) -> "IVFPQParams":
        self._p.search_n_center = v
        return self

    def train_epoch(self, v: int) -> "IVFPQParams":
        self._p.train_epoch = v
        return self

    def seed(self, v: int) -> "IVFPQParams":
        self._p.seed = v
        return self

    def kmeans_accumulate(self, v: int) -> "IVFPQParams":
        self._p.kmeans_accumulate = v
        return self


class PQParams:
    """pq_params.rs:7-41 builder; defaults of PQParams::default() (n_sub 4,
    sub_bits 4, train_epoch 100).  seed / kmeans_accumulate as in
    IVFPQParams."""

    def __init__(self):
        p = HmParams()
        check(lib().hm_params_default(HM_INDEX_PQ, C.byref(p)))
        self._p = p

    def n_sub(self, v: int) -> "PQParams":
        self._p.n_sub = v
        return self

    def sub_bits(self, v: int) -> "PQParams":
        self._p.sub_bits = v
        return self

    def train_epoch(self, v: int) -> "PQParams":
        self._p.train_epoch = v
        return self

    def seed(self, v: int) -> "PQParams":
        self._p.seed = v
        return self

    def kmeans_accumulate(self, v: int) -> "PQParams":
        self._p.kmeans_accumulate = v
        return self


# bincode encodings of the reference's IdxType implementations (node.rs:76-87):
# (width in bytes, signed); String is u64-length-prefixed UTF-8 (width 0).
_ID_TYPES = {"usize": (8, False), "u64": (8, False), "i64": (8, True), "u32": (4, False),
             "i32": (4, True), "u16": (2, False), "i16": (2, True), "u128": (16, False),
             "i128": (16, True), "String": (0, False)}


def _id_codec(idx_type: str):
    if idx_type not in _ID_TYPES:
        raise HoraError(1, f"unsupported id type {idx_type!r}")
    return _ID_TYPES[idx_type]


def _encode_ids(ids, idx_type: str):
    """Option<T> bincode bytes of each id (None -> tag 0) and n+1 offsets."""
    width, signed = _id_codec(idx_type)
    parts = []
    for i in ids:
        if i is None:
            parts.append(b"\x00")
        elif width:
            parts.append(b"\x01" + int(i).to_bytes(width, "little", signed=signed))
        else:
            b = str(i).encode()
            parts.append(b"\x01" + len(b).to_bytes(8, "little") + b)
    off = np.zeros(len(parts) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in parts]) if parts else []
    return b"".join(parts), off


def _decode_ids(blob: bytes, off, idx_type: str):
    width, signed = _id_codec(idx_type)
    out = []
    for j in range(len(off) - 1):
        e = blob[int(off[j]):int(off[j + 1])]
        if e[0] == 0:
            out.append(None)
        elif width:
            out.append(int.from_bytes(e[1:], "little", signed=signed))
        else:
            out.append(e[9:].decode())
    return out


def _f32_rows(a, dim: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if a.shape[1] != dim:
        raise HoraError(2, "different dimensions")
    return a


class _Index:
    _kind = 0

    def __init__(self, dimension: int, params: Optional[Any] = None):
        h = C.c_void_p()
        p = params._p if isinstance(params, (IVFPQParams, PQParams)) else None
        check(lib().hm_index_create(self._kind, dimension,
                                    C.byref(p) if p is not None else None, C.byref(h)))
        self._h = h
        self._dim = int(dimension)
        self._ids: List[Any] = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().hm_index_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- ANNIndex
    def build(self, mt: Metric) -> None:
        """ann_index.rs:30."""
        check(lib().hm_index_build(self._h, int(mt)))

    def add_node(self, item: Node) -> None:
        """ann_index.rs:36."""
        self._add_rows(item.vectors().reshape(1, -1), [item.idx()])

    def add(self, vs, idx) -> None:
        """ann_index.rs:41-43."""
        self.add_node(Node.new_with_idx(vs, idx))

    def madd(self, vss: Sequence, indices: Sequence) -> None:
        """ann_index.rs:48-59."""
        if len(vss) != len(indices):
            raise HoraError(1, "vector's size is different with index")
        rows = np.stack([np.asarray(v, dtype=np.float32) for v in vss]) if len(vss) else \
            np.zeros((0, self._dim), np.float32)
        self._add_rows(rows, list(indices))

    def add_batch(self, rows, ids: Optional[Sequence] = None) -> None:
        """Bulk form of madd over an [n, dim] array (ids default to positions)."""
        rows = _f32_rows(rows, self._dim)
        if ids is None:
            ids = range(len(self._ids), len(self._ids) + rows.shape[0])
        self._add_rows(rows, list(ids))

    def add_device(self, ptr: int, n_rows: int, ids: Optional[Sequence] = None) -> None:
        """Append n_rows rows already in device memory (contiguous f32)."""
        check(lib().hm_index_add_device(self._h, ptr, n_rows))
        if ids is None:
            ids = range(len(self._ids), len(self._ids) + n_rows)
        self._ids.extend(ids)

    def _add_rows(self, rows, ids) -> None:
        rows = _f32_rows(rows, self._dim)
        check(lib().hm_index_add(self._h, rows.ctypes.data, rows.shape[0]))
        self._ids.extend(ids)

    def built(self) -> bool:
        out = C.c_int32()
        check(lib().hm_index_built(self._h, C.byref(out)))
        return bool(out.value)

    def rebuild(self, mt: Metric) -> None:
        """ann_index.rs:69-71: the trait default, which none of the index
        types overrides (raises HoraError "not implement")."""
        check(lib().hm_index_rebuild(self._h, int(mt)))

    def search_batch(self, queries, k: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """node_search_k for a batch: (positions u64 [nq,k], dist f32 [nq,k],
        count u32 [nq]); unused slots hold UINT64_MAX / +inf."""
        q = _f32_rows(queries, self._dim)
        nq = q.shape[0]
        idx = np.empty((nq, max(k, 1)), np.uint64)
        dist = np.empty((nq, max(k, 1)), np.float32)
        cnt = np.empty(max(nq, 1), np.uint32)
        check(lib().hm_index_search(self._h, q.ctypes.data, nq, k, idx.ctypes.data,
                                    dist.ctypes.data, cnt.ctypes.data))
        return idx[:, :k], dist[:, :k], cnt[:nq]

    def node_search_k(self, item: Node, k: int) -> List[Tuple[Node, float]]:
        """ann_index.rs:74: (Node, distance) ascending."""
        idx, dist, cnt = self.search_batch(item.vectors().reshape(1, -1), k)
        m = int(cnt[0])
        pos = idx[0, :m]
        rows = self.get_rows(pos)
        return [(Node(rows[j], self._ids[int(pos[j])]), float(dist[0, j])) for j in range(m)]

    def search_nodes(self, item, k: int) -> List[Tuple[Node, float]]:
        """ann_index.rs:81-84."""
        item = np.asarray(item, dtype=np.float32).reshape(-1)
        if item.size != self._dim:
            raise HoraError(2, "different dimensions")
        return self.node_search_k(Node.new(item), k)

    def search(self, item, k: int) -> List[Any]:
        """ann_index.rs:91-97: ids of the k nearest."""
        item = np.asarray(item, dtype=np.float32).reshape(-1)
        if item.size != self._dim:
            raise HoraError(2, "different dimensions")
        Node.new(item)  # validation (node.rs:158-166)
        idx, _, cnt = self.search_batch(item.reshape(1, -1), k)
        return [self._ids[int(i)] for i in idx[0, : int(cnt[0])]]

    def search_stats(self) -> Tuple[int, int]:
        """(rows rerun on the exact path or replayed, rows uncertified) of the
        last search."""
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().hm_index_search_stats(self._h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def get_rows(self, positions) -> np.ndarray:
        pos = np.ascontiguousarray(positions, dtype=np.uint64)
        out = np.empty((pos.size, self._dim), np.float32)
        check(lib().hm_index_get_rows(self._h, pos.ctypes.data, pos.size, out.ctypes.data))
        return out

    def name(self) -> str:
        return lib().hm_index_name(self._h).decode()

    def nodes_size(self) -> int:
        out = C.c_uint64()
        check(lib().hm_index_nodes_size(self._h, C.byref(out)))
        return int(out.value)

    def clear(self) -> None:
        check(lib().hm_index_clear(self._h))
        self._ids = []

    def idx_info(self) -> str:
        """ann_index.rs:113-115: the trait default text."""
        n = C.c_uint64()
        check(lib().hm_index_idx_info(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(int(n.value) + 1)
        check(lib().hm_index_idx_info(self._h, buf, int(n.value) + 1, None))
        return buf.value.decode()

    def dimension(self) -> int:
        out = C.c_uint64()
        check(lib().hm_index_dimension(self._h, C.byref(out)))
        return int(out.value)

    # ---- SerializableIndex (ann_index.rs:141-158)
    def dump(self, path: str, idx_type: str = "usize") -> None:
        """Write the index in the reference's bincode layout; ids are
        encoded as Option<idx_type>."""
        blob, off = _encode_ids(self._ids, idx_type)
        buf = np.frombuffer(blob, np.uint8) if blob else np.zeros(1, np.uint8)
        check(lib().hm_index_dump(self._h, str(path).encode(), buf.ctypes.data,
                                  off.ctypes.data))

    @classmethod
    def load(cls, path: str, idx_type: str = "usize"):
        """Read an index written by dump (or by the reference's dump)."""
        width, _ = _id_codec(idx_type)
        h = C.c_void_p()
        check(lib().hm_index_load(cls._kind, str(path).encode(), width, C.byref(h)))
        obj = cls.__new__(cls)
        obj._h = h
        d = C.c_uint64()
        check(lib().hm_index_dimension(h, C.byref(d)))
        obj._dim = int(d.value)
        n = C.c_uint64()
        check(lib().hm_index_nodes_size(h, C.byref(n)))
        total = C.c_uint64()
        check(lib().hm_index_loaded_ids(h, None, None, C.byref(total)))
        blob = np.zeros(max(int(total.value), 1), np.uint8)
        off = np.zeros(int(n.value) + 1, np.uint64)
        check(lib().hm_index_loaded_ids(h, blob.ctypes.data, off.ctypes.data, C.byref(total)))
        obj._ids = _decode_ids(blob.tobytes(), off, idx_type)
        return obj

    # ---- device-resident search (inputs/outputs already in HBM)
    def search_device(self, q_ptr: int, nq: int, k: int, idx_ptr: int, dist_ptr: int,
                      count_ptr: int, stream: int = 0) -> None:
        check(lib().hm_index_search_device(self._h, q_ptr, nq, k, idx_ptr, dist_ptr, count_ptr,
                                           stream or None))

    def search_device_noalloc(self, q_ptr: int, nq: int, k: int, idx_ptr: int, dist_ptr: int,
                              count_ptr: int, stream: int = 0) -> None:
        """search_device in the exact form without any allocation (the
        library's fixed per-GPU scratch array; see the C header)."""
        check(lib().hm_index_search_device_noalloc(self._h, q_ptr, nq, k, idx_ptr, dist_ptr,
                                                   count_ptr, stream or None))


class BruteForceIndex(_Index):
    """bruteforce_idx.rs:15-106."""
    _kind = HM_INDEX_BRUTE_FORCE

    def __init__(self, dimension: int, params: Optional[BruteForceParams] = None):
        super().__init__(dimension, None)

    def set_search_mode(self, mode: int, shortlist: int = 0) -> None:
        """EXACT, BF16_RERANK or FP8_RERANK; shortlist 0 = automatic."""
        check(lib().hm_index_set_search_mode(self._h, mode, shortlist))

    def search_ragged_device(self, q_ptr: int, batch_offsets, batch_k, idx_ptr: int,
                             dist_ptr: int, count_ptr: int, stream: int = 0) -> None:
        """Ragged batches with per-batch k on device buffers (see the C header)."""
        off = np.ascontiguousarray(batch_offsets, dtype=np.uint32)
        ks = np.ascontiguousarray(batch_k, dtype=np.uint32)
        if off.size != ks.size + 1:
            raise HoraError(1, "batch_offsets needs one entry more than batch_k")
        check(lib().hm_index_search_ragged_device(self._h, q_ptr, ks.size, off.ctypes.data,
                                                  ks.ctypes.data, idx_ptr, dist_ptr, count_ptr,
                                                  stream or None))

    def distances(self, queries, id_lists) -> List[np.ndarray]:
        """Batched distance service of the graph searches (hnsw_idx.rs:243-254,
        ssg_idx.rs:402-456): for query q, metric(node, q) of every node
        index in id_lists[q]."""
        q = _f32_rows(queries, self._dim)
        if len(id_lists) != q.shape[0]:
            raise HoraError(1, "one id list per query")
        lens = [len(x) for x in id_lists]
        off = np.zeros(q.shape[0] + 1, np.uint64)
        off[1:] = np.cumsum(lens) if lens else []
        ids = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32).reshape(-1)
                                                   for x in id_lists])
                                   if lens and sum(lens) else np.zeros(1, np.uint32),
                                   dtype=np.uint32)
        out = np.empty(max(int(off[-1]), 1), np.float32)
        check(lib().hm_index_distances(self._h, q.ctypes.data, q.shape[0], ids.ctypes.data,
                                       off.ctypes.data, out.ctypes.data))
        return [out[int(off[i]):int(off[i + 1])] for i in range(q.shape[0])]

    def distances_device(self, q_ptr: int, nq: int, ids_ptr: int, off_ptr: int, total: int,
                         out_ptr: int, stream: int = 0) -> None:
        check(lib().hm_index_distances_device(self._h, q_ptr, nq, ids_ptr, off_ptr, total,
                                              out_ptr, stream or None))

    def build_knn_graph(self, k: int):
        """naive_build_knn_graph (knn.rs:14-41) over the indexed rows with the
        build metric: (idx u64 [n,k], dist f32 [n,k], count u32 [n]); row i
        holds its k nearest other rows, largest distance first."""
        n = self.nodes_size()
        idx = np.empty((max(n, 1), max(k, 1)), np.uint64)
        dist = np.empty((max(n, 1), max(k, 1)), np.float32)
        cnt = np.empty(max(n, 1), np.uint32)
        check(lib().hm_index_knn_graph(self._h, k, idx.ctypes.data, dist.ctypes.data,
                                       cnt.ctypes.data))
        return idx[:n, :k], dist[:n, :k], cnt[:n]

    def autotune(self, queries, k: int) -> int:
        """Times the shortlist depths, then the score tile group sizes, on
        these queries and keeps the fastest; returns the depth."""
        q = _f32_rows(queries, self._dim)
        out = C.c_uint32()
        check(lib().hm_index_autotune(self._h, q.ctypes.data, q.shape[0], k, C.byref(out)))
        return int(out.value)

    def tuning(self) -> dict:
        """The shortlist depth (0: automatic) and score tile group (KiB) in use."""
        sl, g = C.c_uint32(), C.c_uint32()
        check(lib().hm_index_tuning(self._h, C.byref(sl), C.byref(g)))
        return {"shortlist": int(sl.value), "score_group_kib": int(g.value)}




class IVFPQIndex(_Index):
    """pq_idx.rs:267-523."""
    _kind = HM_INDEX_IVFPQ

    def __init__(self, dimension: int, params: Optional[IVFPQParams] = None):
        super().__init__(dimension, params if params is not None else IVFPQParams())

    def set_search_n_center(self, nprobe: int) -> None:
        check(lib().hm_index_set_nprobe(self._h, nprobe))

    def train_probe(self, train_queries, label_k: int = 10, hidden: int = 512, epochs: int = 4,
                    batch: int = 1024, lr: float = 1e-3, seed: int = 1) -> float:
        """Train the learned-probe network on the GPU; returns the last
        epoch's mean cross-entropy."""
        q = _f32_rows(train_queries, self._dim)
        loss = C.c_double()
        check(lib().hm_ivfpq_train_probe(self._h, q.ctypes.data, q.shape[0], label_k, hidden,
                                         epochs, batch, lr, seed, C.byref(loss)))
        return float(loss.value)

    SCAN_FORMS = {"auto": 0, "per_probe": 1, "lists": 2, "rows": 3}

    def set_scan_form(self, form: str) -> None:
        """How search scans the probed lists (same results): "auto",
        "per_probe", "lists" or "rows"."""
        check(lib().hm_ivfpq_set_scan_form(self._h, self.SCAN_FORMS[form]))

    def set_rerank(self, depth: int) -> None:
        """Exact rerank of the ADC top-max(depth, k) (0: the reference's ADC
        result; see the C header)."""
        check(lib().hm_ivfpq_set_rerank(self._h, depth))

    def set_probe_policy(self, mode: int, mass: float = 0.9, pmin: int = 1,
                         pmax: int = 64) -> None:
        """0: fixed search_n_center (reference); 1: learned per-query nprobe."""
        check(lib().hm_ivfpq_set_probe_policy(self._h, mode, mass, pmin, pmax))

    def last_probes(self, nq: int, ldp: int):
        probes = np.zeros((nq, ldp), np.uint32)
        counts = np.zeros(nq, np.uint32)
        check(lib().hm_ivfpq_last_probes(self._h, nq, ldp, probes.ctypes.data,
                                         counts.ctypes.data))
        return probes, counts

    def mean_nprobe(self) -> float:
        out = C.c_double()
        check(lib().hm_ivfpq_mean_nprobe(self._h, C.byref(out)))
        return float(out.value)

    def info(self) -> dict:
        out = np.zeros(8, np.uint32)
        check(lib().hm_ivfpq_info(self._h, out.ctypes.data))
        keys = ["nlist", "n_sub", "n_sub_center", "sub_dim_max", "n_trained", "trained",
                "nprobe", "dpad"]
        return dict(zip(keys, out.tolist()))

    def export(self) -> dict:
        """Trained state as arrays (centers, list_offsets, list_ids, codebooks,
        codes); see hm_ivfpq_export."""
        i = self.info()
        nl, m, ks, sdm, n = i["nlist"], i["n_sub"], i["n_sub_center"], i["sub_dim_max"], \
            i["n_trained"]
        cen = np.zeros((nl, self._dim), np.float32)
        off = np.zeros(nl + 1, np.uint64)
        ids = np.zeros(max(n, 1), np.uint64)
        cb = np.zeros((nl, m, ks, sdm), np.float32)
        codes = np.zeros((max(n, 1), m), np.uint8)
        check(lib().hm_ivfpq_export(self._h, cen.ctypes.data, off.ctypes.data, ids.ctypes.data,
                                    cb.ctypes.data, codes.ctypes.data))
        return {"centers": cen, "list_offsets": off, "list_ids": ids[:n], "codebooks": cb,
                "codes": codes[:n]}

    def import_trained(self, state: dict, mt: Metric) -> None:
        """Install a trained state (the arrays of export(); codes in list
        order) on this index, which must hold the same rows; it is then
        built for mt (see hm_ivfpq_import)."""
        cen = np.ascontiguousarray(state["centers"], dtype=np.float32)
        off = np.ascontiguousarray(state["list_offsets"], dtype=np.uint64)
        ids = np.ascontiguousarray(state["list_ids"], dtype=np.uint64)
        cb = np.ascontiguousarray(state["codebooks"], dtype=np.float32)
        codes = np.ascontiguousarray(state["codes"], dtype=np.uint8)
        i = self.info()
        nl, m, ks, sdm = i["nlist"], i["n_sub"], i["n_sub_center"], i["sub_dim_max"]
        if cen.shape != (nl, self._dim) or off.shape != (nl + 1,) or \
                cb.shape != (nl, m, ks, sdm) or codes.shape != (ids.shape[0], m):
            raise HoraError(1, "trained state does not match the index shape")
        check(lib().hm_ivfpq_import(self._h, int(mt), cen.ctypes.data, off.ctypes.data,
                                    ids.ctypes.data if ids.size else None, cb.ctypes.data,
                                    codes.ctypes.data if codes.size else None))


PQIVFIndex = IVFPQIndex  # README.md:57 name


class PQIndex(_Index):
    """pq_idx.rs:21-265: product quantization over the whole base (no
    coarse quantizer, no residual); the sub-space k-means and the LUT use
    the build metric."""
    _kind = HM_INDEX_PQ

    def __init__(self, dimension: int, params: Optional[PQParams] = None):
        super().__init__(dimension, params if params is not None else PQParams())

    def info(self) -> dict:
        out = np.zeros(8, np.uint32)
        check(lib().hm_pq_info(self._h, out.ctypes.data))
        keys = ["n_sub", "n_sub_center", "sub_dim_max", "n_trained", "trained", "dpad",
                "metric", "reserved"]
        return dict(zip(keys, out.tolist()))

    def export(self) -> dict:
        """Trained state: codebook [n_sub][2^sub_bits][sub_dim_max] and codes
        [n][n_sub] (see hm_pq_export)."""
        i = self.info()
        m, ks, sdm, n = i["n_sub"], i["n_sub_center"], i["sub_dim_max"], i["n_trained"]
        cb = np.zeros((m, ks, sdm), np.float32)
        codes = np.zeros((max(n, 1), m), np.uint8)
        check(lib().hm_pq_export(self._h, cb.ctypes.data, codes.ctypes.data))
        return {"codebook": cb, "codes": codes[:n]}

    def import_trained(self, state: dict, mt: Metric) -> None:
        """Install a codebook and the codes of the first len(codes) rows (the
        arrays of export()); the index is then built for mt (see
        hm_pq_import)."""
        cb = np.ascontiguousarray(state["codebook"], dtype=np.float32)
        codes = np.ascontiguousarray(state["codes"], dtype=np.uint8)
        i = self.info()
        m, ks, sdm = i["n_sub"], i["n_sub_center"], i["sub_dim_max"]
        if cb.shape != (m, ks, sdm) or codes.ndim != 2 or codes.shape[1] != m:
            raise HoraError(1, "trained state does not match the index shape")
        check(lib().hm_pq_import(self._h, int(mt), cb.ctypes.data,
                                 codes.ctypes.data if codes.size else None, codes.shape[0]))


def naive_build_knn_graph(rows, mt: Metric, k: int) -> List[List[Tuple[int, float]]]:
    """knn.rs:14-41: graph[n] = the k nearest other rows of row n as
    (index, distance), in the order the reference's heap drains them
    (largest distance first)."""
    rows = np.ascontiguousarray(rows, dtype=np.float32)
    ix = BruteForceIndex(rows.shape[1])
    ix.add_batch(rows)
    ix.build(mt)
    idx, dist, cnt = ix.build_knn_graph(k)
    return [[(int(idx[r, j]), float(dist[r, j])) for j in range(int(cnt[r]))]
            for r in range(rows.shape[0])]


def mlp_loss_grad(W1, b1, W2, b2, X, T=None):
    """Probe-network forward/backward on the GPU for given weights (gradient
    checks).  Returns dict(loss, gW1, gb1, gW2, gb2, logits)."""
    W1, b1, W2, b2, X = (np.ascontiguousarray(a, dtype=np.float32) for a in (W1, b1, W2, b2, X))
    din, hid = W1.shape
    dout = W2.shape[1]
    n = X.shape[0]
    out = {"logits": np.zeros((n, dout), np.float32)}
    ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
    if T is not None:
        T = np.ascontiguousarray(T, dtype=np.float32)
        out.update(loss=np.zeros(n, np.float32), gW1=np.zeros_like(W1), gb1=np.zeros_like(b1),
                   gW2=np.zeros_like(W2), gb2=np.zeros_like(b2))
    check(lib().hm_mlp_loss_grad(din, hid, dout, n, ptr(W1), ptr(b1), ptr(W2), ptr(b2), ptr(X),
                                 ptr(T), ptr(out.get("loss")), ptr(out.get("gW1")),
                                 ptr(out.get("gb1")), ptr(out.get("gW2")), ptr(out.get("gb2")),
                                 ptr(out["logits"])))
    return out
