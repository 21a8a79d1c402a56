"""ctypes front end of the CPU oracle (oracle/hora_oracle.c).

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it; the
product library never does.

The C file restates the reference's scalar arithmetic and std-BinaryHeap
top-k (see the header of hora_oracle.c for the file:line of each function
and for which check pins it).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

EUCLIDEAN, MANHATTAN, DOT, COSINE, ANGULAR, UNKNOWN = 3, 1, 2, 4, 5, 0

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_sz = C.c_size_t


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU)."""
    if force or not os.path.exists(_LIB_PATH) or (
        os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "hora_oracle.c"))
    ):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.orc_metric.argtypes = [_f32p, _f32p, _sz, C.c_int, C.POINTER(C.c_float)]
    L.orc_metric.restype = C.c_int
    L.orc_metric_f64.argtypes = [_f32p, _f32p, _sz, C.c_int]
    L.orc_metric_f64.restype = C.c_double
    L.orc_cosine_reference.argtypes = [_f32p, _f32p, _sz]
    L.orc_cosine_reference.restype = C.c_float
    L.orc_dot_neg.argtypes = [_f32p, _f32p, _sz]
    L.orc_dot_neg.restype = C.c_float
    L.orc_heap_script.argtypes = [_i32p, _f32p, _u64p, _sz, _f32p, _u64p,
                                  C.POINTER(_sz), _f32p, _u64p, C.POINTER(_sz)]
    L.orc_bf_search.argtypes = [_f32p, _sz, _sz, _f32p, _sz, _sz, C.c_int, C.c_int,
                                _u64p, _f32p, _u32p]
    L.orc_distance_rows.argtypes = [_f32p, _sz, _sz, _f32p, _sz, C.c_int, C.c_int, _f32p]
    L.orc_replay_row.argtypes = [_f32p, _sz, _sz, _u64p, _f32p, _u32p]
    L.orc_splitmix64_next.argtypes = [C.POINTER(C.c_uint64)]
    L.orc_splitmix64_next.restype = C.c_uint64
    L.orc_stream_seed.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_stream_seed.restype = C.c_uint64
    L.orc_kmeans_train.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_void_p, _sz, C.c_int, _sz,
                                   C.c_uint64, _f32p, _u32p]
    L.orc_kmeans_init.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_void_p, _sz, _f32p]
    L.orc_kmeans_assign.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_void_p, _f32p, _sz,
                                    C.c_int, C.c_int, _u32p]
    L.orc_kmeans_update.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_void_p, _u32p, _sz,
                                    _f32p, _u64p]
    L.orc_kmeans_split.argtypes = [_f32p, _sz, _sz, _sz, _u64p, C.POINTER(C.c_uint64)]
    L.orc_pq_ranges.argtypes = [_sz, _sz, _u32p, _u32p]
    L.orc_pq_train.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_void_p, _sz, C.c_uint64,
                               C.c_uint64, _f32p, _sz, _u32p]
    L.orc_pq_lut.argtypes = [_f32p, C.c_void_p, _sz, _sz, _f32p, _sz, _sz, _f32p]
    L.orc_pq_encode.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_void_p, _f32p, _sz, _u32p]
    L.orc_ivfpq_search.argtypes = [_f32p, _sz, _sz, C.c_int, _f32p, _sz, _f32p, _sz, _sz, _sz,
                                   _u64p, _u64p, _u32p, _sz, _sz, C.c_int, _u64p, _f32p,
                                   _u32p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_pqindex_train.argtypes = [_f32p, _sz, _sz, _sz, _sz, C.c_int, _sz, C.c_uint64,
                                    _f32p, _sz, _u32p]
    L.orc_pqindex_search.argtypes = [_f32p, _sz, _sz, C.c_int, _f32p, _sz, _sz, _sz, _u32p,
                                     _sz, _sz, C.c_int, _u64p, _f32p, _u32p]
    L.orc_knn_graph.argtypes = [_f32p, _sz, _sz, _sz, C.c_int, C.c_int, _u64p, _f32p, _u32p]
    L.orc_pqindex_lut.argtypes = [_f32p, _sz, _sz, _f32p, _sz, _sz, C.c_int, _f32p]
    _lib = L
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: oracle status {status}")
        self.status = status


def metric(a, b, mt: int) -> float:
    a, b = _f32(a), _f32(b)
    out = C.c_float()
    st = lib().orc_metric(a, b, a.size, mt, C.byref(out))
    if st:
        raise OracleError(st, "metric")
    return out.value


def metric_f64(a, b, mt: int) -> float:
    a, b = _f32(a), _f32(b)
    return lib().orc_metric_f64(a, b, a.size, mt)


def cosine_reference(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return lib().orc_cosine_reference(a, b, a.size)


def dot_neg(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return lib().orc_dot_neg(a, b, a.size)


def heap_script(ops, dists, ids):
    """Run push (op 0) / pop (op 1) operations on the std-BinaryHeap model.
    Returns (layout [(dist,id)...], popped [(dist,id)...])."""
    ops = np.ascontiguousarray(ops, dtype=np.int32)
    d = _f32(dists)
    i = np.ascontiguousarray(ids, dtype=np.uint64)
    n = ops.size
    ld, li = np.zeros(n + 1, np.float32), np.zeros(n + 1, np.uint64)
    pd, pi = np.zeros(n + 1, np.float32), np.zeros(n + 1, np.uint64)
    ln, pn = _sz(), _sz()
    lib().orc_heap_script(ops, d, i, n, ld, li, C.byref(ln), pd, pi, C.byref(pn))
    layout = [(float(ld[j]), int(li[j])) for j in range(ln.value)]
    popped = [(float(pd[j]), int(pi[j])) for j in range(pn.value)]
    return layout, popped


def bf_search(base, queries, k: int, mt: int, nthreads: int = 0):
    """BruteForceIndex::node_search_k for every query.  Returns
    (ids uint64 [nq,k], dist f32 [nq,k], count u32 [nq])."""
    base, queries = _f32(base), _f32(queries)
    n, d = base.shape
    nq = queries.shape[0]
    kk = max(k, 1)
    ids = np.zeros((nq, kk), np.uint64)
    dist = np.zeros((nq, kk), np.float32)
    cnt = np.zeros(nq, np.uint32)
    st = lib().orc_bf_search(base, n, d, queries, nq, k, mt, nthreads, ids, dist, cnt)
    if st:
        raise OracleError(st, "bf_search")
    return ids[:, :k], dist[:, :k], cnt


def distance_rows(base, queries, mt: int, nthreads: int = 0):
    base, queries = _f32(base), _f32(queries)
    n, d = base.shape
    nq = queries.shape[0]
    out = np.zeros((nq, n), np.float32)
    st = lib().orc_distance_rows(base, n, d, queries, nq, mt, nthreads, out)
    if st:
        raise OracleError(st, "distance_rows")
    return out


def replay_row(row, k: int):
    row = _f32(row)
    kk = max(k, 1)
    ids = np.zeros(kk, np.uint64)
    dist = np.zeros(kk, np.float32)
    cnt = np.zeros(1, np.uint32)
    lib().orc_replay_row(row, row.size, k, ids, dist, cnt)
    m = int(cnt[0])
    return ids[:m], dist[:m]


def splitmix64_stream(seed: int, n: int):
    s = C.c_uint64(seed)
    return [lib().orc_splitmix64_next(C.byref(s)) for _ in range(n)]


def stream_seed(seed: int, instance: int) -> int:
    return lib().orc_stream_seed(seed, instance)


def kmeans_train(data, n_center: int, epochs: int, seed: int, mt: int = EUCLIDEAN,
                 begin: int = 0, dim: int | None = None, residual=None):
    data = _f32(data)
    n, stride = data.shape
    dim = stride - begin if dim is None else dim
    res = None if residual is None else _f32(residual)
    cent = np.zeros((n_center, dim), np.float32)
    assign = np.zeros(max(n, 1), np.uint32)
    st = lib().orc_kmeans_train(data, n, stride, begin, dim, _ptr(res), n_center, mt, epochs,
                                seed, cent, assign)
    if st:
        raise OracleError(st, "kmeans_train")
    return cent, assign[:n]


def kmeans_init(data, n_center: int, begin: int = 0, dim: int | None = None, residual=None):
    data = _f32(data)
    n, stride = data.shape
    dim = stride - begin if dim is None else dim
    res = None if residual is None else _f32(residual)
    cent = np.zeros((n_center, dim), np.float32)
    lib().orc_kmeans_init(data, n, stride, begin, dim, _ptr(res), n_center, cent)
    return cent


def kmeans_assign(data, centers, mt: int = EUCLIDEAN, begin: int = 0, residual=None,
                  nthreads: int = 0):
    data, centers = _f32(data), _f32(centers)
    n, stride = data.shape
    nc, dim = centers.shape
    res = None if residual is None else _f32(residual)
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_kmeans_assign(data, n, stride, begin, dim, _ptr(res), centers, nc, mt, nthreads,
                            out)
    return out[:n]


def kmeans_update(data, assign, n_center: int, begin: int = 0, dim: int | None = None,
                  residual=None):
    data = _f32(data)
    n, stride = data.shape
    dim = stride - begin if dim is None else dim
    res = None if residual is None else _f32(residual)
    a = np.ascontiguousarray(assign, dtype=np.uint32)
    cent = np.zeros((n_center, dim), np.float32)
    cnt = np.zeros(n_center, np.uint64)
    lib().orc_kmeans_update(data, n, stride, begin, dim, _ptr(res), a, n_center, cent, cnt)
    return cent, cnt


def kmeans_split(centers, counts, n: int, rng_state: int):
    cent = np.array(centers, dtype=np.float32, copy=True, order="C")
    cnt = np.array(counts, dtype=np.uint64, copy=True)
    s = C.c_uint64(rng_state)
    st = lib().orc_kmeans_split(cent, cent.shape[0], cent.shape[1], n, cnt, C.byref(s))
    if st:
        raise OracleError(st, "kmeans_split")
    return cent, cnt, s.value


def pq_ranges(d: int, n_sub: int):
    b = np.zeros(n_sub, np.uint32)
    e = np.zeros(n_sub, np.uint32)
    st = lib().orc_pq_ranges(d, n_sub, b, e)
    if st:
        raise OracleError(st, "pq_ranges")
    return list(zip(b.tolist(), e.tolist()))


def pq_train(items, n_sub: int, n_center: int, residual, epochs: int, seed: int,
             first_instance: int, sub_dim_max: int):
    items = _f32(items)
    n, d = items.shape
    res = None if residual is None else _f32(residual)
    cb = np.zeros((n_sub, n_center, sub_dim_max), np.float32)
    codes = np.zeros((max(n, 1), n_sub), np.uint32)
    st = lib().orc_pq_train(items, n, d, n_sub, n_center, _ptr(res), epochs, seed,
                            first_instance, cb, sub_dim_max, codes)
    if st:
        raise OracleError(st, "pq_train")
    return cb, codes[:n]


def pq_lut(q, residual, n_sub: int, codebook):
    q = _f32(q)
    cb = _f32(codebook)
    _, n_center, sdm = cb.shape
    res = None if residual is None else _f32(residual)
    lut = np.zeros(n_sub * n_center, np.float32)
    st = lib().orc_pq_lut(q, _ptr(res), q.size, n_sub, cb, n_center, sdm, lut)
    if st:
        raise OracleError(st, "pq_lut")
    return lut.reshape(n_sub, n_center)


def pq_encode(items, n_sub: int, codebook, residual):
    items = _f32(items)
    n, d = items.shape
    cb = _f32(codebook)
    _, n_center, sdm = cb.shape
    res = None if residual is None else _f32(residual)
    codes = np.zeros((max(n, 1), n_sub), np.uint32)
    lib().orc_pq_encode(items, n, d, n_sub, n_center, _ptr(res), cb, sdm, codes)
    return codes[:n]


def ivfpq_search(queries, mt: int, centers, codebooks, list_offsets, list_ids, codes,
                 nprobe: int, k: int, nthreads: int = 0, given_probes=None, given_counts=None):
    """IVFPQIndex::node_search_k on an index given as arrays (see C header).
    given_probes [nq][nprobe] / given_counts [nq]: scan these lists instead
    of the coarse top-nprobe."""
    q = _f32(queries)
    nq, d = q.shape
    cen = _f32(centers)
    nlist = cen.shape[0]
    cb = _f32(codebooks)
    _, n_sub, n_center, sdm = cb.shape
    off = np.ascontiguousarray(list_offsets, dtype=np.uint64)
    lid = np.ascontiguousarray(list_ids, dtype=np.uint64)
    cd = np.ascontiguousarray(codes, dtype=np.uint32)
    kk = max(k, 1)
    ids = np.zeros((nq, kk), np.uint64)
    dist = np.zeros((nq, kk), np.float32)
    cnt = np.zeros(nq, np.uint32)
    probes = np.zeros((nq, max(nprobe, 1)), np.uint32)
    gp = gc = None
    if given_probes is not None:
        gp = np.ascontiguousarray(given_probes, dtype=np.uint32)
        gc = np.ascontiguousarray(given_counts, dtype=np.uint32)
    st = lib().orc_ivfpq_search(q, nq, d, mt, cen, nlist, cb, n_sub, n_center, sdm, off, lid,
                                cd, nprobe, k, nthreads, ids, dist, cnt,
                                probes.ctypes.data_as(C.c_void_p), _ptr(gp), _ptr(gc))
    if st:
        raise OracleError(st, "ivfpq_search")
    return ids[:, :k], dist[:, :k], cnt, probes[:, :nprobe]


def pq_sub_dim_max(d: int, n_sub: int) -> int:
    sd = d // n_sub
    return sd + (1 if d % sd else 0)


def pqindex_train(items, n_sub: int, n_center: int, mt: int, epochs: int, seed: int):
    """Standalone PQIndex build (pq_idx.rs:122-149, 198-202): codebook
    [n_sub][n_center][sub_dim_max], codes [n][n_sub]."""
    items = _f32(items)
    n, d = items.shape
    sdm = pq_sub_dim_max(d, n_sub)
    cb = np.zeros((n_sub, n_center, sdm), np.float32)
    codes = np.zeros((max(n, 1), n_sub), np.uint32)
    st = lib().orc_pqindex_train(items, n, d, n_sub, n_center, mt, epochs, seed, cb, sdm, codes)
    if st:
        raise OracleError(st, "pqindex_train")
    return cb, codes[:n]


def pqindex_search(queries, mt: int, codebook, codes, k: int, nthreads: int = 0):
    """Standalone PQIndex::node_search_k (pq_idx.rs:165-194, 212-233) on a
    given codebook and codes: (ids [nq,k], dist [nq,k], count [nq])."""
    q = _f32(queries)
    nq, d = q.shape
    cb = _f32(codebook)
    n_sub, n_center, sdm = cb.shape
    cd = np.ascontiguousarray(codes, dtype=np.uint32)
    n = cd.shape[0]
    kk = max(k, 1)
    ids = np.zeros((nq, kk), np.uint64)
    dist = np.zeros((nq, kk), np.float32)
    cnt = np.zeros(max(nq, 1), np.uint32)
    st = lib().orc_pqindex_search(q, nq, d, mt, cb, n_sub, n_center, sdm, cd, n, k, nthreads,
                                  ids, dist, cnt)
    if st:
        raise OracleError(st, "pqindex_search")
    return ids[:, :k], dist[:, :k], cnt[:nq]


def pqindex_lut(q, mt: int, codebook):
    q = _f32(q).reshape(-1)
    cb = _f32(codebook)
    n_sub, n_center, sdm = cb.shape
    lut = np.zeros(n_sub * n_center, np.float32)
    st = lib().orc_pqindex_lut(q, q.size, n_sub, cb, n_center, sdm, mt, lut)
    if st:
        raise OracleError(st, "pqindex_lut")
    return lut.reshape(n_sub, n_center)


def knn_graph(base, k: int, mt: int, nthreads: int = 0):
    """naive_build_knn_graph (knn.rs:14-41): (idx [n,k], dist [n,k], count
    [n]), rows largest distance first."""
    base = _f32(base)
    n, d = base.shape
    kk = max(k, 1)
    ids = np.zeros((max(n, 1), kk), np.uint64)
    dist = np.zeros((max(n, 1), kk), np.float32)
    cnt = np.zeros(max(n, 1), np.uint32)
    st = lib().orc_knn_graph(base, n, d, k, mt, nthreads, ids, dist, cnt)
    if st:
        raise OracleError(st, "knn_graph")
    return ids[:n, :k], dist[:n, :k], cnt[:n]
