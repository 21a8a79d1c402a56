"""Restatement of the reference's on-disk index layout (TEST INFRASTRUCTURE:
the checker for hm_index_dump / hm_index_load; the product never imports it).

The reference writes its index structs with serde's derive and
bincode::serialize (bincode 1.3, Cargo.toml:47; the crate is not vendored).
bincode 1.x's default configuration, as its documentation states it:
little-endian; integers at fixed width, usize as u64; sequence and string
lengths as a u64 prefix; enum variants as a u32 index; bool as one byte;
Option as a one-byte tag (0 None, 1 Some) followed by the value; structs as
their fields in declaration order.  Field orders follow
  BruteForceIndex  bruteforce_idx.rs:15-22 (`nodes` has serde(skip))
  PQIndex          pq_idx.rs:21-45
  IVFPQIndex       pq_idx.rs:267-292
  Node             node.rs:93-97
Pinned by a hand-assembled byte string in tests/test_oracle.py.
"""
from __future__ import annotations

import struct

import numpy as np

MAX_ITEM = 100000  # pq_idx.rs:66, 319


def u64(v: int) -> bytes:
    return struct.pack("<Q", int(v))


def u32(v: int) -> bytes:
    return struct.pack("<I", int(v))


def boolean(b: bool) -> bytes:
    return b"\x01" if b else b"\x00"


def f32s(a) -> bytes:
    return np.ascontiguousarray(a, dtype="<f4").tobytes()


def usizes(a) -> bytes:
    a = np.asarray(a, dtype="<u8").reshape(-1)
    return u64(a.size) + a.tobytes()


def option_usize(i) -> bytes:
    return b"\x00" if i is None else b"\x01" + u64(i)


def node(vec, id_bytes: bytes) -> bytes:
    vec = np.asarray(vec, dtype=np.float32).reshape(-1)
    return u64(vec.size) + f32s(vec) + id_bytes


def nodes(rows, ids) -> bytes:
    """Vec<Node>: rows [n][d], ids: list of Option<T> encodings."""
    return u64(len(rows)) + b"".join(node(r, i) for r, i in zip(rows, ids))


def pq_ranges(d: int, n_sub: int):
    # pq_idx.rs:75-88
    sd = d // n_sub
    r = d % sd
    out = []
    for i in range(n_sub):
        if i < r:
            out.append((i * (sd + 1), (i + 1) * (sd + 1)))
        else:
            out.append((r * (sd + 1) + (i - r) * sd, r * (sd + 1) + (i + 1 - r) * sd))
    return out


def bruteforce(rows, ids, mt: int, dim: int) -> bytes:
    return nodes(rows, ids) + u32(mt) + u64(dim)


def pq_struct(dim: int, n_sub: int, sub_bits: int, epochs: int, codebook, trained: bool,
              residual, rows, ids, codes, mt: int) -> bytes:
    """One PQIndex.  codebook [M][KS][sdm] (sub s keeps its own dims) or
    None; codes [n][M] or None; residual a vector or None."""
    ks = 1 << sub_bits
    sub_bytes = (sub_bits + 7) // 8
    rng = pq_ranges(dim, n_sub)
    out = [u64(dim), u64(n_sub), u64(dim // n_sub), u64(n_sub)]
    out += [u64(2) + u64(b) + u64(e) for b, e in rng]
    out += [u64(sub_bits), u64(sub_bytes), u64(ks), u64(sub_bytes * n_sub), u64(epochs)]
    if codebook is None:
        out.append(u64(0))
    else:
        out.append(u64(n_sub))
        for s, (b, e) in enumerate(rng):
            out.append(u64(ks))
            for c in range(ks):
                out.append(u64(e - b) + f32s(codebook[s][c][: e - b]))
    out += [boolean(trained), boolean(residual is not None)]
    out.append(u64(0) if residual is None else u64(len(residual)) + f32s(residual))
    out += [u64(len(rows)), u64(MAX_ITEM), nodes(rows, ids)]
    if codes is None:
        out.append(u64(0))
    else:
        codes = np.asarray(codes)
        out.append(u64(n_sub) + b"".join(usizes(codes[:, s]) for s in range(n_sub)))
    out += [u32(mt), nodes(rows, ids)]
    return b"".join(out)


def pqindex(dim, n_sub, sub_bits, epochs, codebook, codes, rows, ids, mt) -> bytes:
    """A standalone PQIndex after build(mt) (trained), or before (codebook
    None: mt is PQIndex::new's Euclidean)."""
    trained = codebook is not None
    return pq_struct(dim, n_sub, sub_bits, epochs, codebook, trained, None, rows, ids,
                     codes if trained else None, mt if trained else 3)


def ivfpq(dim, n_sub, sub_bits, epochs, nprobe, nlist, centers, list_offsets, list_ids,
          codebooks, codes, rows, ids, mt) -> bytes:
    """IVFPQIndex: centers [nlist][dim], CSR lists (ids ascending per
    list), codebooks [nlist][M][KS][sdm], codes [n][M] in list order; centers
    None for an untrained index."""
    ks = 1 << sub_bits
    sub_bytes = (sub_bits + 7) // 8
    trained = centers is not None
    out = [u64(dim), u64(n_sub), u64(dim // n_sub), u64(sub_bits), u64(sub_bytes), u64(ks),
           u64(sub_bytes * n_sub), u64(epochs), u64(nprobe), u64(nlist)]
    if trained:
        out.append(u64(nlist) + b"".join(u64(dim) + f32s(c) for c in centers))
    else:
        out.append(u64(0))
    out.append(u64(nlist))
    for l in range(nlist):
        if trained:
            out.append(usizes(list_ids[list_offsets[l]:list_offsets[l + 1]]))
        else:
            out.append(u64(0))
    if trained:
        out.append(u64(nlist))
        for l in range(nlist):
            a, b = int(list_offsets[l]), int(list_offsets[l + 1])
            mem = [int(i) for i in list_ids[a:b]]
            out.append(pq_struct(dim, n_sub, sub_bits, epochs, codebooks[l], True, centers[l],
                                 [rows[i] for i in mem], [ids[i] for i in mem], codes[a:b], 3))
    else:
        out.append(u64(0))
    out += [boolean(trained), u64(len(rows)), u64(MAX_ITEM), nodes(rows, ids), u64(0), u32(mt),
            nodes(rows, ids)]
    return b"".join(out)
