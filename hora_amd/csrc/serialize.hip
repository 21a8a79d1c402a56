// serialize.hip -- SerializableIndex::dump / load (ann_index.rs:141-158) in
// the reference's on-disk layout.
//
// The reference derives serde's Serialize for its index structs and writes
// them with bincode::serialize (bincode 1.3, Cargo.toml:47), whose default
// configuration is: little-endian, integers at their fixed width (usize as
// u64), sequence and string lengths as u64, enum variants as u32, bool and
// Option tags as one byte.  The structs, in field order:
//   BruteForceIndex (bruteforce_idx.rs:15-22; `nodes` is skipped):
//     tmp_nodes: Vec<Node>, mt: Metric, dimension: usize
//   PQIndex (pq_idx.rs:21-45):
//     _dimension, _n_sub, _sub_dimension: usize, _dimension_range:
//     Vec<Vec<usize>>, _sub_bits, _sub_bytes, _n_sub_center, _code_bytes,
//     _train_epoch: usize, _centers: Vec<Vec<Vec<f32>>>, _is_trained,
//     _has_residual: bool, _residual: Vec<f32>, _n_items, _max_item: usize,
//     _nodes: Vec<Node>, _assigned_center: Vec<Vec<usize>>, mt: Metric,
//     _nodes_tmp: Vec<Node>
//   IVFPQIndex (pq_idx.rs:267-292):
//     _dimension, _n_sub, _sub_dimension, _sub_bits, _sub_bytes,
//     _n_sub_center, _code_bytes, _train_epoch, _search_n_center,
//     _n_kmeans_center: usize, _centers: Vec<Vec<f32>>, _ivflist:
//     Vec<Vec<usize>>, _pq_list: Vec<PQIndex>, _is_trained: bool, _n_items,
//     _max_item: usize, _nodes: Vec<Node>, _assigned_center:
//     Vec<Vec<usize>>, mt: Metric, _nodes_tmp: Vec<Node>
//   Node (node.rs:93-97): vectors: Vec<f32>, idx: Option<T>
// dump fills _nodes_tmp from _nodes (and each list's PQIndex likewise,
// pq_idx.rs:507-515), so those files hold the rows more than once; load
// rebuilds the nodes from _nodes_tmp (pq_idx.rs:490-505).
//
// The library does not know the caller's id type T: dump writes each node's
// Option<T> from caller-supplied bincode bytes, and load hands the bytes
// back (a fixed-width T, or a length-prefixed one such as String).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "index.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "bincode layout assumes a little-endian host");

namespace hm {

namespace {

constexpr uint64_t MAX_ITEM = 100000;  // pq_idx.rs:66, 319

class Out {
  public:
    explicit Out(const char *path)
    {
        f_ = std::fopen(path, "wb");
        if (!f_) fail(HM_ERR_INVALID_ARG, std::string("unable to write file ") + path);
        buf_.reserve(BUF);
    }
    ~Out()
    {
        if (f_) std::fclose(f_);
    }
    void raw(const void *p, size_t n)
    {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        while (n) {
            const size_t take = std::min(n, BUF - buf_.size());
            buf_.insert(buf_.end(), b, b + take);
            b += take;
            n -= take;
            if (buf_.size() == BUF) flush();
        }
    }
    void u8(uint8_t v) { raw(&v, 1); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void f32s(const float *p, size_t n) { raw(p, n * 4); }
    void close()
    {
        flush();
        const int rc = std::fclose(f_);
        f_ = nullptr;
        if (rc != 0) fail(HM_ERR_INVALID_ARG, "unable to write file");
    }

  private:
    void flush()
    {
        if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
            fail(HM_ERR_INVALID_ARG, "unable to write file");
        buf_.clear();
    }
    static constexpr size_t BUF = 8u << 20;
    std::FILE *f_ = nullptr;
    std::vector<uint8_t> buf_;
};

class In {
  public:
    explicit In(const char *path)
    {
        f_ = std::fopen(path, "rb");
        if (!f_) fail(HM_ERR_INVALID_ARG, std::string("unable to open file ") + path);
        std::setvbuf(f_, nullptr, _IOFBF, 8u << 20);
        if (std::fseek(f_, 0, SEEK_END) == 0) {
            const long end = std::ftell(f_);
            size_ = end > 0 ? (uint64_t)end : 0;
        }
        std::fseek(f_, 0, SEEK_SET);
    }
    ~In()
    {
        if (f_) std::fclose(f_);
    }
    void raw(void *p, size_t n)
    {
        if (n && std::fread(p, 1, n, f_) != n) fail(HM_ERR_INVALID_ARG, "bincode: unexpected end of file");
        pos_ += n;
    }
    uint8_t u8()
    {
        uint8_t v;
        raw(&v, 1);
        return v;
    }
    uint32_t u32()
    {
        uint32_t v;
        raw(&v, 4);
        return v;
    }
    uint64_t u64()
    {
        uint64_t v;
        raw(&v, 8);
        return v;
    }
    bool boolean()
    {
        const uint8_t v = u8();
        if (v > 1) fail(HM_ERR_INVALID_ARG, "bincode: invalid bool");
        return v != 0;
    }
    // a length that must not exceed what the rest of the file can hold
    uint64_t len(uint64_t elem_bytes)
    {
        const uint64_t n = u64();
        if (elem_bytes && n > remaining() / elem_bytes)
            fail(HM_ERR_INVALID_ARG, "bincode: sequence longer than the file");
        return n;
    }
    int32_t metric()
    {
        const uint32_t v = u32();  // metrics.rs:5-13 variant index
        if (v > HM_METRIC_ANGULAR) fail(HM_ERR_INVALID_ARG, "bincode: invalid Metric variant");
        return (int32_t)v;
    }
    bool at_end() const { return pos_ >= size_; }

  private:
    uint64_t remaining() const { return size_ > pos_ ? size_ - pos_ : 0; }
    std::FILE *f_ = nullptr;
    uint64_t size_ = 0, pos_ = 0;
};

// Option<T> of node i: the caller's bytes, or Some(i as usize) when none
// were given.
struct IdSource {
    const uint8_t *bytes;
    const uint64_t *off;
    void write(Out &o, uint64_t i) const
    {
        if (!bytes) {
            o.u8(1);
            o.u64(i);
        } else {
            o.raw(bytes + off[i], off[i + 1] - off[i]);
        }
    }
};

// The ids read back: Option tag + payload per node, in node order.
struct IdSink {
    int32_t width;  // > 0: fixed-width T; 0: u64-length-prefixed T
    std::vector<uint8_t> *bytes;
    std::vector<uint64_t> *off;
    void read(In &in)
    {
        const uint8_t tag = in.u8();
        if (tag > 1) fail(HM_ERR_INVALID_ARG, "bincode: invalid Option tag");
        std::vector<uint8_t> tmp;
        tmp.push_back(tag);
        if (tag == 1) {
            if (width > 0) {
                tmp.resize(1 + (size_t)width);
                in.raw(tmp.data() + 1, (size_t)width);
            } else {
                const uint64_t n = in.len(1);
                tmp.resize(9 + n);
                std::memcpy(tmp.data() + 1, &n, 8);
                in.raw(tmp.data() + 9, n);
            }
        }
        if (bytes) {
            bytes->insert(bytes->end(), tmp.begin(), tmp.end());
            off->push_back(bytes->size());
        }
    }
};

// Vec<Node>: all rows of the index in order (rows == nullptr, read from
// the device in chunks), or the given rows out of `all` (every row of the
// index on the host).
void write_nodes(Out &o, const Index *ix, const uint64_t *rows, uint64_t count,
                 const IdSource &ids, const float *all = nullptr)
{
    const uint64_t d = ix->dimension();
    o.u64(count);
    std::vector<float> buf;
    if (rows) {
        for (uint64_t j = 0; j < count; ++j) {
            o.u64(d);
            o.f32s(all + rows[j] * d, d);
            ids.write(o, rows[j]);
        }
    } else {
        const uint64_t chunk = std::max<uint64_t>(1, (64ull << 20) / (d * 4));
        buf.resize(std::min(count, chunk) * d);
        for (uint64_t r0 = 0; r0 < count; r0 += chunk) {
            const uint64_t c = std::min(chunk, count - r0);
            ix->copy_rows(r0, c, buf.data());
            for (uint64_t r = 0; r < c; ++r) {
                o.u64(d);
                o.f32s(buf.data() + r * d, d);
                ids.write(o, r0 + r);
            }
        }
    }
}

// Vec<Node>: rows appended to `rows` (each of dim floats; dim fixed by the
// first node when 0), ids to `sink` when given.
uint64_t read_nodes(In &in, uint64_t &dim, std::vector<float> *rows, IdSink sink)
{
    const uint64_t n = in.len(9);
    std::vector<float> skip;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t len = in.len(4);
        if (dim == 0) dim = len;
        if (len != dim) fail(HM_ERR_DIM_MISMATCH, "dimension is different");
        if (rows) {
            const size_t at = rows->size();
            rows->resize(at + len);
            in.raw(rows->data() + at, len * 4);
        } else {
            skip.resize(len);
            in.raw(skip.data(), len * 4);
        }
        sink.read(in);
    }
    return n;
}

void sub_ranges(uint64_t dim, uint32_t M, std::vector<uint32_t> &b, std::vector<uint32_t> &e)
{
    // pq_idx.rs:75-88
    const uint32_t sd = (uint32_t)(dim / M), r = (uint32_t)(dim % sd);
    b.resize(M);
    e.resize(M);
    for (uint32_t i = 0; i < M; ++i) {
        if (i < r) {
            b[i] = i * (sd + 1);
            e[i] = (i + 1) * (sd + 1);
        } else {
            b[i] = r * (sd + 1) + (i - r) * sd;
            e[i] = r * (sd + 1) + (i + 1 - r) * sd;
        }
    }
}

// The usize header fields shared by PQIndex and IVFPQIndex.
struct PQShape {
    uint64_t dim, n_sub, sub_dim, sub_bits, sub_bytes, n_center, code_bytes, epochs;
    static PQShape of(uint64_t dim, const hm_params &p)
    {
        PQShape s;
        s.dim = dim;
        s.n_sub = p.n_sub;
        s.sub_dim = dim / p.n_sub;
        s.sub_bits = p.sub_bits;
        s.sub_bytes = (p.sub_bits + 7) / 8;
        s.n_center = 1ull << p.sub_bits;
        s.code_bytes = s.sub_bytes * p.n_sub;
        s.epochs = p.train_epoch;
        return s;
    }
};

void write_ranges(Out &o, const std::vector<uint32_t> &b, const std::vector<uint32_t> &e)
{
    o.u64(b.size());
    for (size_t i = 0; i < b.size(); ++i) {
        o.u64(2);
        o.u64(b[i]);
        o.u64(e[i]);
    }
}

// _centers: Vec<Vec<Vec<f32>>> from a codebook [M][KS][sdm] (sub s keeps its
// own e-b dims).
void write_codebook(Out &o, const float *cb, uint32_t M, uint32_t KS, uint32_t sdm,
                    const std::vector<uint32_t> &b, const std::vector<uint32_t> &e)
{
    o.u64(M);
    for (uint32_t s = 0; s < M; ++s) {
        o.u64(KS);
        for (uint32_t c = 0; c < KS; ++c) {
            o.u64(e[s] - b[s]);
            o.f32s(cb + ((size_t)s * KS + c) * sdm, e[s] - b[s]);
        }
    }
}

// _assigned_center: Vec<Vec<usize>> [M][count] from codes [count][M].
void write_codes(Out &o, const uint8_t *codes, uint64_t count, uint32_t M)
{
    o.u64(M);
    for (uint32_t s = 0; s < M; ++s) {
        o.u64(count);
        for (uint64_t i = 0; i < count; ++i) o.u64(codes[i * M + s]);
    }
}

void read_codebook(In &in, float *cb, uint32_t M, uint32_t KS, uint32_t sdm,
                   const std::vector<uint32_t> &b, const std::vector<uint32_t> &e)
{
    if (in.len(8) != M) fail(HM_ERR_INVALID_ARG, "bincode: _centers has the wrong sub-space count");
    for (uint32_t s = 0; s < M; ++s) {
        if (in.len(8) != KS) fail(HM_ERR_INVALID_ARG, "bincode: _centers has the wrong centre count");
        for (uint32_t c = 0; c < KS; ++c) {
            if (in.len(4) != e[s] - b[s]) fail(HM_ERR_INVALID_ARG, "bincode: codeword length");
            float *dst = cb + ((size_t)s * KS + c) * sdm;
            std::memset(dst, 0, sdm * 4);
            in.raw(dst, (size_t)(e[s] - b[s]) * 4);
        }
    }
}

void read_codes(In &in, uint8_t *codes, uint64_t count, uint32_t M, uint32_t KS)
{
    if (in.len(8) != M) fail(HM_ERR_INVALID_ARG, "bincode: _assigned_center sub-space count");
    for (uint32_t s = 0; s < M; ++s) {
        if (in.len(8) != count) fail(HM_ERR_INVALID_ARG, "bincode: _assigned_center length");
        for (uint64_t i = 0; i < count; ++i) {
            const uint64_t c = in.u64();
            if (c >= KS) fail(HM_ERR_INVALID_ARG, "bincode: code out of range");
            codes[i * M + s] = (uint8_t)c;
        }
    }
}

// ---------------------------------------------------------------- dump

void dump_bruteforce(const Index *ix, Out &o, const IdSource &ids)
{
    write_nodes(o, ix, nullptr, ix->nodes_size(), ids);  // tmp_nodes
    o.u32((uint32_t)ix->metric());                       // mt
    o.u64(ix->dimension());                              // dimension
}

void dump_pq(Index *ix, Out &o, const IdSource &ids)
{
    const hm_params &p = pq_params(ix);
    const PQShape sh = PQShape::of(ix->dimension(), p);
    uint32_t info[8];
    pq_info(ix, info);
    const uint32_t M = info[0], KS = info[1], sdm = info[2];
    const uint64_t nt = info[3], n = ix->nodes_size();
    const bool trained = info[4] != 0;
    std::vector<uint32_t> b, e;
    sub_ranges(sh.dim, M, b, e);
    o.u64(sh.dim);
    o.u64(sh.n_sub);
    o.u64(sh.sub_dim);
    write_ranges(o, b, e);
    o.u64(sh.sub_bits);
    o.u64(sh.sub_bytes);
    o.u64(sh.n_center);
    o.u64(sh.code_bytes);
    o.u64(sh.epochs);
    std::vector<float> cb;
    std::vector<uint8_t> codes;
    if (trained) {
        cb.resize((size_t)M * KS * sdm);
        codes.resize(nt * M + 1);
        pq_export(ix, cb.data(), codes.data());
        write_codebook(o, cb.data(), M, KS, sdm, b, e);
    } else {
        o.u64(0);
    }
    o.u8(trained ? 1 : 0);  // _is_trained
    o.u8(0);                // _has_residual
    o.u64(0);               // _residual
    o.u64(n);               // _n_items
    o.u64(MAX_ITEM);        // _max_item
    write_nodes(o, ix, nullptr, n, ids);  // _nodes
    if (trained) write_codes(o, codes.data(), nt, M);
    else o.u64(0);
    // PQIndex::new sets Euclidean (pq_idx.rs:71); build overwrites it
    o.u32(trained ? (uint32_t)ix->metric() : (uint32_t)HM_METRIC_EUCLIDEAN);
    write_nodes(o, ix, nullptr, n, ids);  // _nodes_tmp
}

void dump_ivfpq(Index *ix, Out &o, const IdSource &ids)
{
    const hm_params &p = ivfpq_params(ix);
    const PQShape sh = PQShape::of(ix->dimension(), p);
    uint32_t info[8];
    ivfpq_info(ix, info);
    const uint32_t nlist = info[0], M = info[1], KS = info[2], sdm = info[3];
    const uint64_t nt = info[4], n = ix->nodes_size(), d = sh.dim;
    const bool trained = info[5] != 0;
    std::vector<uint32_t> b, e;
    sub_ranges(d, M, b, e);
    o.u64(sh.dim);
    o.u64(sh.n_sub);
    o.u64(sh.sub_dim);
    o.u64(sh.sub_bits);
    o.u64(sh.sub_bytes);
    o.u64(sh.n_center);
    o.u64(sh.code_bytes);
    o.u64(sh.epochs);
    o.u64(info[6]);  // _search_n_center
    o.u64(nlist);    // _n_kmeans_center
    std::vector<float> cen, cb, all;
    std::vector<uint64_t> off, lid;
    std::vector<uint8_t> codes;
    if (trained) {
        all.resize(n * d);
        ix->copy_rows(0, n, all.data());
        cen.resize((size_t)nlist * d);
        off.resize(nlist + 1);
        lid.resize(nt + 1);
        cb.resize((size_t)nlist * M * KS * sdm);
        codes.resize(nt * M + 1);
        ivfpq_export(ix, cen.data(), off.data(), lid.data(), cb.data(), codes.data());
        o.u64(nlist);  // _centers
        for (uint32_t l = 0; l < nlist; ++l) {
            o.u64(d);
            o.f32s(cen.data() + (size_t)l * d, d);
        }
    } else {
        o.u64(0);
    }
    o.u64(nlist);  // _ivflist (IVFPQIndex::new makes nlist empty lists)
    for (uint32_t l = 0; l < nlist; ++l) {
        const uint64_t L = trained ? off[l + 1] - off[l] : 0;
        o.u64(L);
        for (uint64_t r = 0; r < L; ++r) o.u64(lid[off[l] + r]);
    }
    if (trained) {
        o.u64(nlist);  // _pq_list: one PQIndex per list (pq_idx.rs:377-393)
        for (uint32_t l = 0; l < nlist; ++l) {
            const uint64_t L = off[l + 1] - off[l];
            o.u64(sh.dim);
            o.u64(sh.n_sub);
            o.u64(sh.sub_dim);
            write_ranges(o, b, e);
            o.u64(sh.sub_bits);
            o.u64(sh.sub_bytes);
            o.u64(sh.n_center);
            o.u64(sh.code_bytes);
            o.u64(sh.epochs);
            write_codebook(o, cb.data() + (size_t)l * M * KS * sdm, M, KS, sdm, b, e);
            o.u8(1);  // _is_trained
            o.u8(1);  // _has_residual
            o.u64(d);  // _residual = the list's centre
            o.f32s(cen.data() + (size_t)l * d, d);
            o.u64(L);  // _n_items
            o.u64(MAX_ITEM);
            write_nodes(o, ix, lid.data() + off[l], L, ids, all.data());  // _nodes
            write_codes(o, codes.data() + off[l] * M, L, M);
            o.u32(HM_METRIC_EUCLIDEAN);  // the lists' PQ keeps Euclidean (pq_idx.rs:71)
            write_nodes(o, ix, lid.data() + off[l], L, ids, all.data());  // _nodes_tmp
        }
    } else {
        o.u64(0);
    }
    o.u8(trained ? 1 : 0);  // _is_trained
    o.u64(n);               // _n_items
    o.u64(MAX_ITEM);        // _max_item
    write_nodes(o, ix, nullptr, n, ids);  // _nodes
    o.u64(0);                             // _assigned_center (never filled)
    o.u32((uint32_t)ix->metric());
    write_nodes(o, ix, nullptr, n, ids);  // _nodes_tmp
}

// ---------------------------------------------------------------- load

void add_rows(Index *ix, const std::vector<float> &rows)
{
    const uint64_t d = ix->dimension(), n = d ? rows.size() / d : 0;
    const uint64_t chunk = std::max<uint64_t>(1, (256ull << 20) / (d * 4));
    for (uint64_t r0 = 0; r0 < n; r0 += chunk)
        ix->add(rows.data() + r0 * d, std::min(chunk, n - r0));
}

std::unique_ptr<Index> load_bruteforce(In &in, IdSink sink)
{
    uint64_t dim = 0;
    std::vector<float> rows;
    read_nodes(in, dim, &rows, sink);  // tmp_nodes
    const int32_t mt = in.metric();
    const uint64_t d = in.u64();
    if (dim && dim != d) fail(HM_ERR_DIM_MISMATCH, "dimension is different");
    auto ix = make_bruteforce(d);
    add_rows(ix.get(), rows);
    if (mt != HM_METRIC_UNKNOWN) ix->build(mt);
    return ix;
}

// Reads the PQIndex header fields into params; checks them against each
// other.
void read_pq_header(In &in, hm_params &p, uint64_t &dim, std::vector<uint32_t> &b,
                    std::vector<uint32_t> &e)
{
    dim = in.u64();
    const uint64_t n_sub = in.u64(), sub_dim = in.u64();
    if (n_sub == 0 || n_sub > 256 || dim == 0 || dim > 4096 || sub_dim != dim / n_sub || sub_dim == 0)
        fail(HM_ERR_INVALID_ARG, "bincode: inconsistent PQ header");
    sub_ranges(dim, (uint32_t)n_sub, b, e);
    if (in.len(8) != n_sub) fail(HM_ERR_INVALID_ARG, "bincode: _dimension_range length");
    for (uint64_t i = 0; i < n_sub; ++i) {
        if (in.len(8) != 2 || in.u64() != b[i] || in.u64() != e[i])
            fail(HM_ERR_INVALID_ARG, "bincode: _dimension_range differs from pq_idx.rs:75-88");
    }
    const uint64_t sub_bits = in.u64(), sub_bytes = in.u64(), n_center = in.u64();
    const uint64_t code_bytes = in.u64(), epochs = in.u64();
    if (sub_bits == 0 || sub_bits > 32 || n_center != (1ull << sub_bits) ||
        sub_bytes != (sub_bits + 7) / 8 || code_bytes != sub_bytes * n_sub)
        fail(HM_ERR_INVALID_ARG, "bincode: inconsistent PQ header");
    p.n_sub = (uint32_t)n_sub;
    p.sub_bits = (uint32_t)sub_bits;
    p.train_epoch = (uint32_t)std::min<uint64_t>(epochs, 0xFFFFFFFFu);
}

std::unique_ptr<Index> load_pq(In &in, IdSink sink)
{
    hm_params p;
    hm_params_default(HM_INDEX_PQ, &p);
    uint64_t dim = 0;
    std::vector<uint32_t> b, e;
    read_pq_header(in, p, dim, b, e);
    auto ix = make_pq(dim, p);  // checks sub_bits <= 8
    uint32_t info[8];
    pq_info(ix.get(), info);
    const uint32_t M = info[0], KS = info[1], sdm = info[2];
    std::vector<float> cb((size_t)M * KS * sdm, 0.0f);
    const uint64_t nc = in.len(8);
    if (nc != 0 && nc != M) fail(HM_ERR_INVALID_ARG, "bincode: _centers has the wrong sub-space count");
    if (nc) {
        for (uint32_t s = 0; s < M; ++s) {
            if (in.len(8) != KS) fail(HM_ERR_INVALID_ARG, "bincode: _centers centre count");
            for (uint32_t c = 0; c < KS; ++c) {
                if (in.len(4) != e[s] - b[s]) fail(HM_ERR_INVALID_ARG, "bincode: codeword length");
                in.raw(cb.data() + ((size_t)s * KS + c) * sdm, (size_t)(e[s] - b[s]) * 4);
            }
        }
    }
    const bool trained = in.boolean();
    if (in.boolean()) fail(HM_ERR_NOT_IMPLEMENTED, "a standalone PQIndex with a residual");
    if (in.len(4) != 0) fail(HM_ERR_NOT_IMPLEMENTED, "a standalone PQIndex with a residual");
    const uint64_t n_items = in.u64();
    (void)in.u64();  // _max_item
    uint64_t d2 = dim;
    read_nodes(in, d2, nullptr, IdSink{sink.width, nullptr, nullptr});  // _nodes
    const uint64_t na = in.len(8);
    std::vector<uint8_t> codes;
    uint64_t n_coded = 0;
    if (na) {
        if (na != M) fail(HM_ERR_INVALID_ARG, "bincode: _assigned_center sub-space count");
        for (uint32_t s = 0; s < M; ++s) {
            const uint64_t cnt = in.len(8);
            if (s == 0) {
                n_coded = cnt;
                codes.assign(cnt * M + 1, 0);
            } else if (cnt != n_coded) {
                fail(HM_ERR_INVALID_ARG, "bincode: _assigned_center lengths differ");
            }
            for (uint64_t i = 0; i < cnt; ++i) {
                const uint64_t c = in.u64();
                if (c >= KS) fail(HM_ERR_INVALID_ARG, "bincode: code out of range");
                codes[i * M + s] = (uint8_t)c;
            }
        }
    }
    const int32_t mt = in.metric();
    std::vector<float> rows;
    const uint64_t n = read_nodes(in, d2, &rows, sink);  // _nodes_tmp
    if (n != n_items) fail(HM_ERR_INVALID_ARG, "bincode: _n_items differs from the nodes");
    add_rows(ix.get(), rows);
    if (trained) {
        if (nc == 0 || n_coded > n) fail(HM_ERR_INVALID_ARG, "bincode: trained PQIndex without codes");
        pq_import(ix.get(), mt, cb.data(), codes.data(), n_coded);
    }
    return ix;
}

std::unique_ptr<Index> load_ivfpq(In &in, IdSink sink)
{
    hm_params p;
    hm_params_default(HM_INDEX_IVFPQ, &p);
    const uint64_t dim = in.u64(), n_sub = in.u64(), sub_dim = in.u64(), sub_bits = in.u64();
    const uint64_t sub_bytes = in.u64(), n_center = in.u64(), code_bytes = in.u64();
    const uint64_t epochs = in.u64(), nprobe = in.u64(), nlist = in.u64();
    if (n_sub == 0 || n_sub > 256 || dim == 0 || dim > 4096 || sub_dim != dim / n_sub ||
        sub_dim == 0 || sub_bits == 0 || sub_bits > 32 || n_center != (1ull << sub_bits) ||
        sub_bytes != (sub_bits + 7) / 8 || code_bytes != sub_bytes * n_sub || nlist == 0 ||
        nlist > 0xFFFFFFFull)
        fail(HM_ERR_INVALID_ARG, "bincode: inconsistent IVFPQ header");
    p.n_sub = (uint32_t)n_sub;
    p.sub_bits = (uint32_t)sub_bits;
    p.n_kmeans_center = (uint32_t)nlist;
    p.search_n_center = (uint32_t)std::min<uint64_t>(nprobe, 0xFFFFFFFFu);
    p.train_epoch = (uint32_t)std::min<uint64_t>(epochs, 0xFFFFFFFFu);
    auto ix = make_ivfpq(dim, p);
    uint32_t info[8];
    ivfpq_info(ix.get(), info);
    const uint32_t M = info[1], KS = info[2], sdm = info[3];
    std::vector<uint32_t> b, e;
    sub_ranges(dim, M, b, e);
    const uint64_t ncen = in.len(8);
    if (ncen != 0 && ncen != nlist) fail(HM_ERR_INVALID_ARG, "bincode: _centers count");
    std::vector<float> cen(ncen * dim);
    for (uint64_t l = 0; l < ncen; ++l) {
        if (in.len(4) != dim) fail(HM_ERR_DIM_MISMATCH, "dimension is different");
        in.raw(cen.data() + l * dim, dim * 4);
    }
    if (in.len(8) != nlist) fail(HM_ERR_INVALID_ARG, "bincode: _ivflist count");
    std::vector<uint64_t> off(nlist + 1, 0), lid;
    for (uint64_t l = 0; l < nlist; ++l) {
        const uint64_t L = in.len(8);
        for (uint64_t r = 0; r < L; ++r) lid.push_back(in.u64());
        off[l + 1] = off[l] + L;
    }
    const uint64_t npq = in.len(8);
    if (npq != 0 && npq != nlist) fail(HM_ERR_INVALID_ARG, "bincode: _pq_list count");
    std::vector<float> cb(npq ? (size_t)nlist * M * KS * sdm : 0, 0.0f);
    std::vector<uint8_t> codes(npq ? off[nlist] * M + 1 : 0);
    for (uint64_t l = 0; l < npq; ++l) {
        hm_params q = p;
        uint64_t d2 = 0;
        std::vector<uint32_t> b2, e2;
        read_pq_header(in, q, d2, b2, e2);
        if (d2 != dim || q.n_sub != M || (1u << q.sub_bits) != KS)
            fail(HM_ERR_INVALID_ARG, "bincode: a list's PQIndex differs from the IVF header");
        read_codebook(in, cb.data() + (size_t)l * M * KS * sdm, M, KS, sdm, b, e);
        (void)in.boolean();  // _is_trained
        (void)in.boolean();  // _has_residual
        const uint64_t rl = in.len(4);
        std::vector<float> skip(rl);
        in.raw(skip.data(), rl * 4);  // _residual (the list's centre)
        const uint64_t L = in.u64();
        if (L != off[l + 1] - off[l]) fail(HM_ERR_INVALID_ARG, "bincode: list length");
        (void)in.u64();  // _max_item
        uint64_t d3 = dim;
        read_nodes(in, d3, nullptr, IdSink{sink.width, nullptr, nullptr});
        read_codes(in, codes.data() + off[l] * M, L, M, KS);
        (void)in.metric();
        read_nodes(in, d3, nullptr, IdSink{sink.width, nullptr, nullptr});
    }
    const bool trained = in.boolean();
    const uint64_t n_items = in.u64();
    (void)in.u64();  // _max_item
    uint64_t d4 = dim;
    read_nodes(in, d4, nullptr, IdSink{sink.width, nullptr, nullptr});  // _nodes
    const uint64_t nac = in.len(8);
    for (uint64_t i = 0; i < nac; ++i) {
        const uint64_t L = in.len(8);
        for (uint64_t r = 0; r < L; ++r) (void)in.u64();
    }
    const int32_t mt = in.metric();
    std::vector<float> rows;
    const uint64_t n = read_nodes(in, d4, &rows, sink);  // _nodes_tmp
    if (n != n_items) fail(HM_ERR_INVALID_ARG, "bincode: _n_items differs from the nodes");
    add_rows(ix.get(), rows);
    if (trained) {
        if (ncen != nlist || npq != nlist) fail(HM_ERR_INVALID_ARG, "bincode: trained IVFPQIndex without lists");
        ivfpq_import(ix.get(), mt, cen.data(), off.data(), lid.data(), cb.data(), codes.data());
    }
    return ix;
}

}  // namespace

void dump_index(Index *ix, const char *path, const uint8_t *id_bytes, const uint64_t *id_off)
{
    Out o(path);
    const IdSource ids{id_bytes, id_off};
    switch (ix->kind()) {
    case HM_INDEX_BRUTE_FORCE: dump_bruteforce(ix, o, ids); break;
    case HM_INDEX_PQ: dump_pq(ix, o, ids); break;
    case HM_INDEX_IVFPQ: dump_ivfpq(ix, o, ids); break;
    default: fail(HM_ERR_INVALID_ARG, "unknown index kind");
    }
    o.close();
}

std::unique_ptr<Index> load_index(int32_t kind, const char *path, int32_t id_width,
                                  std::vector<uint8_t> &id_bytes, std::vector<uint64_t> &id_off)
{
    if (id_width < 0 || id_width > 16) fail(HM_ERR_INVALID_ARG, "id width must be 0..16");
    In in(path);
    id_bytes.clear();
    id_off.assign(1, 0);
    const IdSink sink{id_width, &id_bytes, &id_off};
    std::unique_ptr<Index> ix;
    switch (kind) {
    case HM_INDEX_BRUTE_FORCE: ix = load_bruteforce(in, sink); break;
    case HM_INDEX_PQ: ix = load_pq(in, sink); break;
    case HM_INDEX_IVFPQ: ix = load_ivfpq(in, sink); break;
    default: fail(HM_ERR_INVALID_ARG, "unknown index kind");
    }
    if (!in.at_end()) fail(HM_ERR_INVALID_ARG, "bincode: trailing bytes after the index");
    return ix;
}

}  // namespace hm
