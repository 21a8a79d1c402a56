// wave_heap.h -- the reference's std BinaryHeap (heap.rs:21-68, 95-112)
// run by one whole wave, with the wave's lanes loading many heap slots at
// once.  Entries are (distance bits, item) pairs; the storage policy S
// provides `uint2 ld(uint32_t)` and `void st(uint32_t, uint2)`.  Every lane
// of the wave must call push / pop together (the control flow is uniform).
//
//   push: the new element climbs while it is greater than its parent.  All
//     ancestors are loaded at once (one lane each); their values rise toward
//     the root, so the number of levels climbed is the count of leading
//     lanes whose ancestor is smaller, and the moves are one parallel store.
//   pop: the last element replaces the root; the hole descends to the bottom
//     along the larger child (the right one when left <= right), and the
//     element then climbs back while greater than its parent.  The descent
//     is resolved WH_R levels per round: lane l loads one sibling pair of
//     the WH_R-level subtree below the round's start and the path is walked
//     through register reads.  The moved values fall along the path from the
//     root down, so the climb stops below the first path value the element
//     does not exceed: only path positions above that point are written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hm {

constexpr uint32_t WH_R = 6;  // 2^6 - 1 = 63 sibling pairs per round

__device__ __forceinline__ float wh_val(uint2 v) { return __uint_as_float(v.x); }

__device__ __forceinline__ uint2 wh_lane(uint2 v, uint32_t l)
{
    return make_uint2((uint32_t)__builtin_amdgcn_readlane((int)v.x, (int)l),
                      (uint32_t)__builtin_amdgcn_readlane((int)v.y, (int)l));
}

template <class S> struct WaveHeapT {
    S s;
    uint32_t len;
    uint32_t lane;

    // heap.rs:21-35: append e at position len and sift it up
    __device__ void push(uint2 e)
    {
        const uint32_t p = len++;
        const uint32_t A = 31u - (uint32_t)__clz(p + 1);  // ancestors of p
        uint2 u = make_uint2(0u, 0u);
        bool mv = false;
        if (lane < A) {
            u = s.ld(((p + 1) >> (lane + 1)) - 1);
            mv = wh_val(e) > wh_val(u);
        }
        const uint64_t m = __ballot(mv);
        const uint32_t t = (uint32_t)__builtin_ctzll(~m);
        if (lane < t) s.st(lane == 0 ? p : ((p + 1) >> lane) - 1, u);
        if (lane == 0) s.st(t == 0 ? p : ((p + 1) >> t) - 1, e);
    }

    // heap.rs:54-68, 95-112: remove and return the root (len >= 1)
    __device__ uint2 pop()
    {
        len -= 1;
        const uint2 last = s.ld(len);
        if (len == 0) return last;
        const uint2 root = s.ld(0);
        const uint64_t end = len;
        uint32_t cur = 0, depth = 0, pp = 0;
        uint2 pw = make_uint2(0u, 0u);
        bool done = false;
        while (!done) {
            const uint64_t r0 = cur;
            uint2 a = make_uint2(0u, 0u), b = make_uint2(0u, 0u);
            if (lane < 63) {
                const uint32_t j = 32u - (uint32_t)__clz(lane + 1);
                const uint64_t o = lane + 1 - (1u << (j - 1));
                const uint64_t first = ((r0 + 1) << j) - 1 + 2 * o;
                if (first < end) a = s.ld((uint32_t)first);
                if (first + 1 < end) b = s.ld((uint32_t)first + 1);
            }
            for (uint32_t step = 1; step <= WH_R; ++step) {
                const uint64_t c1 = 2 * (uint64_t)cur + 1;
                if (c1 >= end) {
                    done = true;
                    break;
                }
                const uint32_t l =
                    (1u << (step - 1)) - 1 + (uint32_t)((c1 - (((r0 + 1) << step) - 1)) / 2);
                const uint2 av = wh_lane(a, l);
                uint32_t c;
                uint2 w;
                if (c1 + 1 < end) {
                    const uint2 bv = wh_lane(b, l);
                    if (wh_val(av) <= wh_val(bv)) {
                        c = (uint32_t)c1 + 1;
                        w = bv;
                    } else {
                        c = (uint32_t)c1;
                        w = av;
                    }
                } else {
                    c = (uint32_t)c1;
                    w = av;
                    done = true;
                }
                if (lane == depth) {
                    pp = cur;
                    pw = w;
                }
                ++depth;
                cur = c;
                if (done) break;
            }
        }
        const uint64_t m = __ballot(lane < depth && wh_val(last) > wh_val(pw));
        const uint32_t sd = depth - (uint32_t)__popcll(m);
        if (lane < sd) s.st(pp, pw);
        if (lane == 0)
            s.st(sd == depth ? cur : (uint32_t)__builtin_amdgcn_readlane((int)pp, (int)sd), last);
        return root;
    }
};

// std BinaryHeap on packed (item, distance bits) entries in LDS, position p
// at e[p + 1] (so a sibling pair is one aligned 16-byte word), run by one
// whole wave.
//   push: as WaveHeapT::push (all ancestors compared in one load round).
//   pop: the reference's sift_down_to_bottom followed by sift_up leaves the
//     same heap as a top-down sift of the last element that moves the larger
//     child up (the right one when left <= right) while that child is not
//     smaller than the element, and stops at a leaf: the children below the
//     stop point move up on the way down and back on the way up.  Each round
//     covers WH_R levels of that sift: lane l holds node l of the round's
//     63-node subtree (l's parent is (l - 1) / 2) and loads its sibling pair
//     of children.  Two ballots say where the sift stops and which way it
//     turns; lane l is on the sift's path when none of its ancestors stops
//     and each turns toward it (two mask compares against per-lane ancestor
//     and direction masks), so the path's moves are one parallel store and
//     the stopping lane stores the element.
struct PackedWaveHeap {
    uint2 *e;  // 16-byte aligned, capacity + 2 entries
    uint32_t len;
    uint32_t lane;
    uint32_t lj, lo;    // level of node `lane` in a round, its offset in the level
    uint64_t anc, dir;  // its ancestors, and those whose right child leads to it

    __device__ PackedWaveHeap(uint2 *e_, uint32_t lane_) : e(e_), len(0), lane(lane_)
    {
        lj = 31u - (uint32_t)__clz(lane + 1);
        lo = lane + 1 - (1u << lj);
        anc = 0;
        dir = 0;
        for (uint32_t x = lane < 63 ? lane : 0; x > 0;) {
            const uint32_t p = (x - 1) / 2;
            anc |= 1ull << p;
            if ((x & 1u) == 0) dir |= 1ull << p;
            x = p;
        }
    }

    __device__ static float val(uint2 x) { return __uint_as_float(x.y); }

    // heap.rs:21-35.  Lane l compares with ancestor l + 1; lanes past the
    // root read the unused slot e[0].
    __device__ void push(uint2 x)
    {
        const uint32_t p = len++;
        const uint32_t A = 31u - (uint32_t)__clz(p + 1);  // ancestors of p
        const uint2 u = e[(p + 1) >> min(lane + 1, 31u)];
        const uint64_t m = __ballot(lane < A && val(x) > val(u));
        const uint32_t t = (uint32_t)__builtin_ctzll(~m);  // levels climbed
        if (lane < t) e[(p + 1) >> lane] = u;               // ancestor lane + 1 moves down
        if (lane == 0) e[(p + 1) >> t] = x;
    }

    // heap.rs:54-68, 95-112: remove and return the root (len >= 1).  The
    // first round's loads go out with the reads of the root and the last
    // element; every lane loads (a node without children reads the pair at
    // the clamped index end & ~1, inside the array).
    __device__ uint2 pop()
    {
        len -= 1;
        const uint32_t end = len, ceven = end & ~1u;
        uint32_t mypos = (1u << lj) - 1 + lo;  // round root 0
        uint4 pr = *reinterpret_cast<const uint4 *>(&e[min(2 * mypos + 2, ceven)]);
        const uint2 last = e[end + 1];
        const uint2 root = e[1];
        if (end == 0) return last;
        const float lv = val(last);
        for (;;) {
            const bool ex = lane < 63 && 2 * mypos + 1 < end;
            const uint2 a = make_uint2(pr.x, pr.y), b = make_uint2(pr.z, pr.w);
            const bool right = ex && 2 * mypos + 2 < end && val(a) <= val(b);
            const uint2 w = right ? b : a;
            const bool st = !ex || lv > val(w);  // the sift stops at this node
            const uint64_t ms = __ballot(st), mr = __ballot(right);
            const bool onp = lane < 63 && (ms & anc) == 0 && (mr & anc) == dir;
            if (onp && !st) e[mypos + 1] = w;
            if (__ballot(onp && st)) {
                if (onp && st) e[mypos + 1] = last;
                break;
            }
            // no stop in this round: continue below the path's level-5 node
            const uint32_t l5 = (uint32_t)__builtin_ctzll(__ballot(onp && lj == WH_R - 1));
            const uint32_t child = 2 * mypos + 1 + (right ? 1u : 0u);
            const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)child, (int)l5);
            mypos = ((r0 + 1) << lj) - 1 + lo;
            pr = *reinterpret_cast<const uint4 *>(&e[min(2 * mypos + 2, ceven)]);
        }
        return root;
    }
};

// Storage: separate distance / item arrays (LDS or global).
struct WhSplit {
    float *d;
    uint32_t *i;
    __device__ uint2 ld(uint32_t p) const { return make_uint2(__float_as_uint(d[p]), i[p]); }
    __device__ void st(uint32_t p, uint2 v) const
    {
        d[p] = __uint_as_float(v.x);
        i[p] = v.y;
    }
};

}  // namespace hm
