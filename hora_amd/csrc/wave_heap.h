// wave_heap.h -- the reference's std BinaryHeap (heap.rs:21-68, 95-112)
// run by one whole wave, with the wave's lanes loading many heap slots at
// once.  Entries are (item, distance bits) pairs; the storage policy S
// provides `uint2 ld(uint32_t p)`, `void st(uint32_t p, uint2 v)` for heap
// position p, and `uint4 pair(uint32_t first, uint32_t end)`: the entries at
// positions first and first + 1, from any in-bounds location when first >=
// end (those lanes' values are not used).  Every lane of the wave must call
// push / pop together (the control flow is uniform).
//
//   push: the new element climbs while it is greater than its parent.  All
//     ancestors are loaded at once (one lane each); their values rise toward
//     the root, so the number of levels climbed is the count of leading
//     lanes whose ancestor is smaller, and the moves are one parallel store.
//   pop: the last element replaces the root; sift_down_to_bottom moves the
//     hole to the bottom along the larger child (the right one when left <=
//     right) and sift_up lets the element climb back while greater than its
//     parent.  That leaves the same heap as a top-down sift of the element
//     that moves the chosen child up while the child is not smaller than the
//     element and stops at a leaf: below the stop point, the children that
//     moved up on the way down move back on the way up.  Each round covers
//     WH_R levels of that sift: lane l holds node l of the round's 63-node
//     subtree (l's parent is (l - 1) / 2) and loads its pair of children.
//     Two ballots say where the sift stops and which way it turns; lane l is
//     on the sift's path when none of its ancestors stops and each turns
//     toward it (two compares against per-lane ancestor and direction
//     masks), so the path's moves are one parallel store and the stopping
//     lane stores the element.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hm {

constexpr uint32_t WH_R = 6;  // 2^6 - 1 = 63 nodes per pop round

__device__ __forceinline__ uint2 wh_lane(uint2 v, uint32_t l)
{
    return make_uint2((uint32_t)__builtin_amdgcn_readlane((int)v.x, (int)l),
                      (uint32_t)__builtin_amdgcn_readlane((int)v.y, (int)l));
}

template <class S> struct WaveHeap {
    S s;
    uint32_t len;
    uint32_t lane;
    uint32_t lj, lo;    // level of node `lane` in a round, its offset in the level
    uint64_t anc, dir;  // its ancestors, and those whose right child leads to it

    __device__ WaveHeap(S s_, uint32_t lane_) : s(s_), len(0), lane(lane_)
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

    // heap.rs:21-35: append x at position len and sift it up.  Lane l
    // compares with ancestor l + 1 (lanes past the root read the root).
    __device__ void push(uint2 x)
    {
        const uint32_t p = len++;
        const uint32_t A = 31u - (uint32_t)__clz(p + 1);  // ancestors of p
        const uint2 u = s.ld(lane < A ? ((p + 1) >> (lane + 1)) - 1 : 0u);
        // both words in one load: left to itself the compiler loads the item
        // under the move branch, a second LDS round trip
        asm volatile("" ::"v"(u.x), "v"(u.y));
        const uint64_t m = __ballot(lane < A && val(x) > val(u));
        const uint32_t t = (uint32_t)__builtin_ctzll(~m);  // levels climbed
        if (lane < t) s.st(((p + 1) >> lane) - 1, u);       // ancestor lane + 1 moves down
        if (lane == 0) s.st(((p + 1) >> t) - 1, x);
    }

    // heap.rs:54-68, 95-112: remove and return the root (len >= 1).  The
    // first round's loads go out with the reads of the root and the last
    // element.
    __device__ uint2 pop()
    {
        len -= 1;
        const uint32_t end = len;
        uint32_t mypos = (1u << lj) - 1 + lo;  // round root 0
        uint4 pr = s.pair(2 * mypos + 1, end);
        const uint2 last = s.ld(end);
        const uint2 root = s.ld(0);
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
            if (onp && !st) s.st(mypos, w);
            if (__ballot(onp && st)) {
                if (onp && st) s.st(mypos, last);
                break;
            }
            // no stop in this round: continue below the path's level-5 node
            const uint32_t l5 = (uint32_t)__builtin_ctzll(__ballot(onp && lj == WH_R - 1));
            const uint32_t child = 2 * mypos + 1 + (right ? 1u : 0u);
            const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)child, (int)l5);
            mypos = ((r0 + 1) << lj) - 1 + lo;
            pr = s.pair(2 * mypos + 1, end);
        }
        return root;
    }
};

// The same heap for at most 64 entries held in registers: lane p holds
// position p, and the loads of the LDS form become cross-lane reads.  push
// and pop make the same moves as WaveHeap's (a pop's sift ends inside one
// 63-node round, since every position is below 64).
struct WaveHeapReg {
    uint32_t it, dv;  // this lane's entry: item, distance bits
    uint32_t len;
    uint32_t lane;
    uint64_t anc, dir;  // as WaveHeap's, for node `lane` of the 63-node round

    __device__ explicit WaveHeapReg(uint32_t lane_) : it(0), dv(0), len(0), lane(lane_)
    {
        anc = 0;
        dir = 0;
        for (uint32_t x = lane < 63 ? lane : 0; x > 0;) {
            const uint32_t p = (x - 1) / 2;
            anc |= 1ull << p;
            if ((x & 1u) == 0) dir |= 1ull << p;
            x = p;
        }
    }

    __device__ static float fv(uint32_t b) { return __uint_as_float(b); }
    __device__ static uint32_t xl(uint32_t v, uint32_t src)  // v of lane src (per lane)
    {
        return (uint32_t)__shfl((int)v, (int)src, 64);
    }
    // entry / value at a wave-uniform position, value at a per-lane one
    __device__ uint2 at(uint32_t p) const { return wh_lane(make_uint2(it, dv), p); }
    __device__ float val_at(uint32_t p) const
    {
        return fv((uint32_t)__builtin_amdgcn_readlane((int)dv, (int)p));
    }
    __device__ float val_lane(uint32_t p) const { return fv(xl(dv, min(p, 63u))); }

    // heap.rs:21-35 (len < 64): lane l < A compares with ancestor l + 1; the
    // position j steps above p takes its parent's entry for j < t and x at t
    __device__ void push(uint2 x)
    {
        const uint32_t p = len++, P = p + 1;
        const uint32_t A = 31u - (uint32_t)__clz(P);
        const float u = fv(xl(dv, lane < A ? (P >> (lane + 1)) - 1 : 0u));
        const uint64_t m = __ballot(lane < A && fv(x.y) > u);
        const uint32_t t = (uint32_t)__builtin_ctzll(~m);
        const uint32_t Q = lane + 1, bq = 32u - (uint32_t)__clz(Q), bp = 32u - (uint32_t)__clz(P);
        const uint32_t j = bp - bq;
        const bool on = bq <= bp && (P >> j) == Q && j <= t;
        const uint32_t par = lane > 0 ? (lane - 1) / 2 : 0u;
        const uint32_t pit = xl(it, par), pdv = xl(dv, par);
        if (on) {
            it = j < t ? pit : x.x;
            dv = j < t ? pdv : x.y;
        }
    }

    // heap.rs:54-68, 95-112 (len >= 1): the top-down sift of the last entry
    __device__ uint2 pop()
    {
        len -= 1;
        const uint32_t end = len;
        const uint2 root = at(0), last = at(end);
        if (end == 0) return last;
        const float lv = fv(last.y);
        const uint32_t c1 = min(2 * lane + 1, 63u), c2 = min(2 * lane + 2, 63u);
        const uint32_t ait = xl(it, c1), adv = xl(dv, c1), bit = xl(it, c2), bdv = xl(dv, c2);
        const bool ex = lane < 63 && 2 * lane + 1 < end;
        const bool right = ex && 2 * lane + 2 < end && fv(adv) <= fv(bdv);
        const uint32_t wit = right ? bit : ait, wdv = right ? bdv : adv;
        const bool st = !ex || lv > fv(wdv);  // the sift stops at this node
        const uint64_t ms = __ballot(st), mr = __ballot(right);
        const bool onp = lane < 63 && (ms & anc) == 0 && (mr & anc) == dir;
        if (onp) {
            it = st ? last.x : wit;
            dv = st ? last.y : wdv;
        }
        return root;
    }
};

// Storage: the heap in LDS with position p at e[p + 1], so the children
// (2q + 1, 2q + 2) of a node are one aligned 16-byte word.  e: 16-byte
// aligned, capacity + 1 entries.
struct WhPackedLds {
    uint2 *e;
    __device__ uint2 ld(uint32_t p) const { return e[p + 1]; }
    __device__ void st(uint32_t p, uint2 v) const { e[p + 1] = v; }
    __device__ uint4 pair(uint32_t first, uint32_t end) const
    {
        return *reinterpret_cast<const uint4 *>(&e[min(first + 1, end & ~1u)]);
    }
};

}  // namespace hm
