// vp_walk.h — index arithmetic of the walk tree, shared by the device walks
// (vp_device.h), the tree build (vp_bvh.hip) and a host check of the
// traversal (tests/walk_tree_check.cpp).
//
// The walk tree is an implicit complete 4-ary tree over the stored
// primitives.  With L levels (the smallest L > kLeafBlockLevels with
// 4^L >= N), node k of
// level l has heap index base(l) + k, base(l) = (4^l - 1) / 3, and its
// children are nodes 4k .. 4k + 3 of level l + 1; below the last level
// (l = L - 1) the four children of node k are the leaves 4k .. 4k + 3, i.e.
// stored primitives (slots >= N are empty).  A node is one 128-byte record
// holding the boxes of its four children, one float4 per bound and axis:
//   r[0] = min.x of children 0..3, r[1] = min.y, r[2] = min.z,
//   r[3] = max.x,                  r[4] = max.y, r[5] = max.z, r[6..7] unused.
// An empty child has min = max = +inf on every axis, which no ray or point
// test accepts.  Parents and siblings follow from the index, so walks need no
// stack (see Walk4).
//
// A scene whose boxes lose little to it stores the tree instead in 64-byte
// records of binary16 bounds (the 16-bit records), each bound rounded
// outward (min toward -inf, max toward +inf) from the 32-bit record:
//   bytes  0.. 7 min.x of children 0..3, 8..15 min.y, 16..23 min.z,
//   bytes 24..31 max.x,                  32..39 max.y, 40..47 max.z,
//   bytes 48..63 unused.
// A decoded 16-bit box contains the 32-bit one, and since the rounding is
// monotone a decoded parent box still contains its decoded children, so a
// walk of the 16-bit tree meets every leaf a walk of the 32-bit tree meets
// (and perhaps a few more, which the leaf tests reject).  The device node
// tests decode in the FMA (binary16 operands of v_fma_mix_f32 convert
// exactly), so the slab distances are those of the decoded box.
#pragma once

#include "vp_math.h"

namespace vp {

constexpr uint32_t kW4RecordF4 = 8;    // float4 per node record
constexpr uint32_t kW4hRecordB = 64;   // bytes per 16-bit node record
constexpr uint32_t kW4MaxLevels = 15;  // 4^15 = 2^30 leaves

// ---- binary16 with directed rounding (16-bit records) ----

// Bits of the largest binary16 magnitude <= |x| (x not NaN); inexact tells
// whether it differs from |x|.  Magnitudes above the largest finite value
// 65504 give 65504 (inexact), +inf gives +inf.
VP_HD uint32_t f16_mag_floor(float x, bool& inexact) {
    const uint32_t u = f2u(x) & 0x7fffffffu;
    const float a = u2f(u);
    if (u >= 0x7f800000u) {
        inexact = false;
        return 0x7c00u;
    }
    if (a >= 65504.0f) {
        inexact = a != 65504.0f;
        return 0x7bffu;
    }
    if (a >= 0x1p-14f) {
        // normal: biased exponent 1..30 and the top 10 bits of the fraction
        inexact = (u & 0x1fffu) != 0u;
        return (((u >> 23) - 112u) << 10) | ((u >> 13) & 0x3ffu);
    }
    // subnormal: m 2^-24 with m < 1024 (the scaling is exact)
    const float s = a * 0x1p24f;
    const uint32_t m = static_cast<uint32_t>(s);
    inexact = static_cast<float>(m) != s;
    return m;
}
// largest binary16 <= x, as bits (x not NaN)
VP_HD uint16_t f16_down(float x) {
    bool inexact = false;
    const uint32_t m = f16_mag_floor(x, inexact);
    if (!(x < 0.0f)) return static_cast<uint16_t>(m);
    // negative: the magnitude rounds up (65504 + 1 ulp is the bit pattern of inf)
    return static_cast<uint16_t>(0x8000u | (m + (inexact ? 1u : 0u)));
}
// smallest binary16 >= x, as bits (x not NaN)
VP_HD uint16_t f16_up(float x) { return static_cast<uint16_t>(f16_down(-x) ^ 0x8000u); }
// the value of binary16 bits h (not NaN), exactly
VP_HD float f16_value(uint16_t h) {
    const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, f = h & 0x3ffu;
    if (e == 31u) return u2f(sign | 0x7f800000u);
    if (e == 0u) return u2f(sign | f2u(static_cast<float>(f) * 0x1p-24f));
    return u2f(sign | ((e + 112u) << 23) | (f << 13));
}

// Levels of the tree the ray walks leave out (walk4_block_levels): their
// leaves are handed to the leaf tests in blocks of 4^kLeafBlockLevels.
#ifndef VP_BLOCK_LEVELS
#define VP_BLOCK_LEVELS 1
#endif
constexpr uint32_t kLeafBlockLevels = VP_BLOCK_LEVELS;
constexpr uint32_t kLeafBlock = 1u << (2u * kLeafBlockLevels);  // leaves per block
// More levels than the ray walks leave out, so they always have a level to
// walk.
VP_HD uint32_t w4_levels(uint32_t n) {
    uint32_t l = kLeafBlockLevels + 1u;
    while (l < kW4MaxLevels && (1ull << (2 * l)) < n) ++l;
    return l;
}
// (4^l - 1) / 3 = 0b0101...01 with l ones
VP_HD uint32_t w4_level_base(uint32_t l) { return 0x55555555u & ((1u << (2u * l)) - 1u); }
// node records of a tree of `levels` levels
VP_HD uint32_t w4_records(uint32_t levels) { return w4_level_base(levels); }

constexpr uint32_t kWalkEnd = 0xffffffffu;

// Position of a walk: the node to load next (kWalkEnd once the walk is
// over), its level, and a trail of 4 bits per level above it: bit c of level
// l is set while child c of the level-l node on the current path is still to
// be visited.  Backtracking takes the deepest level with a pending child and
// computes that child's index from the current node's (an ancestor k levels
// up has the level-local index shifted right by 2k), so a walk keeps no stack
// and loads every visited node exactly once.  Children are taken in
// ascending order, so leaves are met in ascending stored index.
struct Walk4 {
    uint32_t n, lvl;
    uint64_t trail;
};
VP_HD Walk4 walk4_begin() { return Walk4{0u, 0u, 0ull}; }

// Moves the walk past node w.n, whose crossed children are m (4 bits).  On
// the last level the crossed children are leaves: returns them as a mask
// over the stored primitives first .. first + 3; elsewhere returns 0.
// (A branch-free form, descending as a backtrack with this node's children
// added to the trail, measured 12 % slower on the scattering workload: the
// lanes of a wavefront mostly take the same branch.)
VP_HD uint32_t walk4_advance(Walk4& w, uint32_t levels, uint32_t m, uint32_t& first) {
    const bool last = w.lvl + 1u == levels;
    const uint32_t local = w.n - w4_level_base(w.lvl);
    first = 4u * local;
    if (!last && m != 0u) {
        // descend into the first crossed child; the others become pending
        const uint32_t sh = 4u * w.lvl;
        w.trail = (w.trail & ~(0xfull << sh)) | (static_cast<uint64_t>(m & (m - 1u)) << sh);
        w.n = w4_level_base(w.lvl + 1u) + 4u * local + static_cast<uint32_t>(__builtin_ctz(m));
        w.lvl += 1u;
    } else {
        // backtrack to the deepest pending child
        const uint64_t t = w.trail & ((1ull << (4u * w.lvl)) - 1ull);
        if (t == 0ull) {
            w.n = kWalkEnd;
        } else {
            const uint32_t l = static_cast<uint32_t>(63 - __builtin_clzll(t)) >> 2;
            const uint32_t c = static_cast<uint32_t>(__builtin_ctz(static_cast<uint32_t>(t >> (4u * l)) & 0xfu));
            w.trail = t & ~(1ull << (4u * l + c));
            const uint32_t anc = local >> (2u * (w.lvl - l));
            w.n = w4_level_base(l + 1u) + 4u * anc + c;
            w.lvl = l + 1u;
        }
    }
    return last ? m : 0u;
}

// Levels a ray walk of leaf blocks runs over (the tree's L levels less the
// last kLeafBlockLevels): passed to walk4_advance as its level count, it makes
// the walk stop kLeafBlockLevels levels above the last and return the crossed
// children of the nodes there, each the root of block c of leaves
// kLeafBlock c .. kLeafBlock (c + 1) - 1.  The boxes below are then never
// tested: a block is crossed when its box is, and its primitives go to the
// leaf tests, which are exact.  That trades leaf tests, which run at the full
// width of the wave-shared queue, for dependent node loads a walk would make.
VP_HD uint32_t walk4_block_levels(uint32_t levels) { return levels - kLeafBlockLevels; }

// Children still pending in walk w (trail bits of the levels above w.n).
VP_HD uint64_t walk4_pending(const Walk4& w) {
    return w.n == kWalkEnd ? 0ull : w.trail & ((1ull << (4u * w.lvl)) - 1ull);
}

// Splits off the shallowest pending subtree of walk w as a walk of its own
// (part), which starts at that child with an empty trail and so ends once the
// subtree is done; w no longer visits it.  Every subtree is still walked
// exactly once, by one of the two.  Returns false when nothing is pending.
VP_HD bool walk4_split(Walk4& w, Walk4& part) {
    const uint64_t pend = walk4_pending(w);
    if (pend == 0ull) return false;
    const uint32_t bit = static_cast<uint32_t>(__builtin_ctzll(pend));
    const uint32_t l = bit >> 2, c = bit & 3u;
    w.trail &= ~(1ull << bit);
    const uint32_t anc = (w.n - w4_level_base(w.lvl)) >> (2u * (w.lvl - l));
    part = Walk4{w4_level_base(l + 1u) + 4u * anc + c, l + 1u, 0ull};
    return true;
}

}  // namespace vp
