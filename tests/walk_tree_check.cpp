// Host check of the walk-tree traversal (cuda-volprim_amd/csrc/vp_walk.h).
//
// Builds the implicit 4-ary tree over random leaf boxes with the rule of the
// device build (a child box is the union of its children, empty children are
// +inf boxes), walks it with walk4_advance for random rays and points, and
// checks against brute force over the leaves:
//   * the walk meets exactly the leaves whose own box passes the test, in
//     ascending order (the device walks rely on both);
//   * every node index it loads is inside the tree and no node is loaded
//     twice;
//   * a walk whose segment end is lowered while it runs (as the argmin walk
//     lowers it) still meets every leaf that passes at the final end;
//   * a walk split among several walkers (walk4_split, as the wavefront
//     shares out running walks) meets every leaf exactly once;
//   * the near/far form of the node test decides every child box (empty
//     ones included) exactly as the min/max form;
//   * the 16-bit tree (every bound rounded outward to binary16, as the
//     device build stores it) contains each 32-bit box, and its walks meet
//     exactly the leaves whose rounded box passes, a superset of the 32-bit
//     walk's leaves.
// Prints "ok <walks>" and exits 0, or prints the first mismatch and exits 1.
// Usage: walk_tree_check <seed> <walks per size>
//        walk_tree_check f16 <in> <out>: reads float32 values from <in> and
//        writes f16_down and f16_up of each (uint16 pairs) and f16_value of
//        both (float32 pairs) to <out>.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "vp_walk.h"

namespace {

using vp::Walk4;

struct Box {
    float lo[3], hi[3];
};

bool slab(const Box& b, const float o[3], const float inv[3], float tmax) {
    float tn = 0.0f, tf = tmax;
    float t0[3], t1[3];
    for (int a = 0; a < 3; ++a) {
        const float oi = o[a] * inv[a];
        t0[a] = std::fma(b.lo[a], inv[a], -oi);
        t1[a] = std::fma(b.hi[a], inv[a], -oi);
    }
    // the device formula: max of the near planes and 0, min of the far planes and tmax
    tn = std::fmax(std::fmax(std::fmin(t0[0], t1[0]), std::fmin(t0[1], t1[1])), std::fmax(std::fmin(t0[2], t1[2]), 0.0f));
    tf = std::fmin(std::fmin(std::fmax(t0[0], t1[0]), std::fmax(t0[1], t1[1])), std::fmin(std::fmax(t0[2], t1[2]), tmax));
    return tn <= tf;
}

// The near/far form of the device walks (w4_ray_mask_nf): the sign of each
// direction component picks the near bound, and the near and far distances
// are taken without a per-axis min/max.  Must decide exactly as slab().
bool slab_nf(const Box& b, const float o[3], const float inv[3], float tmax) {
    float tn = 0.0f, tf = tmax;
    float nr[3], fr[3];
    for (int a = 0; a < 3; ++a) {
        const float oi = o[a] * inv[a];
        const bool neg = std::signbit(inv[a]);
        nr[a] = std::fma(neg ? b.hi[a] : b.lo[a], inv[a], -oi);
        fr[a] = std::fma(neg ? b.lo[a] : b.hi[a], inv[a], -oi);
    }
    tn = std::fmax(std::fmax(nr[0], nr[1]), std::fmax(nr[2], 0.0f));
    tf = std::fmin(std::fmin(fr[0], fr[1]), std::fmin(fr[2], tmax));
    return tn <= tf;
}

bool inside(const Box& b, const float x[3]) {
    for (int a = 0; a < 3; ++a)
        if (!(x[a] >= b.lo[a] && x[a] <= b.hi[a])) return false;
    return true;
}

struct Tree {
    uint32_t levels = 0;
    std::vector<Box> child;  // 4 per record
};

Tree build(const std::vector<Box>& leaves) {
    const uint32_t n = static_cast<uint32_t>(leaves.size());
    Tree t;
    t.levels = vp::w4_levels(n);
    t.child.resize(4 * static_cast<size_t>(vp::w4_records(t.levels)));
    const Box empty{{INFINITY, INFINITY, INFINITY}, {INFINITY, INFINITY, INFINITY}};
    const uint32_t L = t.levels;
    for (uint32_t k = 0; k < (1u << (2 * (L - 1))); ++k)
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t j = 4 * k + c;
            t.child[4 * (vp::w4_level_base(L - 1) + k) + c] = j < n ? leaves[j] : empty;
        }
    for (uint32_t l = L - 1; l-- > 0;)
        for (uint32_t k = 0; k < (1u << (2 * l)); ++k)
            for (uint32_t c = 0; c < 4; ++c) {
                const uint32_t r = vp::w4_level_base(l + 1) + 4 * k + c;
                Box u{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
                bool any = false;
                for (uint32_t g = 0; g < 4; ++g) {
                    const Box& b = t.child[4 * r + g];
                    if (!(b.lo[0] < INFINITY)) continue;
                    any = true;
                    for (int a = 0; a < 3; ++a) {
                        u.lo[a] = std::fmin(u.lo[a], b.lo[a]);
                        u.hi[a] = std::fmax(u.hi[a], b.hi[a]);
                    }
                }
                t.child[4 * (vp::w4_level_base(l) + k) + c] = any ? u : empty;
            }
    return t;
}

// the 16-bit tree of t: every bound rounded outward to binary16, decoded
Tree half_of(const Tree& t) {
    Tree h = t;
    for (Box& b : h.child)
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = vp::f16_value(vp::f16_down(b.lo[a]));
            b.hi[a] = vp::f16_value(vp::f16_up(b.hi[a]));
        }
    return h;
}

int f16_mode(const char* in, const char* out) {
    FILE* f = std::fopen(in, "rb");
    if (!f) return 2;
    std::vector<float> x;
    float v;
    while (std::fread(&v, sizeof v, 1, f) == 1) x.push_back(v);
    std::fclose(f);
    FILE* g = std::fopen(out, "wb");
    if (!g) return 2;
    for (float xv : x) {
        const uint16_t h[2] = {vp::f16_down(xv), vp::f16_up(xv)};
        const float d[2] = {vp::f16_value(h[0]), vp::f16_value(h[1])};
        std::fwrite(h, sizeof h, 1, g);
        std::fwrite(d, sizeof d, 1, g);
    }
    std::fclose(g);
    return 0;
}

int fail(const char* what, uint32_t n, uint32_t walk) {
    std::printf("FAIL %s: n=%u walk=%u\n", what, n, walk);
    return 1;
}

// Walks the tree; mask(node) gives the crossed children of a node, and
// after each node on_node() may lower the segment end.  Returns 0 and the
// leaves met, or 1 on an index error.
template <typename Mask>
int walk(const Tree& t, Mask&& mask, std::vector<uint32_t>& met) {
    const uint32_t records = vp::w4_records(t.levels);
    std::vector<uint8_t> seen(records, 0);
    Walk4 w = vp::walk4_begin();
    while (w.n != vp::kWalkEnd) {
        if (w.n >= records || w.lvl >= t.levels) return 1;
        if (w.n < vp::w4_level_base(w.lvl) || w.n >= vp::w4_level_base(w.lvl + 1)) return 1;
        if (seen[w.n]++) return 1;
        const uint32_t m = mask(w.n);
        uint32_t first = 0;
        for (uint32_t lm = vp::walk4_advance(w, t.levels, m, first); lm; lm &= lm - 1)
            met.push_back(first + static_cast<uint32_t>(__builtin_ctz(lm)));
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc > 3 && std::string(argv[1]) == "f16") return f16_mode(argv[2], argv[3]);
    const uint32_t seed = argc > 1 ? static_cast<uint32_t>(std::atoi(argv[1])) : 1u;
    const uint32_t walks = argc > 2 ? static_cast<uint32_t>(std::atoi(argv[2])) : 200u;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f), H(0.01f, 0.25f), T(0.0f, 1.0f);
    const uint32_t sizes[] = {1, 2, 3, 4, 5, 15, 16, 17, 63, 64, 65, 255, 256, 257, 1000, 4096, 5000, 20000};
    uint64_t total = 0;
    for (uint32_t n : sizes) {
        std::vector<Box> leaves(n);
        for (auto& b : leaves)
            for (int a = 0; a < 3; ++a) {
                const float c = U(rng), h = H(rng);
                b.lo[a] = c - h;
                b.hi[a] = c + h;
            }
        const Tree t = build(leaves);
        const Tree th = half_of(t);
        for (size_t i = 0; i < t.child.size(); ++i)
            for (int a = 0; a < 3; ++a)
                if (!(th.child[i].lo[a] <= t.child[i].lo[a] && th.child[i].hi[a] >= t.child[i].hi[a]))
                    return fail("16-bit box does not contain the 32-bit box", n, static_cast<uint32_t>(i));
        for (uint32_t k = 0; k < walks; ++k) {
            // ray: origin inside or outside the leaves' region, direction random
            float o[3], d[3], inv[3];
            const float spread = (k & 1) ? 1.2f : 3.0f;
            float d2 = 0.0f;
            for (int a = 0; a < 3; ++a) {
                o[a] = spread * U(rng);
                d[a] = U(rng);
                if (k % 7 == 3 && a == 1) d[a] = 0.0f;  // axis-parallel rays
                d2 += d[a] * d[a];
            }
            if (d2 == 0.0f) d[0] = d2 = 1.0f;
            for (int a = 0; a < 3; ++a) {
                d[a] /= std::sqrt(d2);
                const float x = std::fabs(d[a]) < 1e-20f ? std::copysign(1e-20f, d[a]) : d[a];
                inv[a] = 1.0f / x;
            }
            const float tmax = (k % 3 == 0) ? 1e20f : 4.0f * T(rng);
            std::vector<uint32_t> met, want;
            for (uint32_t j = 0; j < n; ++j)
                if (slab(leaves[j], o, inv, tmax)) want.push_back(j);
            bool nf_same = true;
            if (walk(t, [&](uint32_t node) {
                    uint32_t m = 0;
                    for (uint32_t c = 0; c < 4; ++c) {
                        const bool h = slab(t.child[4 * node + c], o, inv, tmax);
                        nf_same &= h == slab_nf(t.child[4 * node + c], o, inv, tmax);
                        m |= h ? (1u << c) : 0u;
                    }
                    return m;
                }, met))
                return fail("ray walk index", n, k);
            if (!nf_same) return fail("near/far node test decides differently", n, k);
            if (met != want) return fail("ray walk leaves", n, k);
            // the 16-bit tree: exactly the leaves whose rounded box passes,
            // which include every leaf of the 32-bit walk
            {
                const uint32_t lb = vp::w4_level_base(th.levels - 1);
                std::vector<uint32_t> met16, want16;
                for (uint32_t j = 0; j < n; ++j)
                    if (slab(th.child[4 * lb + j], o, inv, tmax)) want16.push_back(j);
                if (walk(th, [&](uint32_t node) {
                        uint32_t m = 0;
                        for (uint32_t c = 0; c < 4; ++c) {
                            const bool h = slab(th.child[4 * node + c], o, inv, tmax);
                            nf_same &= h == slab_nf(th.child[4 * node + c], o, inv, tmax);
                            m |= h ? (1u << c) : 0u;
                        }
                        return m;
                    }, met16))
                    return fail("16-bit ray walk index", n, k);
                if (!nf_same) return fail("near/far node test decides differently on the 16-bit tree", n, k);
                if (met16 != want16) return fail("16-bit ray walk leaves", n, k);
                if (!std::includes(met16.begin(), met16.end(), want.begin(), want.end()))
                    return fail("16-bit ray walk lost a leaf of the 32-bit walk", n, k);
            }
            // the same ray with the end lowered as the walk goes (argmin style):
            // every leaf passing at the final end must be met
            float tcur = 1e20f;
            std::vector<uint32_t> met2;
            if (walk(t, [&](uint32_t node) {
                    uint32_t m = 0;
                    for (uint32_t c = 0; c < 4; ++c) m |= slab(t.child[4 * node + c], o, inv, tcur) ? (1u << c) : 0u;
                    tcur = std::fmax(tmax, tcur * 0.7f);
                    return m;
                }, met2))
                return fail("shrinking walk index", n, k);
            for (uint32_t j : want) {
                bool found = false;
                for (uint32_t v : met2) found |= v == j;
                if (!found) return fail("shrinking walk missed a leaf", n, k);
            }
            // the same ray walked by several walkers that split pending
            // subtrees off each other at random (the wavefront's work
            // sharing): every leaf exactly once, none lost
            {
                std::vector<Walk4> ws{vp::walk4_begin()};
                std::vector<uint32_t> got;
                const uint32_t records = vp::w4_records(t.levels);
                std::vector<uint8_t> seen(records, 0);
                std::uniform_int_distribution<uint32_t> pick(0, 1 << 20);
                while (!ws.empty()) {
                    const uint32_t i = pick(rng) % ws.size();
                    if (pick(rng) % 3 == 0 && ws.size() < 64) {
                        Walk4 part;
                        if (vp::walk4_split(ws[i], part)) ws.push_back(part);
                        continue;
                    }
                    Walk4& w = ws[i];
                    if (w.n == vp::kWalkEnd) {
                        ws.erase(ws.begin() + i);
                        continue;
                    }
                    if (w.n >= records || seen[w.n]++) return fail("split walk index", n, k);
                    uint32_t m = 0;
                    for (uint32_t c = 0; c < 4; ++c) m |= slab(t.child[4 * w.n + c], o, inv, tmax) ? (1u << c) : 0u;
                    uint32_t first = 0;
                    for (uint32_t lm = vp::walk4_advance(w, t.levels, m, first); lm; lm &= lm - 1)
                        got.push_back(first + static_cast<uint32_t>(__builtin_ctz(lm)));
                }
                std::sort(got.begin(), got.end());
                if (got != want) return fail("split walk leaves", n, k);
            }
            // point query
            float x[3];
            for (int a = 0; a < 3; ++a) x[a] = 1.1f * U(rng);
            std::vector<uint32_t> pm, pw;
            for (uint32_t j = 0; j < n; ++j)
                if (inside(leaves[j], x)) pw.push_back(j);
            if (walk(t, [&](uint32_t node) {
                    uint32_t m = 0;
                    for (uint32_t c = 0; c < 4; ++c) m |= inside(t.child[4 * node + c], x) ? (1u << c) : 0u;
                    return m;
                }, pm))
                return fail("point walk index", n, k);
            if (pm != pw) return fail("point walk leaves", n, k);
            std::vector<uint32_t> pm16;
            if (walk(th, [&](uint32_t node) {
                    uint32_t m = 0;
                    for (uint32_t c = 0; c < 4; ++c) m |= inside(th.child[4 * node + c], x) ? (1u << c) : 0u;
                    return m;
                }, pm16))
                return fail("16-bit point walk index", n, k);
            if (!std::includes(pm16.begin(), pm16.end(), pw.begin(), pw.end()))
                return fail("16-bit point walk lost a leaf", n, k);
            total += 3;
        }
    }
    std::printf("ok %llu\n", static_cast<unsigned long long>(total));
    return 0;
}
