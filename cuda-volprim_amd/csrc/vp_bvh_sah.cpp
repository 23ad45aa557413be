// vp_bvh_sah.cpp — optional host-built trees: a surface-area-heuristic BVH
// (VP_SCENE_BVH_SAH) and a balanced median-split tree (VP_SCENE_BVH_MEDIAN).
//
// The default scene build is the GPU LBVH of vp_bvh.hip.  Its splits follow
// Morton-code bits, which leaves sibling boxes overlapping when primitive
// sizes vary; shadow rays then visit more nodes.  This builder makes a binned
// SAH tree over the same conservative leaf boxes (32 bins per axis, one
// primitive per leaf) on the host and lays it out exactly like the LBVH:
// depth-first nodes with skip links, primitives stored in leaf order.  It
// replaces the LBVH's nodes, stored records and order; the Morton codes the
// GPU build computed stay available for export.  Render results do not
// depend on the tree (every query is consumed order-independently or in
// stored-index order), so both builds give the same images for the same
// stored order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vp_internal.h"

namespace {

struct Box {
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    void grow(const Box& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    double area() const {
        const double dx = std::max(0.0f, hi[0] - lo[0]), dy = std::max(0.0f, hi[1] - lo[1]),
                     dz = std::max(0.0f, hi[2] - lo[2]);
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Tree {
    // binary tree in build order; leaf: prim >= 0
    struct N {
        Box box;
        int32_t left = -1, right = -1, prim = -1;
    };
    std::vector<N> nodes;
};

constexpr int kBins = 32;

// Top-down binned SAH over idx[a, b); returns the node index.
int32_t build(Tree& T, std::vector<uint32_t>& idx, uint32_t a, uint32_t b, const std::vector<Box>& boxes,
              const std::vector<float>& cen) {
    struct Item {
        uint32_t a, b;
        int32_t node;
    };
    const int32_t root = static_cast<int32_t>(T.nodes.size());
    T.nodes.emplace_back();
    std::vector<Item> stack{{a, b, root}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        Box nb, cb;
        for (uint32_t i = it.a; i < it.b; ++i) {
            nb.grow(boxes[idx[i]]);
            Box c;
            for (int k = 0; k < 3; ++k) c.lo[k] = c.hi[k] = cen[3 * idx[i] + k];
            cb.grow(c);
        }
        T.nodes[it.node].box = nb;
        const uint32_t cnt = it.b - it.a;
        if (cnt == 1) {
            T.nodes[it.node].prim = static_cast<int32_t>(idx[it.a]);
            continue;
        }
        // best split over the three axes
        double best = 1e300;
        int best_axis = -1, best_bin = 0;
        for (int ax = 0; ax < 3; ++ax) {
            const float ext = cb.hi[ax] - cb.lo[ax];
            if (!(ext > 0.0f)) continue;
            const float scale = kBins / ext;
            Box bb[kBins];
            uint32_t bc[kBins] = {};
            for (uint32_t i = it.a; i < it.b; ++i) {
                int k = static_cast<int>((cen[3 * idx[i] + ax] - cb.lo[ax]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                bb[k].grow(boxes[idx[i]]);
                ++bc[k];
            }
            double right_area[kBins];
            uint32_t right_cnt[kBins];
            Box acc;
            uint32_t c = 0;
            for (int k = kBins - 1; k > 0; --k) {
                acc.grow(bb[k]);
                c += bc[k];
                right_area[k] = acc.area();
                right_cnt[k] = c;
            }
            Box left;
            uint32_t lc = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                left.grow(bb[k]);
                lc += bc[k];
                if (lc == 0 || right_cnt[k + 1] == 0) continue;
                const double cost = left.area() * lc + right_area[k + 1] * right_cnt[k + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = ax;
                    best_bin = k;
                }
            }
        }
        uint32_t mid;
        if (best_axis >= 0) {
            const float scale = kBins / (cb.hi[best_axis] - cb.lo[best_axis]);
            auto* first = idx.data() + it.a;
            auto* last = idx.data() + it.b;
            auto* m = std::partition(first, last, [&](uint32_t p) {
                int k = static_cast<int>((cen[3 * p + best_axis] - cb.lo[best_axis]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                return k <= best_bin;
            });
            mid = static_cast<uint32_t>(m - idx.data());
        } else {
            mid = it.a;  // all centres coincide
        }
        if (mid == it.a || mid == it.b) {
            // no usable split: halves in input order
            mid = it.a + cnt / 2;
        }
        const int32_t l = static_cast<int32_t>(T.nodes.size());
        T.nodes.emplace_back();
        const int32_t r = static_cast<int32_t>(T.nodes.size());
        T.nodes.emplace_back();
        T.nodes[it.node].left = l;
        T.nodes[it.node].right = r;
        stack.push_back({mid, it.b, r});
        stack.push_back({it.a, mid, l});
    }
    return root;
}

// Balanced median splits over idx[0, n): a range of at most `cap` leaf slots
// (cap a power of two) splits into cap/2 primitives and the rest, so every
// split falls on the boundary of a 4-wide walk-tree record (whose groups
// are 4^k consecutive leaves starting at a multiple of 4^k) and every
// record's children hold disjoint, spatially split quarters of its
// primitives.  The split axis is the one whose two halves have the smaller
// summed box surface; within an axis the order is (centre, input index), a
// strict order, so the tree does not depend on how ties are broken.
int32_t build_median(Tree& T, std::vector<uint32_t>& idx, uint32_t n, const std::vector<Box>& boxes,
                     const std::vector<float>& cen) {
    struct Item {
        uint32_t a, b, cap;
        int32_t node;
    };
    uint32_t cap0 = 1;
    while (cap0 < n) cap0 *= 4u;
    const int32_t root = static_cast<int32_t>(T.nodes.size());
    T.nodes.emplace_back();
    std::vector<Item> stack{{0u, n, cap0, root}};
    std::vector<uint32_t> trial[3];
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const uint32_t cnt = it.b - it.a;
        while (it.cap > 1u && cnt <= it.cap / 2u) it.cap /= 2u;
        if (cnt == 1) {
            T.nodes[it.node].box = boxes[idx[it.a]];
            T.nodes[it.node].prim = static_cast<int32_t>(idx[it.a]);
            continue;
        }
        const uint32_t nl = it.cap / 2u;  // cnt > nl here
        double best = 1e300;
        int best_axis = 0;
        Box nb;
        for (uint32_t i = it.a; i < it.b; ++i) nb.grow(boxes[idx[i]]);
        T.nodes[it.node].box = nb;
        for (int ax = 0; ax < 3; ++ax) {
            std::vector<uint32_t>& t = trial[ax];
            t.assign(idx.begin() + it.a, idx.begin() + it.b);
            std::nth_element(t.begin(), t.begin() + nl, t.end(), [&](uint32_t p, uint32_t q) {
                const float cp = cen[3 * p + ax], cq = cen[3 * q + ax];
                return cp < cq || (cp == cq && p < q);
            });
            Box l, r;
            for (uint32_t i = 0; i < nl; ++i) l.grow(boxes[t[i]]);
            for (uint32_t i = nl; i < cnt; ++i) r.grow(boxes[t[i]]);
            const double cost = l.area() + r.area();
            if (cost < best) {
                best = cost;
                best_axis = ax;
            }
        }
        std::copy(trial[best_axis].begin(), trial[best_axis].end(), idx.begin() + it.a);
        const int32_t l = static_cast<int32_t>(T.nodes.size());
        T.nodes.emplace_back();
        const int32_t r = static_cast<int32_t>(T.nodes.size());
        T.nodes.emplace_back();
        T.nodes[it.node].left = l;
        T.nodes[it.node].right = r;
        stack.push_back({it.a + nl, it.b, it.cap / 2u, r});
        stack.push_back({it.a, it.a + nl, it.cap / 2u, l});
    }
    return root;
}

}  // namespace

vp_status vp_build_scene_sah(vp_context* ctx, const vp_primitive* in, uint32_t n, vp_scene* sc, bool median) {
    // records and conservative boxes exactly as the GPU setup kernel makes them
    std::vector<vp::PrimRecord> rec(n);
    std::vector<Box> boxes(n);
    std::vector<float> cen(3 * static_cast<size_t>(n));
    for (uint32_t i = 0; i < n; ++i) {
        const vp_primitive& p = in[i];
        const vp::f3 c = vp::f3{p.center[0], p.center[1], p.center[2]};
        const vp::f3 s = vp::f3{p.scale[0], p.scale[1], p.scale[2]};
        rec[i] = vp::make_prim(c, p.quat_wxyz[0], p.quat_wxyz[1], p.quat_wxyz[2], p.quat_wxyz[3], s,
                               vp::f3{p.albedo[0], p.albedo[1], p.albedo[2]}, p.mass);
        vp::conservative_box(rec[i], p.scale, boxes[i].lo, boxes[i].hi);
        for (int k = 0; k < 3; ++k) cen[3 * i + k] = p.center[k];
    }
    Tree T;
    T.nodes.reserve(2 * static_cast<size_t>(n));
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    if (median)
        build_median(T, idx, n, boxes, cen);
    else
        build(T, idx, 0, n, boxes, cen);
    // depth-first layout with skip links; leaves numbered in visiting order
    const size_t nn = 2 * static_cast<size_t>(n) - 1;
    std::vector<float4> nodes(2 * nn);
    std::vector<uint32_t> order;
    order.reserve(n);
    std::vector<vp::PrimRecord> stored;
    stored.reserve(n);
    uint32_t depth = 0;
    struct Frame {
        int32_t node;
        uint32_t level;
    };
    std::vector<Frame> st{{0, 1}};
    uint32_t pos = 0;
    std::vector<uint32_t> size_of(T.nodes.size(), 1);
    // subtree sizes (children are created after their parent)
    for (size_t k = T.nodes.size(); k-- > 0;)
        if (T.nodes[k].left >= 0) size_of[k] = 1 + size_of[T.nodes[k].left] + size_of[T.nodes[k].right];
    while (!st.empty()) {
        const Frame f = st.back();
        st.pop_back();
        const auto& N = T.nodes[f.node];
        depth = std::max(depth, f.level);
        const uint32_t skip = pos + size_of[f.node];
        uint32_t prim_field = vp::kInnerNode;
        if (N.prim >= 0) {
            prim_field = static_cast<uint32_t>(order.size());
            order.push_back(static_cast<uint32_t>(N.prim));
            stored.push_back(rec[N.prim]);
        }
        float sk, pf;
        std::memcpy(&sk, &skip, 4);
        std::memcpy(&pf, &prim_field, 4);
        nodes[2 * pos] = make_float4(N.box.lo[0], N.box.lo[1], N.box.lo[2], sk);
        nodes[2 * pos + 1] = make_float4(N.box.hi[0], N.box.hi[1], N.box.hi[2], pf);
        ++pos;
        if (N.left >= 0) {
            st.push_back({N.right, f.level + 1});
            st.push_back({N.left, f.level + 1});
        }
    }
    hipStream_t s = ctx->stream;
    VP_HIP_TRY(hipMemcpyAsync(sc->nodes, nodes.data(), sizeof(float4) * nodes.size(), hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipMemcpyAsync(sc->prims, stored.data(), sizeof(vp::PrimRecord) * n, hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipMemcpyAsync(sc->order, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipStreamSynchronize(s));
    sc->depth = depth;
    return VP_OK;
}
