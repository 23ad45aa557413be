// vp_bvh.hip — GPU construction of the scene: primitive records, Morton
// codes, a stable LDS radix sort, the Karras (2012) binary radix tree, a
// bottom-up bounding-box refit and the depth-first "rope" node layout the
// traversal kernels walk without a stack.
//
// Replaces the reference's Morton sort (src/thesis/host/app/renderer.cpp:90-121,
// include/thesis/host/utils/math.h:103-160) and its OptiX GAS/IAS build
// (renderer.cpp:123-173, include/thesis/host/optix/gas.h:47-79).  Morton codes
// are bit-exact with the reference's morton3D on primitive centres; ties are
// broken by input index (a stable sort), which the reference leaves to
// std::sort.
#include "vp_internal.h"

namespace vp {
namespace {

constexpr uint32_t kSortBlock = 256;
constexpr uint32_t kSortItems = 8;
constexpr uint32_t kSortTile = kSortBlock * kSortItems;

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Records, centres and conservative leaf boxes (conservative_box).
__global__ void k_setup(const vp_primitive* in, uint32_t n, PrimRecord* rec, float* centers, float* box) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const vp_primitive p = in[i];
    const f3 c = f3{p.center[0], p.center[1], p.center[2]};
    const f3 s = f3{p.scale[0], p.scale[1], p.scale[2]};
    const PrimRecord r = make_prim(c, p.quat_wxyz[0], p.quat_wxyz[1], p.quat_wxyz[2], p.quat_wxyz[3], s,
                                   f3{p.albedo[0], p.albedo[1], p.albedo[2]}, p.mass);
    rec[i] = r;
    centers[3 * i + 0] = c.x;
    centers[3 * i + 1] = c.y;
    centers[3 * i + 2] = c.z;
    conservative_box(r, p.scale, box + 6 * i, box + 6 * i + 3);
}

// Centre bounds: per-block partial min/max, then one block folds them.
__global__ void k_bounds_partial(const float* centers, uint32_t n, float* partial) {
    __shared__ float smin[3][256], smax[3][256];
    const uint32_t t = threadIdx.x;
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const float c = centers[3 * i + k];
            mn[k] = minf(mn[k], c);
            mx[k] = maxf(mx[k], c);
        }
    for (int k = 0; k < 3; ++k) {
        smin[k][t] = mn[k];
        smax[k][t] = mx[k];
    }
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 3; ++k) {
                smin[k][t] = minf(smin[k][t], smin[k][t + s]);
                smax[k][t] = maxf(smax[k][t], smax[k][t + s]);
            }
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 3; ++k) {
            partial[6 * blockIdx.x + k] = smin[k][0];
            partial[6 * blockIdx.x + 3 + k] = smax[k][0];
        }
}

__global__ void k_bounds_final(const float* partial, uint32_t nb, float* bounds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (uint32_t b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) {
            mn[k] = minf(mn[k], partial[6 * b + k]);
            mx[k] = maxf(mx[k], partial[6 * b + 3 + k]);
        }
    for (int k = 0; k < 3; ++k) {
        bounds[k] = mn[k];
        bounds[3 + k] = 1.0f / maxf(mx[k] + (-mn[k]), 1e-30f);  // inverse extent
    }
}

// morton3D of the reference (host/utils/math.h:117-131).
__global__ void k_morton(const float* centers, uint32_t n, const float* bounds, uint32_t* codes, uint32_t* idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float nrm = (centers[3 * i + k] + (-bounds[k])) * bounds[3 + k];
        q[k] = static_cast<uint32_t>(clampf(nrm, 0.0f, 1.0f) * 1023.0f);
    }
    codes[i] = (expand_bits(q[2]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[0]);
    idx[i] = i;
}

// ---- stable LSD radix sort, 8-bit digits ---------------------------------
// A block owns a tile of 2048 keys; item r of thread t is key base + r*256 + t,
// so processing rounds in order and lanes by id visits keys in index order.
__global__ void __launch_bounds__(kSortBlock) k_hist(const uint32_t* keys, uint32_t n, uint32_t shift, uint32_t* hist,
                                                     uint32_t nblocks) {
    __shared__ uint32_t h[256];
    const uint32_t t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortBlock + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[t * nblocks + blockIdx.x] = h[t];
}

// Exclusive scan of hist (digit-major, block-minor) in one 1024-thread block.
__global__ void __launch_bounds__(1024) k_scan(uint32_t* hist, uint32_t count) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (count + 1023) / 1024;
    const uint32_t b = t * chunk;
    const uint32_t e = min(count, b + chunk);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += hist[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;  // exclusive prefix of this chunk
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(kSortBlock) k_scatter(const uint32_t* keys, const uint32_t* vals, uint32_t n,
                                                        uint32_t shift, const uint32_t* offsets, uint32_t nblocks,
                                                        uint32_t* keys_out, uint32_t* vals_out) {
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_wave[kSortBlock / 64][256];
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t >> 6;
    const uint32_t lane = t & 63u;
    s_base[t] = offsets[t * nblocks + blockIdx.x];
    for (uint32_t w = 0; w < kSortBlock / 64; ++w) s_wave[w][t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortBlock + t;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t val = valid ? vals[i] : 0u;
        const uint32_t digit = (key >> shift) & 255u;
        unsigned long long m = __ballot(valid);
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            m &= bit ? bal : ~bal;
        }
        const uint32_t rank = __popcll(m & lt);
        const uint32_t cnt = __popcll(m);
        if (valid && rank == 0) s_wave[wave][digit] = cnt;
        __syncthreads();
        if (valid) {
            uint32_t pre = s_base[digit];
            for (uint32_t w = 0; w < wave; ++w) pre += s_wave[w][digit];
            keys_out[pre + rank] = key;
            vals_out[pre + rank] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t w = 0; w < kSortBlock / 64; ++w) {
            add += s_wave[w][t];
            s_wave[w][t] = 0;
        }
        s_base[t] += add;
        __syncthreads();
    }
}

// ---- Karras 2012 radix tree ---------------------------------------------
// Keys are (code << 32 | index), unique, so delta = clz64(k_i ^ k_j).
__device__ __forceinline__ int delta(const uint32_t* codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned long long a = (static_cast<unsigned long long>(codes[i]) << 32) | static_cast<uint32_t>(i);
    const unsigned long long b = (static_cast<unsigned long long>(codes[j]) << 32) | static_cast<uint32_t>(j);
    return __clzll(a ^ b);
}

constexpr uint32_t kLeafBit = 0x80000000u;

__global__ void k_karras(const uint32_t* codes, int n, uint32_t* left, uint32_t* right, uint32_t* range_lo,
                         uint32_t* range_hi, int* parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(codes, n, i, j);
    int s = 0;
    for (int div = 2;; div <<= 1) {
        const int t = (l + div - 1) / div;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = min(i, j), hi = max(i, j);
    const uint32_t lc = (lo == gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
    const uint32_t rc = (hi == gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    range_lo[i] = lo;
    range_hi[i] = hi;
    // parent array over the unified index space: inner i -> i, leaf k -> n-1+k
    parent[(lc & kLeafBit) ? (n - 1 + (lc & ~kLeafBit)) : lc] = i;
    parent[(rc & kLeafBit) ? (n - 1 + (rc & ~kLeafBit)) : rc] = i;
}

// Inner boxes are written by other threads of the same launch: read them with
// agent-scope atomic loads so no stale L1 line is used.
__device__ __forceinline__ void box_of(uint32_t child, const float* leaf_box, const float* inner_box, float* b) {
    if (child & kLeafBit) {
        const float* src = leaf_box + 6 * (child & ~kLeafBit);
        for (int k = 0; k < 6; ++k) b[k] = src[k];
    } else {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(inner_box + 6 * child);
        for (int k = 0; k < 6; ++k)
            b[k] = __uint_as_float(__hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
}

// Bottom-up refit: the second thread to reach an inner node builds its box.
__global__ void k_refit(int n, const int* parent, const uint32_t* left, const uint32_t* right, const float* leaf_box,
                        float* inner_box, uint32_t* flags) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int node = parent[n - 1 + k];
    while (node >= 0) {
        __threadfence();
        const uint32_t old = atomicAdd(&flags[node], 1u);
        if (old == 0) return;
        __threadfence();
        float a[6], b[6];
        box_of(left[node], leaf_box, inner_box, a);
        box_of(right[node], leaf_box, inner_box, b);
        uint32_t* dst = reinterpret_cast<uint32_t*>(inner_box + 6 * node);
        for (int q = 0; q < 3; ++q) {
            __hip_atomic_store(dst + q, __float_as_uint(minf(a[q], b[q])), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(dst + 3 + q, __float_as_uint(maxf(a[3 + q], b[3 + q])), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        node = parent[node];
    }
}

__device__ __forceinline__ uint32_t subtree_size(uint32_t child, const uint32_t* range_lo, const uint32_t* range_hi) {
    if (child & kLeafBit) return 1u;
    return 2u * (range_hi[child] - range_lo[child] + 1u) - 1u;
}

// Depth-first position of every node (walk to the root) and the rope layout.
__global__ void k_emit(int n, const int* parent, const uint32_t* left, const uint32_t* right, const uint32_t* range_lo,
                       const uint32_t* range_hi, const float* leaf_box, const float* inner_box, float4* nodes,
                       uint32_t* max_depth) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * n - 1) return;
    // unified index: v < n-1 inner node v; else leaf v-(n-1)
    const bool leaf = v >= n - 1;
    const uint32_t self = leaf ? (kLeafBit | (uint32_t)(v - (n - 1))) : (uint32_t)v;
    uint32_t pos = 0, depth = 0;
    int u = v;
    uint32_t cu = self;
    while (true) {
        const int par = parent[u];
        if (par < 0) break;
        if (left[par] == cu) pos += 1u;
        else pos += 1u + subtree_size(left[par], range_lo, range_hi);
        ++depth;
        u = par;
        cu = (uint32_t)par;
    }
    const uint32_t size = subtree_size(self, range_lo, range_hi);
    const float* b = leaf ? (leaf_box + 6 * (v - (n - 1))) : (inner_box + 6 * v);
    nodes[2 * pos] = make_float4(b[0], b[1], b[2], __uint_as_float(pos + size));
    nodes[2 * pos + 1] = make_float4(b[3], b[4], b[5], __uint_as_float(leaf ? (uint32_t)(v - (n - 1)) : kInnerNode));
    atomicMax(max_depth, depth);
}

__global__ void k_gather(const uint32_t* order, uint32_t n, const PrimRecord* rec, const float* box, PrimRecord* rec_out,
                         float* box_out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = order[k];
    rec_out[k] = rec[i];
    for (int q = 0; q < 6; ++q) box_out[6 * k + q] = box[6 * i + q];
}

__global__ void k_init_parent(int* parent, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) parent[i] = -1;
}

__global__ void k_single_leaf(const float* box, float4* nodes) {
    nodes[0] = make_float4(box[0], box[1], box[2], __uint_as_float(1u));
    nodes[1] = make_float4(box[3], box[4], box[5], __uint_as_float(0u));
}

// ---- walk tree (implicit 4-ary, see vp_internal.h) ----

// Leaf boxes in stored order, taken from the rope tree's leaves.
__global__ void k_w4_leaf_boxes(const float4* nodes, uint32_t nn, float* box) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const float4 a = nodes[2 * i], b = nodes[2 * i + 1];
    const uint32_t prim = __float_as_uint(b.w);
    if (prim == kInnerNode) return;
    float* o = box + 6 * static_cast<size_t>(prim);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
    o[3] = b.x;
    o[4] = b.y;
    o[5] = b.z;
}

__device__ __forceinline__ void w4_store(float4* w4, uint32_t idx, const float lo[3][4], const float hi[3][4]) {
    float4* r = w4 + static_cast<size_t>(kW4RecordF4) * idx;
    for (int a = 0; a < 3; ++a) {
        r[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        r[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    r[6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r[7] = r[6];
}

// Last level: node k holds the leaves 4k .. 4k + 3.
__global__ void k_w4_bottom(const float* box, uint32_t n, uint32_t levels, float4* w4) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (1u << (2u * (levels - 1u)))) return;
    float lo[3][4], hi[3][4];
    for (uint32_t c = 0; c < 4; ++c) {
        const uint32_t j = 4u * k + c;
        for (int a = 0; a < 3; ++a) {
            lo[a][c] = j < n ? box[6 * static_cast<size_t>(j) + a] : INFINITY;
            hi[a][c] = j < n ? box[6 * static_cast<size_t>(j) + 3 + a] : INFINITY;
        }
    }
    w4_store(w4, w4_level_base(levels - 1u) + k, lo, hi);
}

// Level l above the last: the box of child c of node k is the union of the
// children of node 4k + c of level l + 1 (empty children are skipped).
__global__ void k_w4_level(uint32_t l, float4* w4) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (1u << (2u * l))) return;
    float lo[3][4], hi[3][4];
    for (uint32_t c = 0; c < 4; ++c) {
        const float4* r = w4 + static_cast<size_t>(kW4RecordF4) * (w4_level_base(l + 1u) + 4u * k + c);
        const float4 mnx = r[0];
        const float4 full = make_float4(mnx.x < INFINITY ? 1.0f : 0.0f, mnx.y < INFINITY ? 1.0f : 0.0f,
                                        mnx.z < INFINITY ? 1.0f : 0.0f, mnx.w < INFINITY ? 1.0f : 0.0f);
        const bool any = full.x + full.y + full.z + full.w > 0.0f;
        for (int a = 0; a < 3; ++a) {
            const float4 mn = r[a], mx = r[3 + a];
            lo[a][c] = fminf(fminf(mn.x, mn.y), fminf(mn.z, mn.w));
            float h = -INFINITY;
            if (full.x > 0.0f) h = fmaxf(h, mx.x);
            if (full.y > 0.0f) h = fmaxf(h, mx.y);
            if (full.z > 0.0f) h = fmaxf(h, mx.z);
            if (full.w > 0.0f) h = fmaxf(h, mx.w);
            hi[a][c] = any ? h : INFINITY;
        }
    }
    w4_store(w4, w4_level_base(l) + k, lo, hi);
}

// 16-bit records from the 32-bit ones (one thread per record): min bounds
// rounded toward -inf, max bounds toward +inf, packed in the byte order of
// vp_walk.h.  For the last level's records (index >= leaf_base) it also sums
// the surfaces ex ey + ey ez + ez ex of the non-empty leaf boxes, of both
// forms, into sums[0] (32-bit) and sums[1] (16-bit).
__global__ void k_w4_half(const float4* w4, uint32_t records, uint32_t leaf_base, uint4* out, double* sums) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double s32 = 0.0, s16 = 0.0;
    if (k < records) {
        const float4* r = w4 + static_cast<size_t>(kW4RecordF4) * k;
        float b[6][4];
        uint32_t words[12];
        for (int g = 0; g < 6; ++g) {
            const float4 v = r[g];
            b[g][0] = v.x;
            b[g][1] = v.y;
            b[g][2] = v.z;
            b[g][3] = v.w;
            for (int c = 0; c < 4; c += 2) {
                const uint32_t h0 = g < 3 ? f16_down(b[g][c]) : f16_up(b[g][c]);
                const uint32_t h1 = g < 3 ? f16_down(b[g][c + 1]) : f16_up(b[g][c + 1]);
                words[2 * g + c / 2] = h0 | (h1 << 16);
            }
        }
        uint4* o = out + 4 * static_cast<size_t>(k);
        o[0] = make_uint4(words[0], words[1], words[2], words[3]);
        o[1] = make_uint4(words[4], words[5], words[6], words[7]);
        o[2] = make_uint4(words[8], words[9], words[10], words[11]);
        o[3] = make_uint4(0u, 0u, 0u, 0u);
        if (k >= leaf_base) {
            for (int c = 0; c < 4; ++c) {
                if (!(b[0][c] < INFINITY)) continue;
                double e32[3], e16[3];
                for (int a = 0; a < 3; ++a) {
                    e32[a] = static_cast<double>(b[3 + a][c]) - b[a][c];
                    e16[a] = static_cast<double>(f16_value(f16_up(b[3 + a][c]))) - f16_value(f16_down(b[a][c]));
                }
                s32 += e32[0] * e32[1] + e32[1] * e32[2] + e32[2] * e32[0];
                s16 += e16[0] * e16[1] + e16[1] * e16[2] + e16[2] * e16[0];
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        s32 += __shfl_xor(s32, off);
        s16 += __shfl_xor(s16, off);
    }
    if ((threadIdx.x & 63u) == 0u && s16 != 0.0) {
        atomicAdd(sums, s32);
        atomicAdd(sums + 1, s16);
    }
}

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace vp

vp_status vp_build_scene_device(vp_context* ctx, const vp_primitive* host_prims, uint32_t n, vp_scene* sc) {
    using namespace vp;
    hipStream_t s = ctx->stream;
    const uint32_t nb_sort = cdiv(n, kSortTile);
    const uint32_t nbounds = min(cdiv(n, 256u), 1024u);
    DevBuf d_in, d_rec, d_centers, d_box, d_partial, d_bounds, d_codes, d_idx, d_codes2, d_idx2, d_hist, d_box_sorted;
    DevBuf d_left, d_right, d_lo, d_hi, d_parent, d_inner, d_flags, d_depth;
    VP_HIP_TRY(hipMalloc(&d_in.p, sizeof(vp_primitive) * n));
    VP_HIP_TRY(hipMalloc(&d_rec.p, sizeof(PrimRecord) * n));
    VP_HIP_TRY(hipMalloc(&d_centers.p, sizeof(float) * 3 * n));
    VP_HIP_TRY(hipMalloc(&d_box.p, sizeof(float) * 6 * n));
    VP_HIP_TRY(hipMalloc(&d_box_sorted.p, sizeof(float) * 6 * n));
    VP_HIP_TRY(hipMalloc(&d_partial.p, sizeof(float) * 6 * nbounds));
    VP_HIP_TRY(hipMalloc(&d_bounds.p, sizeof(float) * 6));
    VP_HIP_TRY(hipMalloc(&d_codes.p, sizeof(uint32_t) * n));
    VP_HIP_TRY(hipMalloc(&d_idx.p, sizeof(uint32_t) * n));
    VP_HIP_TRY(hipMalloc(&d_codes2.p, sizeof(uint32_t) * n));
    VP_HIP_TRY(hipMalloc(&d_idx2.p, sizeof(uint32_t) * n));
    VP_HIP_TRY(hipMalloc(&d_hist.p, sizeof(uint32_t) * 256 * nb_sort));
    VP_HIP_TRY(hipMalloc(&d_depth.p, sizeof(uint32_t)));
    // the records are padded to a multiple of the leaf block with zero records
    // (zero inverse scales: no ray enters them), so a walk may hand every slot
    // of a leaf block to the leaf tests
    const size_t n_slots = (static_cast<size_t>(n) + kLeafBlock - 1u) / kLeafBlock * kLeafBlock;
    VP_HIP_TRY(hipMalloc(&sc->prims, sizeof(PrimRecord) * n_slots));
    if (n_slots > n) VP_HIP_TRY(hipMemsetAsync(sc->prims + n, 0, sizeof(PrimRecord) * (n_slots - n), s));
    VP_HIP_TRY(hipMalloc(&sc->nodes, sizeof(float4) * 2 * (2 * n - 1)));
    VP_HIP_TRY(hipMalloc(&sc->order, sizeof(uint32_t) * n));
    VP_HIP_TRY(hipMalloc(&sc->morton, sizeof(uint32_t) * n));
    VP_HIP_TRY(hipMemcpyAsync(d_in.p, host_prims, sizeof(vp_primitive) * n, hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipMemsetAsync(d_depth.p, 0, sizeof(uint32_t), s));

    const uint32_t g = cdiv(n, 256u);
    hipLaunchKernelGGL(k_setup, dim3(g), dim3(256), 0, s, (const vp_primitive*)d_in.p, n, (PrimRecord*)d_rec.p,
                       (float*)d_centers.p, (float*)d_box.p);
    hipLaunchKernelGGL(k_bounds_partial, dim3(nbounds), dim3(256), 0, s, (const float*)d_centers.p, n,
                       (float*)d_partial.p);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, (const float*)d_partial.p, nbounds,
                       (float*)d_bounds.p);
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, s, (const float*)d_centers.p, n, (const float*)d_bounds.p,
                       (uint32_t*)d_codes.p, (uint32_t*)d_idx.p);
    VP_HIP_TRY(hipMemcpyAsync(sc->morton, d_codes.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));

    uint32_t* kin = (uint32_t*)d_codes.p;
    uint32_t* vin = (uint32_t*)d_idx.p;
    uint32_t* kout = (uint32_t*)d_codes2.p;
    uint32_t* vout = (uint32_t*)d_idx2.p;
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_hist, dim3(nb_sort), dim3(kSortBlock), 0, s, kin, n, shift, (uint32_t*)d_hist.p,
                           nb_sort);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (uint32_t*)d_hist.p, 256 * nb_sort);
        hipLaunchKernelGGL(k_scatter, dim3(nb_sort), dim3(kSortBlock), 0, s, kin, vin, n, shift,
                           (const uint32_t*)d_hist.p, nb_sort, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // after 4 passes the sorted data is back in kin/vin
    VP_HIP_TRY(hipMemcpyAsync(sc->order, vin, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, s, (const uint32_t*)vin, n, (const PrimRecord*)d_rec.p,
                       (const float*)d_box.p, sc->prims, (float*)d_box_sorted.p);

    if (n == 1) {
        hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(1), 0, s, (const float*)d_box_sorted.p, sc->nodes);
        sc->depth = 0;
    } else {
        VP_HIP_TRY(hipMalloc(&d_left.p, sizeof(uint32_t) * (n - 1)));
        VP_HIP_TRY(hipMalloc(&d_right.p, sizeof(uint32_t) * (n - 1)));
        VP_HIP_TRY(hipMalloc(&d_lo.p, sizeof(uint32_t) * (n - 1)));
        VP_HIP_TRY(hipMalloc(&d_hi.p, sizeof(uint32_t) * (n - 1)));
        VP_HIP_TRY(hipMalloc(&d_parent.p, sizeof(int) * (2 * n - 1)));
        VP_HIP_TRY(hipMalloc(&d_inner.p, sizeof(float) * 6 * (n - 1)));
        VP_HIP_TRY(hipMalloc(&d_flags.p, sizeof(uint32_t) * (n - 1)));
        VP_HIP_TRY(hipMemsetAsync(d_flags.p, 0, sizeof(uint32_t) * (n - 1), s));
        hipLaunchKernelGGL(k_init_parent, dim3(cdiv(2 * n - 1, 256u)), dim3(256), 0, s, (int*)d_parent.p,
                           (int)(2 * n - 1));
        hipLaunchKernelGGL(k_karras, dim3(cdiv(n - 1, 256u)), dim3(256), 0, s, (const uint32_t*)kin, (int)n,
                           (uint32_t*)d_left.p, (uint32_t*)d_right.p, (uint32_t*)d_lo.p, (uint32_t*)d_hi.p,
                           (int*)d_parent.p);
        hipLaunchKernelGGL(k_refit, dim3(g), dim3(256), 0, s, (int)n, (const int*)d_parent.p,
                           (const uint32_t*)d_left.p, (const uint32_t*)d_right.p, (const float*)d_box_sorted.p,
                           (float*)d_inner.p, (uint32_t*)d_flags.p);
        hipLaunchKernelGGL(k_emit, dim3(cdiv(2 * n - 1, 256u)), dim3(256), 0, s, (int)n, (const int*)d_parent.p,
                           (const uint32_t*)d_left.p, (const uint32_t*)d_right.p, (const uint32_t*)d_lo.p,
                           (const uint32_t*)d_hi.p, (const float*)d_box_sorted.p, (const float*)d_inner.p,
                           sc->nodes, (uint32_t*)d_depth.p);
        uint32_t depth = 0;
        VP_HIP_TRY(hipMemcpyAsync(&depth, d_depth.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        VP_HIP_TRY(hipStreamSynchronize(s));
        sc->depth = depth;
    }
    VP_HIP_TRY(hipGetLastError());
    VP_HIP_TRY(hipStreamSynchronize(s));
    return VP_OK;
}

vp_status vp_build_walk_tree(vp_context* ctx, vp_scene* sc, uint32_t walk_format) {
    using namespace vp;
    hipStream_t s = ctx->stream;
    const uint32_t n = sc->n;
    const uint32_t levels = w4_levels(n);
    const size_t records = w4_records(levels);
    if (sc->w4) {
        VP_HIP_TRY(hipStreamSynchronize(s));
        VP_HIP_TRY(hipFree(sc->w4));
        sc->w4 = nullptr;
    }
    DevBuf d_box;
    VP_HIP_TRY(hipMalloc(&d_box.p, sizeof(float) * 6 * static_cast<size_t>(n)));
    VP_HIP_TRY(hipMalloc(&sc->w4, sizeof(float4) * kW4RecordF4 * records));
    const uint32_t nn = 2 * n - 1;
    hipLaunchKernelGGL(k_w4_leaf_boxes, dim3(cdiv(nn, 256u)), dim3(256), 0, s, (const float4*)sc->nodes, nn,
                       (float*)d_box.p);
    const uint32_t bottom = 1u << (2u * (levels - 1u));
    hipLaunchKernelGGL(k_w4_bottom, dim3(cdiv(bottom, 256u)), dim3(256), 0, s, (const float*)d_box.p, n, levels,
                       sc->w4);
    for (uint32_t l = levels - 1u; l-- > 0;)
        hipLaunchKernelGGL(k_w4_level, dim3(cdiv(1u << (2u * l), 256u)), dim3(256), 0, s, l, sc->w4);
    VP_HIP_TRY(hipGetLastError());
    sc->w4_levels = levels;
    sc->w4_half = 0u;
    sc->w4_growth = 0.0;
    if (walk_format == VP_SCENE_WALK_F32) {
        VP_HIP_TRY(hipStreamSynchronize(s));
        return VP_OK;
    }
    // the 16-bit records, kept when forced or when the leaf boxes' summed
    // surface grows by at most 1/16 (more leaf tests and node visits, against
    // half the bytes per node visit)
    DevBuf d_half, d_sums;
    VP_HIP_TRY(hipMalloc(&d_half.p, static_cast<size_t>(kW4hRecordB) * records));
    VP_HIP_TRY(hipMalloc(&d_sums.p, 2 * sizeof(double)));
    VP_HIP_TRY(hipMemsetAsync(d_sums.p, 0, 2 * sizeof(double), s));
    hipLaunchKernelGGL(k_w4_half, dim3(cdiv(static_cast<uint32_t>(records), 256u)), dim3(256), 0, s,
                       (const float4*)sc->w4, static_cast<uint32_t>(records), w4_level_base(levels - 1u),
                       (uint4*)d_half.p, (double*)d_sums.p);
    VP_HIP_TRY(hipGetLastError());
    double sums[2] = {0.0, 0.0};
    VP_HIP_TRY(hipMemcpyAsync(sums, d_sums.p, sizeof(sums), hipMemcpyDeviceToHost, s));
    VP_HIP_TRY(hipStreamSynchronize(s));
    const double growth = sums[0] > 0.0 ? sums[1] / sums[0] - 1.0 : 0.0;
    sc->w4_growth = growth;
    if (walk_format == VP_SCENE_WALK_F16 || (growth <= 1.0 / 16.0)) {
        VP_HIP_TRY(hipFree(sc->w4));
        sc->w4 = static_cast<float4*>(d_half.p);
        d_half.p = nullptr;
        sc->w4_half = 1u;
    }
    return VP_OK;
}
