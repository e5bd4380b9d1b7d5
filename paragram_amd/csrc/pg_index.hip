// LBVH build for gfx950: exact point AABB and clipping box, 30-bit Morton
// keys, stable LSD radix sort staged through LDS, Karras hierarchy, and a
// bottom-up refit that fuses the AABB, the max-weight bound and the leaf
// collapse into the traversal-form node records.
//
// Replaces the reference's cuBQL gpuBuilder + create_boxes_kernel
// (pwr_bvh.cu:24-39,217-339) and the max-weight propagation kernels
// (pwr_bvh.cu:42-157).  Leaves name contiguous ranges of the sorted points,
// so no primitive-id indirection is needed.
#include "pg_common.hpp"
#include "pg_kernels.hpp"

#include <float.h>

namespace pg {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint64_t lanemask_lt64() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ------------------------------------------------------------------ AABB
__global__ void __launch_bounds__(kReduceThreads) bbox_partial_kernel(const float *__restrict__ pts, int64_t n,
                                                                      float *__restrict__ partial) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            const float v = pts[3 * i + k];
            mn[k] = fminf(mn[k], v);
            mx[k] = fmaxf(mx[k], v);
        }
    }
    __shared__ float red[6][kReduceThreads];
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = mn[k];
        red[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(kReduceThreads) bbox_final_kernel(const float *__restrict__ partial, int nparts,
                                                                    IndexHeader *__restrict__ hdr, int64_t n,
                                                                    int has_w, int leaf_size) {
    __shared__ float red[6][kReduceThreads];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], partial[6 * i + k]);
            mx[k] = fmaxf(mx[k], partial[6 * i + 3 + k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = mn[k];
        red[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) {
            const float lo = red[k][0], hi = red[3 + k][0];
            hdr->raw_lo[k] = lo;
            hdr->raw_hi[k] = hi;
            // voronoi_ultra.cu:237-238 then voronoi_convex_cell.cuh:187-193,
            // two FP32 roundings in this order.
            const float a = __fsub_rn(lo, 1.0f);
            const float b = __fadd_rn(hi, 1.0f);
            hdr->box_lo[k] = __fsub_rn(a, 1e-5f);
            hdr->box_hi[k] = __fadd_rn(b, 1e-5f);
        }
        hdr->n = (uint32_t)n;
        hdr->has_weights = (uint32_t)has_w;
        hdr->leaf_size = (uint32_t)leaf_size;
        hdr->root_ref = (n > leaf_size) ? 0u : leaf_ref(0u, (uint32_t)n);
        hdr->has_duplicates = 0u;
    }
}

// ------------------------------------------------------------------ Morton
__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// 30-bit Morton key of each point on a cubic 1024^3 grid over the AABB (one
// scale for all axes so flat clouds keep their aspect).
__global__ void morton_kernel(const float *__restrict__ pts, int64_t n, const IndexHeader *__restrict__ hdr,
                              uint32_t *__restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ex = fmaxf(fmaxf(hdr->raw_hi[0] - hdr->raw_lo[0], hdr->raw_hi[1] - hdr->raw_lo[1]),
                           hdr->raw_hi[2] - hdr->raw_lo[2]);
    const float scale = ex > 0.0f ? 1024.0f / ex : 0.0f;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float t = (pts[3 * i + k] - hdr->raw_lo[k]) * scale;
        t = fminf(fmaxf(t, 0.0f), 1023.0f);
        q[k] = (uint32_t)t;
    }
    keys[i] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
}

// ------------------------------------------------------------------ radix sort
// Pass histogram: digit-major counts hist[d * nblocks + b] for the tile of
// block b.
__global__ void __launch_bounds__(kSortThreads) radix_hist_kernel(const uint32_t *__restrict__ keys, int64_t n,
                                                                  int shift, uint32_t *__restrict__ hist, int nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        const int64_t i = base + it * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: item (it, tid) of the tile has position it*256 + tid; its
// rank among equal digits of the tile is the running count of earlier
// iterations, plus equal digits of lower waves in this iteration, plus equal
// digits of lower lanes in its own wave (8 ballots give the peer mask).
__global__ void __launch_bounds__(kSortThreads) radix_scatter_kernel(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, int64_t n, int shift,
    const uint32_t *__restrict__ hist_scanned, int nblocks, uint32_t *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out) {
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[kWaves][256];
    __shared__ uint32_t gbase[256];
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    run[tid] = 0;
    gbase[tid] = hist_scanned[(int64_t)tid * nblocks + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    const uint64_t lt = lanemask_lt64();
    for (int it = 0; it < kSortItems; ++it) {
        for (int w = 0; w < kWaves; ++w) wc[w][tid] = 0;
        __syncthreads();
        const int64_t i = base + it * kSortThreads + tid;
        const bool valid = i < n;
        uint32_t key = 0, val = 0, d = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in ? vals_in[i] : (uint32_t)i;
            d = (key >> shift) & 255u;
        }
        uint64_t peers = wave_ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = wave_ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wc[w][d];
            const uint32_t dst = gbase[d] + pos;
            if ((int64_t)dst < n) {
                keys_out[dst] = key;
                vals_out[dst] = val;
            }
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kWaves; ++w) add += wc[w][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ reorder
__global__ void reorder_kernel(const float *__restrict__ pts, const float *__restrict__ w, int64_t n,
                               const uint32_t *__restrict__ perm, float4 *__restrict__ sorted,
                               uint32_t *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    if ((int64_t)j >= n) return;
    sorted[i] = make_float4(pts[3 * (int64_t)j], pts[3 * (int64_t)j + 1], pts[3 * (int64_t)j + 2], w ? w[j] : 0.0f);
    inv[j] = (uint32_t)i;
}

// ------------------------------------------------------------------ Karras
__device__ __forceinline__ int karras_delta(const uint32_t *__restrict__ keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// T. Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", HPG 2012: internal node i finds its range by exponential
// and binary search on the common-prefix length, then its split.
__global__ void karras_kernel(const uint32_t *__restrict__ keys, int64_t n, uint32_t *__restrict__ children,
                              uint32_t *__restrict__ parents, uint2 *__restrict__ ranges) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = karras_delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (karras_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = karras_delta(keys, n, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t split = i + s * d + (d < 0 ? -1 : 0);
    const int64_t first = i < j ? i : j, last = i < j ? j : i;
    uint32_t left, right;
    if (first == split) {
        left = 0x80000000u | (uint32_t)split;
        parents[(n - 1) + split] = (uint32_t)i;
    } else {
        left = (uint32_t)split;
        parents[split] = (uint32_t)i;
    }
    if (last == split + 1) {
        right = 0x80000000u | (uint32_t)(split + 1);
        parents[(n - 1) + split + 1] = (uint32_t)i;
    } else {
        right = (uint32_t)(split + 1);
        parents[split + 1] = (uint32_t)i;
    }
    children[2 * i] = left;
    children[2 * i + 1] = right;
    ranges[i] = make_uint2((uint32_t)first, (uint32_t)last);
    if (i == 0) parents[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ float load_coherent(const float *p) {
    return __uint_as_float(__hip_atomic_load((const uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ uint32_t load_coherent_u(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up refit in two phases.  The hierarchy's caches are per XCD, so a
// device-scope fence writes back and invalidates L2; the phases keep those
// fences off the bulk of the nodes.
//   1. One workgroup per chunk of kRefitChunk Morton-consecutive points: a
//      thread per point climbs; the second arrival at a node whose range lies
//      inside the chunk builds it (arrival counters and child boxes in LDS,
//      workgroup-scope fences).  A Karras node's index is one end of its
//      range, so the chunk's nodes have indices inside the chunk.
//   2. The nodes whose ranges cross a chunk boundary (a few per boundary):
//      a thread starts at each such node without crossing children and
//      climbs with device-scope arrival counters.
constexpr int kRefitChunk = 1024;

__device__ __forceinline__ bool refit_crossing(uint2 r) { return r.x / kRefitChunk != r.y / kRefitChunk; }

// Builds node p's traversal record and AABB (max weight in lo.w) from its
// children; box_of(ch) gives an internal child's AABB.
template <typename BoxOf>
__device__ __forceinline__ void refit_node(uint32_t p, const float4 *__restrict__ pts, const uint32_t *__restrict__ children,
                                           const uint2 *__restrict__ ranges, int leaf_size, Node *__restrict__ nodes,
                                           BoxOf box_of, float4 &blo, float4 &bhi) {
    float4 lo[2], hi[2];
    uint32_t ref[2];
    for (int c = 0; c < 2; ++c) {
        const uint32_t ch = children[2 * p + c];
        uint32_t first, last;
        if (ch & 0x80000000u) {
            const uint32_t prim = ch & 0x7fffffffu;
            const float4 q = pts[prim];
            lo[c] = q;
            hi[c] = make_float4(q.x, q.y, q.z, 0.0f);
            first = last = prim;
        } else {
            box_of(ch, lo[c], hi[c]);
            const uint2 r = ranges[ch];
            first = r.x;
            last = r.y;
        }
        const uint32_t count = last - first + 1;
        ref[c] = (count <= (uint32_t)leaf_size) ? leaf_ref(first, count) : ch;
    }
    Node nd;
    nd.lo0 = lo[0];
    nd.hi0 = make_float4(hi[0].x, hi[0].y, hi[0].z, __uint_as_float(ref[0]));
    nd.lo1 = lo[1];
    nd.hi1 = make_float4(hi[1].x, hi[1].y, hi[1].z, __uint_as_float(ref[1]));
    nodes[p] = nd;
    blo = make_float4(fminf(lo[0].x, lo[1].x), fminf(lo[0].y, lo[1].y), fminf(lo[0].z, lo[1].z), fmaxf(lo[0].w, lo[1].w));
    bhi = make_float4(fmaxf(hi[0].x, hi[1].x), fmaxf(hi[0].y, hi[1].y), fmaxf(hi[0].z, hi[1].z), 0.0f);
}

__global__ void __launch_bounds__(kRefitChunk)
    refit_chunk_kernel(const float4 *__restrict__ pts, int64_t n, const uint32_t *__restrict__ children,
                       const uint32_t *__restrict__ parents, const uint2 *__restrict__ ranges, float4 *__restrict__ boxes,
                       Node *__restrict__ nodes, int leaf_size) {
    __shared__ float4 slo[kRefitChunk], shi[kRefitChunk];
    __shared__ uint32_t arrive[kRefitChunk];
    const int64_t c0 = (int64_t)blockIdx.x * kRefitChunk;
    const int64_t c1 = (c0 + kRefitChunk < n ? c0 + kRefitChunk : n) - 1;  // chunk points [c0, c1]
    arrive[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t t = c0 + threadIdx.x;
    if (t > c1) return;
    uint32_t p = parents[(n - 1) + t];
    auto box_of = [&](uint32_t ch, float4 &lo, float4 &hi) {
        lo = slo[ch - c0];
        hi = shi[ch - c0];
    };
    while (p != 0xFFFFFFFFu && (int64_t)p < n - 1) {
        const uint2 r = ranges[p];
        if ((int64_t)r.x < c0 || (int64_t)r.y > c1) return;  // crosses the chunk: phase 2
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (atomicAdd(&arrive[p - c0], 1u) == 0u) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        float4 blo, bhi;
        refit_node(p, pts, children, ranges, leaf_size, nodes, box_of, blo, bhi);
        slo[p - c0] = blo;
        shi[p - c0] = bhi;
        boxes[2 * (int64_t)p] = blo;
        boxes[2 * (int64_t)p + 1] = bhi;
        p = parents[p];
    }
}

__device__ __forceinline__ int refit_crossing_children(uint32_t p, const uint32_t *__restrict__ children,
                                                       const uint2 *__restrict__ ranges) {
    int k = 0;
    for (int c = 0; c < 2; ++c) {
        const uint32_t ch = children[2 * p + c];
        if (!(ch & 0x80000000u) && refit_crossing(ranges[ch])) k++;
    }
    return k;
}

__global__ void refit_top_kernel(const float4 *__restrict__ pts, int64_t n, const uint32_t *__restrict__ children,
                                 const uint32_t *__restrict__ parents, const uint2 *__restrict__ ranges,
                                 float4 *__restrict__ boxes, uint32_t *__restrict__ flags, Node *__restrict__ nodes,
                                 int leaf_size) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (!refit_crossing(ranges[i]) || refit_crossing_children((uint32_t)i, children, ranges) != 0) return;
    auto box_of = [&](uint32_t ch, float4 &lo, float4 &hi) {
        const float *b = (const float *)&boxes[2 * (int64_t)ch];
        lo = make_float4(load_coherent(b + 0), load_coherent(b + 1), load_coherent(b + 2), load_coherent(b + 3));
        hi = make_float4(load_coherent(b + 4), load_coherent(b + 5), load_coherent(b + 6), 0.0f);
    };
    uint32_t p = (uint32_t)i;
    // a crossing path is at most as deep as the hierarchy (< 64 levels)
    for (int level = 0; level < 128; ++level) {
        float4 blo, bhi;
        refit_node(p, pts, children, ranges, leaf_size, nodes, box_of, blo, bhi);
        boxes[2 * (int64_t)p] = blo;
        boxes[2 * (int64_t)p + 1] = bhi;
        const uint32_t q = parents[p];
        if (q == 0xFFFFFFFFu || (int64_t)q >= n - 1) return;
        if (refit_crossing_children(q, children, ranges) == 2) {  // the second arrival builds q
            __threadfence();
            if (atomicAdd(&flags[q], 1u) == 0u) return;
            __threadfence();
        }
        p = q;
    }
}

// Points identical in position and weight have identical Morton keys, so
// each duplicate group lies inside one run of equal keys: scan the run.
// Groups of exact duplicates (equal position and weight) among the
// Morton-sorted points, for the rows (every duplicate of a face owner is a
// neighbour): head[i] is the smallest sorted index of i's group and next[]
// chains every member from the head (0xFFFFFFFF ends a chain).  Duplicates
// share their Morton key, so only equal-key runs need a look.  A run of at
// most kDupWindow points is scanned by each member (the chain then runs in
// sorted order); the members of a longer run (a cluster denser than the
// Morton grid, or many coincident points) go through a hash table instead,
// so the work stays linear whatever the run length.
constexpr int kDupWindow = 32;
constexpr uint32_t kDupLong = 0xFFFFFFFEu;  // head[] mark of a member of a long run
constexpr uint32_t kDupEnd = 0xFFFFFFFFu;

__device__ __forceinline__ bool same_point(float4 a, float4 b) {
    return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
}
// Hash consistent with same_point (-0 and +0 compare equal, so both hash as +0).
__device__ __forceinline__ uint32_t point_hash(float4 p) {
    const float v[4] = {p.x == 0.0f ? 0.0f : p.x, p.y == 0.0f ? 0.0f : p.y, p.z == 0.0f ? 0.0f : p.z,
                        p.w == 0.0f ? 0.0f : p.w};
    uint32_t h = 0x9E3779B9u;
    for (int k = 0; k < 4; ++k) {
        h = (h ^ __float_as_uint(v[k])) * 0x85EBCA6Bu;
        h ^= h >> 13;
    }
    return h * 0xC2B2AE35u;
}
__device__ __forceinline__ uint32_t dup_slot(uint32_t h, uint32_t table_size) {
    return (uint32_t)(((uint64_t)h * table_size) >> 32);
}

__global__ void dup_runs_kernel(const uint32_t *__restrict__ keys, const float4 *__restrict__ pts, int64_t n,
                                uint32_t *__restrict__ head, uint32_t *__restrict__ next,
                                IndexHeader *__restrict__ hdr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    // the run [s, e] of key k around i, as far as the window reaches; every
    // member of a run reaches the same verdict: longer than the window or not
    int64_t s = i, e = i;
    for (int t = 0; t < kDupWindow && s > 0 && keys[s - 1] == k; ++t) --s;
    bool lng = s > 0 && keys[s - 1] == k;
    for (int t = 0; t < kDupWindow && !lng && e + 1 < n && keys[e + 1] == k; ++t) ++e;
    lng = lng || (e + 1 < n && keys[e + 1] == k) || e - s + 1 > kDupWindow;
    if (lng) {
        head[i] = kDupLong;
        next[i] = kDupEnd;
        return;
    }
    const float4 p = pts[i];
    int64_t h = i;
    for (int64_t j = s; j < i; ++j) {
        if (same_point(pts[j], p)) {
            h = j;
            break;
        }
    }
    uint32_t nx = kDupEnd;
    for (int64_t j = i + 1; j <= e; ++j) {
        if (same_point(pts[j], p)) {
            nx = (uint32_t)j;
            break;
        }
    }
    head[i] = (uint32_t)h;
    next[i] = nx;
    if (h != i || nx != kDupEnd) hdr->has_duplicates = 1u;
}

// Members of long runs: the table (table_size = 2n slots, all kDupEnd) maps
// each distinct point to the smallest index of its group; linear probing,
// and the table never holds more than n entries, so a probe ends.
__global__ void dup_insert_kernel(const float4 *__restrict__ pts, int64_t n, const uint32_t *__restrict__ head,
                                  uint32_t *__restrict__ table, uint32_t table_size) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || head[i] != kDupLong) return;
    const float4 p = pts[i];
    uint32_t sl = dup_slot(point_hash(p), table_size);
    for (uint32_t probe = 0; probe < table_size; ++probe) {
        uint32_t v = table[sl];
        if (v == kDupEnd) {
            v = atomicCAS(&table[sl], kDupEnd, (uint32_t)i);
            if (v == kDupEnd) return;
        }
        // any member identifies the group; keep the smallest index
        if (same_point(pts[v], p)) {
            atomicMin(&table[sl], (uint32_t)i);
            return;
        }
        sl = sl + 1 == table_size ? 0 : sl + 1;
    }
}

// Members of long runs: head = the group's smallest index; every other
// member links itself in right after the head (next[head] was set to the
// end by dup_runs_kernel).
__global__ void dup_link_kernel(const float4 *__restrict__ pts, int64_t n, uint32_t *__restrict__ head,
                                uint32_t *__restrict__ next, const uint32_t *__restrict__ table, uint32_t table_size,
                                IndexHeader *__restrict__ hdr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || head[i] != kDupLong) return;
    const float4 p = pts[i];
    uint32_t sl = dup_slot(point_hash(p), table_size);
    uint32_t h = (uint32_t)i;
    for (uint32_t probe = 0; probe < table_size; ++probe) {
        const uint32_t v = table[sl];
        if (v == kDupEnd) break;  // (not reached: i itself was inserted)
        if (same_point(pts[v], p)) {
            h = v;
            break;
        }
        sl = sl + 1 == table_size ? 0 : sl + 1;
    }
    head[i] = h;
    if (h != (uint32_t)i) {
        next[i] = atomicExch(&next[h], (uint32_t)i);
        hdr->has_duplicates = 1u;
    }
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Leaf statistics of the hierarchy (the debug print of pwr_bvh.cu:342-453):
// out = {leaves, fewest points in a leaf, most, points in all leaves}.
__global__ void leaf_stats_init_kernel(unsigned long long *out) {
    if (threadIdx.x < 4) out[threadIdx.x] = threadIdx.x == 1 ? ~0ull : 0ull;
}
// A node whose range holds at most leaf_size points was collapsed into a
// leaf of its parent (its record is written but never visited); ranges grow
// towards the root, so a node is visited iff its range is larger.
__global__ void leaf_stats_kernel(const Node *__restrict__ nodes, const uint2 *__restrict__ ranges, int64_t n_internal,
                                  const IndexHeader *__restrict__ hdr, unsigned long long *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t refs[2] = {0u, 0u};
    int m = 0;
    if (i < n_internal && ranges[i].y - ranges[i].x + 1u > hdr->leaf_size) {
        refs[m++] = __float_as_uint(nodes[i].hi0.w);
        refs[m++] = __float_as_uint(nodes[i].hi1.w);
    } else if (i == 0 && is_leaf(hdr->root_ref)) {
        refs[m++] = hdr->root_ref;  // no more points than a leaf holds: the root is a leaf
    }
    for (int k = 0; k < m; ++k) {
        if (!is_leaf(refs[k])) continue;
        const unsigned long long c = leaf_count(refs[k]);
        atomicAdd(&out[0], 1ull);
        atomicMin(&out[1], c);
        atomicMax(&out[2], c);
        atomicAdd(&out[3], c);
    }
}

// ------------------------------------------------------------------ host
static inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

hipError_t fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

// Sort, hierarchy and refit over keys already written to the workspace.
static hipError_t build_tail(int64_t n, const float *points, const float *weights, int leaf_size, char *base,
                             const IndexLayout &L, hipStream_t stream) {
    IndexHeader *hdr = (IndexHeader *)(base + L.header);
    float4 *sorted = (float4 *)(base + L.points);
    uint32_t *perm = (uint32_t *)(base + L.perm);
    uint32_t *inv = (uint32_t *)(base + L.inv);
    Node *nodes = (Node *)(base + L.nodes);
    uint32_t *keys = (uint32_t *)(base + L.keys);
    uint32_t *keys_alt = (uint32_t *)(base + L.keys_alt);
    uint32_t *vals_alt = (uint32_t *)(base + L.vals_alt);
    uint32_t *hist = (uint32_t *)(base + L.hist);
    uint32_t *children = (uint32_t *)(base + L.children);
    uint32_t *parents = (uint32_t *)(base + L.parents);
    uint2 *ranges = (uint2 *)(base + L.ranges);
    float4 *boxes = (float4 *)(base + L.boxes);
    uint32_t *flags = (uint32_t *)(base + L.flags);
    uint32_t *scan_tmp = (uint32_t *)(base + L.scan_tmp);
    uint32_t *dup_head = (uint32_t *)(base + L.dup_head);
    uint32_t *dup_next = (uint32_t *)(base + L.dup_next);

    // 4 stable passes of 8 bits: keys -> alt -> keys -> alt -> keys
    const int nb = L.sort_blocks;
    const uint32_t *kin = keys;
    const uint32_t *vin = nullptr;
    for (int pass = 0; pass < 4; ++pass) {
        uint32_t *kout = (pass % 2 == 0) ? keys_alt : keys;
        uint32_t *vout = (pass % 2 == 0) ? vals_alt : perm;
        const int shift = 8 * pass;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(kSortThreads), 0, stream, kin, n, shift, hist, nb);
        hipError_t e = exclusive_scan_u32(hist, hist, (int64_t)256 * nb, scan_tmp, hist + (int64_t)256 * nb, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(kSortThreads), 0, stream, kin, vin, n, shift, hist, nb,
                           kout, vout);
        kin = kout;
        vin = vout;
    }
    hipLaunchKernelGGL(reorder_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, points, weights, n, perm, sorted,
                       inv);
    const uint32_t table_size = (uint32_t)(2 * n);
    uint32_t *dup_table = (uint32_t *)(base + L.dup_table);
    hipLaunchKernelGGL(dup_runs_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, keys, sorted, n, dup_head,
                       dup_next, hdr);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(table_size, 256)), dim3(256), 0, stream, dup_table,
                       (int64_t)table_size, kDupEnd);
    hipLaunchKernelGGL(dup_insert_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, sorted, n, dup_head,
                       dup_table, table_size);
    hipLaunchKernelGGL(dup_link_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, sorted, n, dup_head, dup_next,
                       dup_table, table_size, hdr);
    if (n >= 2) {
        hipLaunchKernelGGL(karras_kernel, dim3(grid_for(n - 1, 256)), dim3(256), 0, stream, keys, n, children, parents,
                           ranges);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n - 1, 256)), dim3(256), 0, stream, flags, n - 1, 0u);
        hipLaunchKernelGGL(refit_chunk_kernel, dim3(grid_for(n, kRefitChunk)), dim3(kRefitChunk), 0, stream, sorted, n,
                           children, parents, ranges, boxes, nodes, leaf_size);
        if (n > kRefitChunk)
            hipLaunchKernelGGL(refit_top_kernel, dim3(grid_for(n - 1, 256)), dim3(256), 0, stream, sorted, n, children,
                               parents, ranges, boxes, flags, nodes, leaf_size);
    }
    return hipGetLastError();
}


hipError_t leaf_stats(int64_t n, const void *ws, unsigned long long *out, hipStream_t stream) {
    const IndexLayout L = make_index_layout(n);
    const char *base = (const char *)ws;
    const int64_t ni = n >= 2 ? n - 1 : 0;
    hipLaunchKernelGGL(leaf_stats_init_kernel, dim3(1), dim3(64), 0, stream, out);
    hipLaunchKernelGGL(leaf_stats_kernel, dim3(grid_for(ni > 0 ? ni : 1, 256)), dim3(256), 0, stream,
                       (const Node *)(base + L.nodes), (const uint2 *)(base + L.ranges), ni,
                       (const IndexHeader *)(base + L.header), out);
    return hipGetLastError();
}

hipError_t build_index(int64_t n, const float *points, const float *weights, int leaf_size, void *ws,
                       hipStream_t stream) {
    const IndexLayout L = make_index_layout(n);
    char *base = (char *)ws;
    IndexHeader *hdr = (IndexHeader *)(base + L.header);
    uint32_t *keys = (uint32_t *)(base + L.keys);
    float *reduce = (float *)(base + L.reduce);
    hipLaunchKernelGGL(bbox_partial_kernel, dim3(L.reduce_blocks), dim3(kReduceThreads), 0, stream, points, n, reduce);
    hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(kReduceThreads), 0, stream, reduce, L.reduce_blocks, hdr, n,
                       weights ? 1 : 0, leaf_size);
    hipLaunchKernelGGL(morton_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, points, n, hdr, keys);

    return build_tail(n, points, weights, leaf_size, base, L, stream);
}

// ------------------------------------------------------------------ batches
// Index over a batch of clouds (contiguous input ranges cloud_offsets[c] ..
// cloud_offsets[c + 1]): per-cloud AABB and clipping box, keys cloud id ||
// per-cloud Morton code, one sort and hierarchy for all points.  A cloud's
// points then occupy its own input range in sorted order, under one subtree.

__device__ __forceinline__ int cloud_of(const int64_t *__restrict__ off, int n_clouds, int64_t i) {
    int lo = 0, hi = n_clouds - 1;  // largest c with off[c] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kReduceThreads)
    batch_bbox_kernel(const float *__restrict__ pts, const int64_t *__restrict__ off, int n_clouds, int has_w,
                      int leaf_size, IndexHeader *__restrict__ clouds, IndexHeader *__restrict__ hdr, int64_t n) {
    const int c = blockIdx.x;
    const int64_t a = off[c], b = off[c + 1];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            const float v = pts[3 * i + k];
            mn[k] = fminf(mn[k], v);
            mx[k] = fmaxf(mx[k], v);
        }
    }
    __shared__ float red[6][kReduceThreads];
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = mn[k];
        red[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        IndexHeader h;
        for (int k = 0; k < 3; ++k) {
            const float lo = b > a ? red[k][0] : 0.0f, hi = b > a ? red[3 + k][0] : 0.0f;
            h.raw_lo[k] = lo;
            h.raw_hi[k] = hi;
            // the single-cloud box, in the same FP32 order (voronoi_ultra.cu:237-238,
            // voronoi_convex_cell.cuh:187-193)
            h.box_lo[k] = __fsub_rn(__fsub_rn(lo, 1.0f), 1e-5f);
            h.box_hi[k] = __fadd_rn(__fadd_rn(hi, 1.0f), 1e-5f);
        }
        h.root_ref = 0xFFFFFFFFu;  // set after the hierarchy is built
        h.n = (uint32_t)(b - a);
        h.has_weights = (uint32_t)has_w;
        h.leaf_size = (uint32_t)leaf_size;
        h.has_duplicates = 0u;
        clouds[c] = h;
        if (c == 0) {
            h.n = (uint32_t)n;
            hdr[0] = h;  // batch-wide fields (duplicates flag, sizes); its box is unused
        }
    }
}

__global__ void batch_morton_kernel(const float *__restrict__ pts, int64_t n, const int64_t *__restrict__ off,
                                    int n_clouds, const IndexHeader *__restrict__ clouds, int cloud_bits,
                                    int axis_bits, uint32_t *__restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cloud_of(off, n_clouds, i);
    const IndexHeader &h = clouds[c];
    const float ex = fmaxf(fmaxf(h.raw_hi[0] - h.raw_lo[0], h.raw_hi[1] - h.raw_lo[1]), h.raw_hi[2] - h.raw_lo[2]);
    const float cells = (float)(1u << axis_bits);
    const float scale = ex > 0.0f ? cells / ex : 0.0f;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float t = (pts[3 * i + k] - h.raw_lo[k]) * scale;
        t = fminf(fmaxf(t, 0.0f), cells - 1.0f);
        q[k] = (uint32_t)t;
    }
    const uint32_t m = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    keys[i] = cloud_bits > 0 ? (((uint32_t)c << (32 - cloud_bits)) | m) : m;
}

// Root of each cloud's subtree: the node whose range is exactly the cloud's
// sorted range (it exists: the cloud id is the keys' leading bits).
__global__ void batch_roots_kernel(int64_t n, const int64_t *__restrict__ off, int n_clouds,
                                   const uint32_t *__restrict__ parents, const uint2 *__restrict__ ranges,
                                   int leaf_size, const IndexHeader *__restrict__ hdr,
                                   IndexHeader *__restrict__ clouds) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_clouds) return;
    const int64_t a = off[c], cnt = off[c + 1] - a;
    uint32_t root = 0xFFFFFFFFu;
    if (cnt > 0 && cnt <= leaf_size) {
        root = leaf_ref((uint32_t)a, (uint32_t)cnt);
    } else if (cnt > 0) {
        uint32_t p = parents[(n - 1) + a];
        root = 0u;  // never kept: the walk meets the cloud's node (a valid node either way)
        for (int level = 0; level < 128 && p != 0xFFFFFFFFu; ++level) {
            const uint2 r = ranges[p];
            if ((int64_t)r.x == a && (int64_t)r.y == a + cnt - 1) {
                root = p;
                break;
            }
            p = parents[p];
        }
    }
    clouds[c].root_ref = root;
    clouds[c].has_duplicates = hdr->has_duplicates;
}

hipError_t build_batch_index(int64_t n, int n_clouds, const float *points, const float *weights,
                             const int64_t *cloud_offsets, int leaf_size, void *ws, hipStream_t stream) {
    const BatchLayout B = make_batch_layout(n, n_clouds);
    const IndexLayout &L = B.idx;
    char *base = (char *)ws;
    IndexHeader *hdr = (IndexHeader *)(base + L.header);
    IndexHeader *clouds = (IndexHeader *)(base + B.clouds);
    uint32_t *keys = (uint32_t *)(base + L.keys);
    hipLaunchKernelGGL(batch_bbox_kernel, dim3(n_clouds), dim3(kReduceThreads), 0, stream, points, cloud_offsets,
                       n_clouds, weights ? 1 : 0, leaf_size, clouds, hdr, n);
    hipLaunchKernelGGL(batch_morton_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, points, n, cloud_offsets,
                       n_clouds, clouds, batch_cloud_bits(n_clouds), batch_axis_bits(n_clouds), keys);
    if (hipError_t e = build_tail(n, points, weights, leaf_size, base, L, stream)) return e;
    hipLaunchKernelGGL(batch_roots_kernel, dim3(grid_for(n_clouds, 256)), dim3(256), 0, stream, n, cloud_offsets,
                       n_clouds, (const uint32_t *)(base + L.parents), (const uint2 *)(base + L.ranges), leaf_size, hdr,
                       clouds);
    return hipGetLastError();
}

}  // namespace pg
