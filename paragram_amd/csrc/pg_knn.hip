// k nearest neighbours of every point over the LBVH (the reference's
// cuBQL_knn_kernel, pwr_bvh.cu:159-215: the K nearest by squared distance,
// the query's own id excluded, short lists padded with the last id found or
// 0).  Used for the learned radius features and as initial guesses.
//
// One lane per query in Morton order (neighbouring lanes walk neighbouring
// parts of the hierarchy).  Depth-first, nearer child first; the farther
// child waits on a per-lane stack in LDS as (parent << 1 | child) and is
// re-tested against the current K-th distance when it comes out.  A Karras
// hierarchy over 32-bit keys and < 2^26 points is less than 64 levels deep,
// and every level pushes at most one entry, so 64 entries never overflow.
// The first pass keeps 32 (half the LDS, so more waves per CU); the rare
// query that needs more is queued in the index workspace and redone by a
// second pass with 64.
#include "pg_common.hpp"
#include "pg_kernels.hpp"

#include <float.h>

namespace pg {

constexpr int kKnnThreads = 128;
constexpr int kKnnStack = 64;      // deep pass: never overflows (see above)
constexpr int kKnnStackFast = 32;  // first pass: 16 KB of LDS per block instead of 32 KB
#ifndef PG_KNN_LOADS
#define PG_KNN_LOADS 2
#endif
constexpr int kLoads = PG_KNN_LOADS;  // leaf points loaded ahead

__device__ __forceinline__ float box_dist2(float4 lo, float4 hi, float qx, float qy, float qz) {
    const float dx = fmaxf(fmaxf(lo.x - qx, qx - hi.x), 0.0f);
    const float dy = fmaxf(fmaxf(lo.y - qy, qy - hi.y), 0.0f);
    const float dz = fmaxf(fmaxf(lo.z - qz, qz - hi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// S: stack entries.  The first pass (kDeep false) runs every query with a
// 32-entry stack; a query whose path needs more appends itself to `spill`
// and the deep pass (kDeep, 64 entries) redoes it, so no query's result
// depends on the stack size.
template <int K, int S, bool kDeep>
__device__ __forceinline__ void knn_body(int64_t n, IndexHeader *__restrict__ hdr, const float4 *__restrict__ pts,
                                         const uint32_t *__restrict__ perm, const Node *__restrict__ nodes, int k_out,
                                         int32_t *__restrict__ out_ids, float *__restrict__ out_d2,
                                         const IndexHeader *__restrict__ clouds, const int64_t *__restrict__ cloud_cells,
                                         int n_clouds, int sorted_rows, uint32_t *__restrict__ spill) {
    __shared__ uint32_t stack[S][kKnnThreads];
    int64_t i;
    if constexpr (kDeep) {
        const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= (int64_t)hdr->knn_spill) return;
        i = spill[t];
    } else {
        i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n) return;
    }
    const float4 q = pts[i];
    float d[K];
    int32_t id[K];
#pragma unroll
    for (int t = 0; t < K; ++t) d[t] = FLT_MAX, id[t] = -1;
    // a node is pruned when its box is farther than the K-th distance (with
    // a relative margin for the rounding of the two squared distances)
    auto far = [&](float bd) { return bd * (1.0f - 1e-6f) > d[K - 1]; };
    auto offer = [&](const float4 &p, uint32_t j) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float dd = dx * dx + dy * dy + dz * dz;
        if (dd < d[K - 1]) {  // sorted insertion, ties keep the earlier point
            d[K - 1] = dd;
            id[K - 1] = (int32_t)j;
#pragma unroll
            for (int u = K - 1; u > 0; --u) {
                if (d[u] < d[u - 1]) {
                    const float td = d[u];
                    d[u] = d[u - 1];
                    d[u - 1] = td;
                    const int32_t ti = id[u];
                    id[u] = id[u - 1];
                    id[u - 1] = ti;
                }
            }
        }
    };
    // a batch searches the query's own cloud (its subtree)
    uint32_t ref = hdr->root_ref;
    if (clouds != nullptr) {
        int lo = 0, hi = n_clouds - 1;  // largest c with cloud_cells[c] <= i
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cloud_cells[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        ref = clouds[lo].root_ref;
    }
    int sp = 0;
    bool spilled = false;
    while (true) {
        if (is_leaf(ref)) {
            const uint32_t f = leaf_first(ref), c = leaf_count(ref);
            // kLoads loads in flight, then as many offers (leaf order kept)
            for (uint32_t t = 0; t < c; t += kLoads) {
                float4 p[kLoads];
#pragma unroll
                for (int u = 0; u < kLoads; ++u) p[u] = pts[f + min(t + (uint32_t)u, c - 1u)];
#pragma unroll
                for (int u = 0; u < kLoads; ++u) {
                    const uint32_t j = f + t + (uint32_t)u;
                    if (t + (uint32_t)u < c && (int64_t)j != i) offer(p[u], j);
                }
            }
        } else {
            const Node nd = nodes[ref];
            const float b0 = box_dist2(nd.lo0, nd.hi0, q.x, q.y, q.z);
            const float b1 = box_dist2(nd.lo1, nd.hi1, q.x, q.y, q.z);
            const bool near0 = b0 <= b1;
            const float bn = near0 ? b0 : b1, bf = near0 ? b1 : b0;
            if (!far(bf)) {
                if (sp == S) {
                    spilled = true;
                    break;
                }
                stack[sp++][threadIdx.x] = (ref << 1) | (near0 ? 1u : 0u);
            }
            if (!far(bn)) {
                ref = __float_as_uint(near0 ? nd.hi0.w : nd.hi1.w);
                continue;
            }
        }
        bool have = false;
        while (sp > 0 && !have) {
            const uint32_t e = stack[--sp][threadIdx.x];
            const Node nd = nodes[e >> 1];
            const float4 lo = (e & 1u) ? nd.lo1 : nd.lo0, hi = (e & 1u) ? nd.hi1 : nd.hi0;
            if (!far(box_dist2(lo, hi, q.x, q.y, q.z))) {
                ref = __float_as_uint(hi.w);
                have = true;
            }
        }
        if (!have) break;
    }
    if (!kDeep && spilled) {
        spill[atomicAdd(&hdr->knn_spill, 1u)] = (uint32_t)i;
        return;
    }
    // short lists repeat the last id found; with none found the reference
    // writes 0, a batch writes the point itself (0 may be another cloud's).
    // sorted_rows: row i (sorted order) of sorted ids, self as padding, and
    // (out_d2 non-null) out_d2[i] = the squared distance of the row's last
    // neighbour, FLT_MAX when the row is short (it then holds every other
    // point the search could reach).
    if (sorted_rows) {
        const int64_t o = i * k_out;
        int32_t last = (int32_t)i;
        float dk = FLT_MAX;
        for (int t = 0; t < k_out; ++t) {
            if (id[t] >= 0) last = id[t];
            out_ids[o + t] = last;
            dk = id[t] >= 0 ? d[t] : FLT_MAX;
        }
        if (out_d2) out_d2[i] = dk;
        return;
    }
    const int64_t o = (int64_t)perm[i] * k_out;
    int32_t last = clouds != nullptr ? (int32_t)perm[i] : 0;
    for (int t = 0; t < k_out; ++t) {
        if (id[t] >= 0) last = (int32_t)perm[id[t]];
        out_ids[o + t] = last;
        if (out_d2) out_d2[o + t] = id[t] >= 0 ? d[t] : FLT_MAX;
    }
}

#define PG_KNN_ARGS                                                                                                 \
    int64_t n, IndexHeader *__restrict__ hdr, const float4 *__restrict__ pts, const uint32_t *__restrict__ perm,    \
        const Node *__restrict__ nodes, int k_out, int32_t *__restrict__ out_ids, float *__restrict__ out_d2,       \
        const IndexHeader *__restrict__ clouds, const int64_t *__restrict__ cloud_cells, int n_clouds,               \
        int sorted_rows, uint32_t *__restrict__ spill
// first pass: 128 VGPRs and 16 KB of LDS per block, four waves per SIMD
template <int K>
__global__ void __launch_bounds__(kKnnThreads) __attribute__((amdgpu_waves_per_eu(4, 4)))
    knn_kernel(PG_KNN_ARGS) {
    knn_body<K, kKnnStackFast, false>(n, hdr, pts, perm, nodes, k_out, out_ids, out_d2, clouds, cloud_cells, n_clouds,
                                      sorted_rows, spill);
}
template <int K>
__global__ void __launch_bounds__(kKnnThreads) knn_deep_kernel(PG_KNN_ARGS) {
    knn_body<K, kKnnStack, true>(n, hdr, pts, perm, nodes, k_out, out_ids, out_d2, clouds, cloud_cells, n_clouds,
                                 sorted_rows, spill);
}
#undef PG_KNN_ARGS

__global__ void knn_reset_kernel(IndexHeader *hdr) { hdr->knn_spill = 0u; }

template <int K>
static void launch_knn(int64_t n, IndexHeader *hdr, const float4 *pts, const uint32_t *perm, const Node *nodes, int k,
                       int32_t *ids, float *d2, const IndexHeader *clouds, const int64_t *cloud_cells, int n_clouds,
                       int sorted, uint32_t *spill, hipStream_t stream) {
    const unsigned grid = (unsigned)((n + kKnnThreads - 1) / kKnnThreads);
    hipLaunchKernelGGL(knn_reset_kernel, dim3(1), dim3(1), 0, stream, hdr);
    hipLaunchKernelGGL(knn_kernel<K>, dim3(grid), dim3(kKnnThreads), 0, stream, n, hdr, pts, perm, nodes, k, ids, d2,
                       clouds, cloud_cells, n_clouds, sorted, spill);
    // the deep pass reads the spill count on the device; its grid covers
    // every query in case all of them spilled (blocks past the count exit)
    hipLaunchKernelGGL(knn_deep_kernel<K>, dim3(grid), dim3(kKnnThreads), 0, stream, n, hdr, pts, perm, nodes, k, ids,
                       d2, clouds, cloud_cells, n_clouds, sorted, spill);
}

hipError_t knn(int64_t n, void *index_ws, int k, int32_t *ids, float *d2, hipStream_t stream, int n_clouds,
               bool sorted_rows) {
    if (n <= 0) return hipSuccess;
    const IndexLayout L = make_index_layout(n);
    char *ws = (char *)index_ws;
    const IndexHeader *clouds = nullptr;
    const int64_t *cloud_cells = nullptr;
    if (n_clouds > 0) {
        const BatchLayout B = make_batch_layout(n, n_clouds);
        clouds = (const IndexHeader *)(ws + B.clouds);
        cloud_cells = (const int64_t *)(ws + B.offsets);
    }
    IndexHeader *hdr = (IndexHeader *)(ws + L.header);
    const float4 *pts = (const float4 *)(ws + L.points);
    const uint32_t *perm = (const uint32_t *)(ws + L.perm);
    const Node *nodes = (const Node *)(ws + L.nodes);
    uint32_t *spill = (uint32_t *)(ws + L.vals_alt);
    if (k <= 8)
        launch_knn<8>(n, hdr, pts, perm, nodes, k, ids, d2, clouds, cloud_cells, n_clouds, sorted_rows ? 1 : 0, spill,
                      stream);
    else
        launch_knn<16>(n, hdr, pts, perm, nodes, k, ids, d2, clouds, cloud_cells, n_clouds, sorted_rows ? 1 : 0, spill,
                       stream);
    return hipGetLastError();
}

}  // namespace pg
