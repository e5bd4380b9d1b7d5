// CSR assembly for gfx950: device-wide exclusive scan, row emission in
// original point order, the symmetry check, and the union symmetrisation.
//
// Replaces gather_adjacency_kernel, count_symmetrized_adjacency_sizes and
// populate_symmetric_adjacency (adjacency.cuh:7-159) and the torch
// cat + cumsum of voronoi_ultra.cu:347-356.  Rows come out sorted ascending,
// so the output does not depend on thread scheduling (the reference's
// atomic write heads at adjacency.cuh:27,148,155 make its row order vary).
#include "pg_common.hpp"
#include "pg_kernels.hpp"

namespace pg {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

// Exclusive scan of one value per thread over a block of kThreads threads.
template <int kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *smem, uint32_t &total) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint32_t x = v;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) smem[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < kThreads / kWave; ++w) {
            const uint32_t t = smem[w];
            smem[w] = run;
            run += t;
        }
        smem[kThreads / kWave] = run;
    }
    __syncthreads();
    const uint32_t res = smem[wave] + x - v;
    total = smem[kThreads / kWave];
    __syncthreads();
    return res;
}

__global__ void __launch_bounds__(kScanThreads) scan_reduce_kernel(const uint32_t *__restrict__ in, int64_t n,
                                                                   uint32_t *__restrict__ sums) {
    __shared__ uint32_t smem[kScanThreads / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_exclusive_scan<kScanThreads>(s, smem, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) scan_sums_kernel(uint32_t *__restrict__ sums, int64_t nb,
                                                         uint32_t *__restrict__ total) {
    __shared__ uint32_t smem[1024 / kWave + 1];
    uint32_t carry = 0;
    for (int64_t b = 0; b < nb; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t t;
        const uint32_t ex = block_exclusive_scan<1024>(v, smem, t);
        if (i < nb) sums[i] = carry + ex;
        carry += t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(kScanThreads) scan_final_kernel(const uint32_t *__restrict__ in,
                                                                  uint32_t *__restrict__ out, int64_t n,
                                                                  const uint32_t *__restrict__ sums) {
    __shared__ uint32_t smem[kScanThreads / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<kScanThreads>(s, smem, total) + sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, uint32_t *total,
                              hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, in, n, tmp);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, stream, tmp, nb, total);
    hipLaunchKernelGGL(scan_final_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, in, out, n, tmp);
    return hipGetLastError();
}

// ------------------------------------------------------------------ rows
__device__ __forceinline__ int64_t src_off(const int64_t *src_offsets, int64_t stride, uint32_t i) {
    return src_offsets ? src_offsets[i] : (int64_t)i * stride;
}

// counts of a sorted range in sorted order (for per-chunk packing)
__global__ void chunk_counts_kernel(const int32_t *__restrict__ counts, const uint32_t *__restrict__ perm,
                                    int64_t begin, int64_t end, uint32_t *__restrict__ out) {
    const int64_t i = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    out[i - begin] = (uint32_t)counts[perm[i]];
}

__global__ void pack_rows_kernel(int64_t begin, int64_t end, const uint32_t *__restrict__ perm,
                                 const int32_t *__restrict__ counts, const uint32_t *__restrict__ local_offsets,
                                 const int32_t *__restrict__ rows, const int64_t *__restrict__ row_offsets,
                                 int32_t *__restrict__ packed) {
    const int64_t i = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const int c = counts[perm[i]];
    const int32_t *src = rows + row_offsets[i];
    int32_t *dst = packed + local_offsets[i - begin];
    for (int k = 0; k < c; ++k) dst[k] = src[k];
}

// One thread per output row (original order).
__global__ void emit_rows_kernel(int64_t n, const uint32_t *__restrict__ inv, const int32_t *__restrict__ counts,
                                 const int32_t *__restrict__ offsets, const int32_t *__restrict__ src,
                                 const int64_t *__restrict__ src_offsets, int64_t src_stride,
                                 int32_t *__restrict__ adj) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int c = counts[j];
    if (c <= 0) return;
    const int32_t *s = src + src_off(src_offsets, src_stride, inv[j]);
    int32_t *d = adj + offsets[j];
    for (int k = 0; k < c; ++k) d[k] = s[k];
}

__device__ __forceinline__ bool row_contains(const int32_t *row, int c, int32_t x) {
    int lo = 0, hi = c;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < c && row[lo] == x;
}

// Counts directed edges u->v whose reverse v->u is missing, and entries that
// point at empty cells (adjacency.cuh:106-107), plus copies the edge total.
// One thread per cell in sorted (Morton) order: the rows a wave reads, its
// own and its neighbours', are Morton neighbours in the row buffer.
__global__ void check_symmetry_kernel(int64_t n, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ inv,
                                      const int32_t *__restrict__ counts, const int32_t *__restrict__ src,
                                      const int64_t *__restrict__ src_offsets, int64_t src_stride,
                                      const int32_t *__restrict__ offsets, unsigned long long *__restrict__ stats) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t asym = 0, dropped = 0;
    if (t < n) {
        const uint32_t u = perm[t];
        const int c = counts[u];
        const int32_t *ru = src + src_off(src_offsets, src_stride, (uint32_t)t);
        for (int k = 0; k < c; ++k) {
            const int32_t v = ru[k];
            const int cv = counts[v];
            if (cv <= 0) {
                dropped++;
                continue;
            }
            const int32_t *rv = src + src_off(src_offsets, src_stride, inv[v]);
            if (!row_contains(rv, cv, (int32_t)u)) asym++;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        asym += __shfl_xor(asym, o);
        dropped += __shfl_xor(dropped, o);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (asym) atomicAdd(&stats[kStatAsymmetric], (unsigned long long)asym);
        if (dropped) atomicAdd(&stats[kStatDropped], (unsigned long long)dropped);
    }
    if (t == 0) stats[15] = (unsigned long long)(uint32_t)offsets[n];
}

// Union symmetrisation on the emitted CSR (adjacency.cuh:78-159 semantics):
//   phase 0: keep[u] = entries v of row u with counts[v] > 0;
//            extra[v] += 1 for each such u->v whose reverse is missing.
//   phase 1: new_counts = keep + extra            (host scans -> new_offsets)
//   phase 2: write kept entries, append missing reverse entries via heads.
//   phase 3: sort every row (rows are short; insertion sort).
__global__ void union_phase0_kernel(int64_t n, const int32_t *__restrict__ counts, const int32_t *__restrict__ offsets,
                                    const int32_t *__restrict__ adj, int32_t *__restrict__ keep,
                                    uint32_t *__restrict__ extra) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int c = counts[u];
    const int32_t *ru = adj + offsets[u];
    int k0 = 0;
    for (int k = 0; k < c; ++k) {
        const int32_t v = ru[k];
        if (counts[v] <= 0) continue;
        k0++;
        if (!row_contains(adj + offsets[v], counts[v], (int32_t)u)) atomicAdd(&extra[v], 1u);
    }
    keep[u] = k0;
}
__global__ void union_phase1_kernel(int64_t n, const int32_t *__restrict__ keep, const uint32_t *__restrict__ extra,
                                    int32_t *__restrict__ new_counts) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    new_counts[u] = keep[u] + (int32_t)extra[u];
}
__global__ void union_phase2_kernel(int64_t n, const int32_t *__restrict__ counts, const int32_t *__restrict__ offsets,
                                    const int32_t *__restrict__ adj, const int32_t *__restrict__ keep,
                                    const int32_t *__restrict__ new_offsets, uint32_t *__restrict__ heads,
                                    int32_t *__restrict__ new_adj) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int c = counts[u];
    const int32_t *ru = adj + offsets[u];
    int32_t *dst = new_adj + new_offsets[u];
    int w = 0;
    for (int k = 0; k < c; ++k) {
        const int32_t v = ru[k];
        if (counts[v] <= 0) continue;
        dst[w++] = v;
        if (!row_contains(adj + offsets[v], counts[v], (int32_t)u)) {
            const uint32_t pos = atomicAdd(&heads[v], 1u);
            new_adj[(int64_t)new_offsets[v] + keep[v] + pos] = (int32_t)u;
        }
    }
}
__global__ void union_phase3_kernel(int64_t n, const int32_t *__restrict__ new_offsets, int32_t *__restrict__ new_adj) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    int32_t *r = new_adj + new_offsets[u];
    const int c = new_offsets[u + 1] - new_offsets[u];
    for (int i = 1; i < c; ++i) {
        const int32_t x = r[i];
        int k = i;
        while (k > 0 && r[k - 1] > x) {
            r[k] = r[k - 1];
            k--;
        }
        r[k] = x;
    }
}

// Cloud-local form of a batch's CSR (every neighbour lies in its row's
// cloud): ids minus the cloud's first point, and each cloud's offsets
// (n_c + 1 of them, from 0) laid out one cloud after another.  Offsets out
// of order only give wrong values: every index is clamped.
__device__ __forceinline__ int cloud_of(const int64_t *off, int n_clouds, int64_t u) {
    int lo = 0, hi = n_clouds - 1;  // largest c with off[c] <= u
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__global__ void batch_local_adj_kernel(int64_t n, int n_clouds, const int64_t *__restrict__ cloud_off,
                                       const int32_t *__restrict__ offsets, const int32_t *__restrict__ adj,
                                       int32_t *__restrict__ out) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int32_t s = (int32_t)cloud_off[cloud_of(cloud_off, n_clouds, u)];
    for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) out[e] = adj[e] - s;
}
__global__ void batch_local_offsets_kernel(int64_t n, int n_clouds, const int64_t *__restrict__ cloud_off,
                                           const int32_t *__restrict__ offsets, int32_t *__restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n + n_clouds) return;
    int lo = 0, hi = n_clouds - 1;  // largest c with off[c] + c <= p
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cloud_off[mid] + mid <= p) lo = mid;
        else hi = mid - 1;
    }
    int64_t u = p - lo, a = cloud_off[lo];
    u = u < 0 ? 0 : (u > n ? n : u);
    a = a < 0 ? 0 : (a > n ? n : a);
    out[p] = offsets[u] - offsets[a];
}

// Work quads of the cells [begin, end) of a batch from its cloud offsets on
// the device (pg_batch_quads's table, without a host copy): quads[c] = the
// first quad of cloud c, quads[n_clouds] = all quads.  One block: each
// thread sums a run of clouds, then the runs are scanned in LDS.
constexpr int kQuadThreads = 1024;
__global__ void __launch_bounds__(kQuadThreads) batch_quads_kernel(const int64_t *__restrict__ off, int n_clouds,
                                                                   int64_t begin, int64_t end,
                                                                   int64_t *__restrict__ quads) {
    __shared__ int64_t part[kQuadThreads];
    const int per = (n_clouds + kQuadThreads - 1) / kQuadThreads;
    const int c0 = threadIdx.x * per, c1 = min(c0 + per, n_clouds);
    auto quads_of = [&](int c) {
        const int64_t lo = off[c] > begin ? off[c] : begin;
        const int64_t hi = off[c + 1] < end ? off[c + 1] : end;
        return hi > lo ? (hi - lo + 3) / 4 : (int64_t)0;
    };
    int64_t sum = 0;
    for (int c = c0; c < c1; ++c) sum += quads_of(c);
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < kQuadThreads; d <<= 1) {  // inclusive scan of the runs
        const int64_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t q = part[threadIdx.x] - sum;
    for (int c = c0; c < c1; ++c) {
        quads[c] = q;
        q += quads_of(c);
    }
    if (threadIdx.x == kQuadThreads - 1) quads[n_clouds] = part[threadIdx.x];
}

static inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

hipError_t batch_quads(const int64_t *off, int n_clouds, int64_t begin, int64_t end, int64_t *quads,
                       hipStream_t stream) {
    hipLaunchKernelGGL(batch_quads_kernel, dim3(1), dim3(kQuadThreads), 0, stream, off, n_clouds, begin, end, quads);
    return hipGetLastError();
}

hipError_t batch_localize(int64_t n, int n_clouds, const int64_t *cloud_off, const int32_t *offsets,
                          const int32_t *adj, int32_t *local_adj, int32_t *local_offsets, hipStream_t stream) {
    if (n > 0)
        hipLaunchKernelGGL(batch_local_adj_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, n, n_clouds, cloud_off,
                           offsets, adj, local_adj);
    hipLaunchKernelGGL(batch_local_offsets_kernel, dim3(grid_for(n + n_clouds, 256)), dim3(256), 0, stream, n, n_clouds,
                       cloud_off, offsets, local_offsets);
    return hipGetLastError();
}

hipError_t gather_counts_sorted(const int32_t *counts, const uint32_t *perm, int64_t begin, int64_t end,
                                uint32_t *out, hipStream_t stream) {
    if (end <= begin) return hipSuccess;
    hipLaunchKernelGGL(chunk_counts_kernel, dim3(grid_for(end - begin, 256)), dim3(256), 0, stream, counts, perm,
                       begin, end, out);
    return hipGetLastError();
}

hipError_t pack_rows(int64_t begin, int64_t end, const uint32_t *perm, const int32_t *counts,
                     const uint32_t *local_offsets, const int32_t *rows, const int64_t *row_offsets,
                     int32_t *packed, hipStream_t stream) {
    if (end <= begin) return hipSuccess;
    hipLaunchKernelGGL(pack_rows_kernel, dim3(grid_for(end - begin, 256)), dim3(256), 0, stream, begin, end, perm,
                       counts, local_offsets, rows, row_offsets, packed);
    return hipGetLastError();
}

hipError_t emit_rows(int64_t n, const uint32_t *inv, const int32_t *counts, const int32_t *offsets,
                     const int32_t *src, const int64_t *src_offsets, int64_t src_stride, int32_t *adjacency,
                     hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(emit_rows_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, n, inv, counts, offsets, src,
                       src_offsets, src_stride, adjacency);
    return hipGetLastError();
}

hipError_t check_symmetry(int64_t n, const uint32_t *perm, const uint32_t *inv, const int32_t *counts,
                          const int32_t *src, const int64_t *src_offsets, int64_t src_stride, const int32_t *offsets,
                          unsigned long long *stats, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(check_symmetry_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, n, perm, inv, counts, src,
                       src_offsets, src_stride, offsets, stats);
    return hipGetLastError();
}

hipError_t symmetrize_union(int64_t n, const int32_t *counts, const int32_t *offsets, const int32_t *adj,
                            int32_t *keep, uint32_t *extra, int32_t *new_counts, const int32_t *new_offsets,
                            int32_t *new_adj, int phase, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const dim3 g(grid_for(n, 256)), b(256);
    switch (phase) {
    case 0:
        hipLaunchKernelGGL(union_phase0_kernel, g, b, 0, stream, n, counts, offsets, adj, keep, extra);
        break;
    case 1:
        hipLaunchKernelGGL(union_phase1_kernel, g, b, 0, stream, n, keep, extra, new_counts);
        break;
    case 2:
        hipLaunchKernelGGL(union_phase2_kernel, g, b, 0, stream, n, counts, offsets, adj, keep, new_offsets, extra,
                           new_adj);
        break;
    case 3:
        hipLaunchKernelGGL(union_phase3_kernel, g, b, 0, stream, n, new_offsets, new_adj);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pg
