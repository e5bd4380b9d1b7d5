// Cell construction for gfx950: shared definitions and the launcher.  The
// kernels (capacity tiers 0 and 1) are in pg_group_cell.hpp.
//
// Replaces compute_bvh_voronoi_kernel (voronoi_ultra.cu:39-141) and
// compute_bvh_power_diagram_kernel (laguerre_ultra.cu:43-174) together with
// their clipper (voronoi_convex_cell.cuh, faster_convex_cell.cuh) and
// traversal (voronoi_bvh.cuh:149-288, pwr_bvh.cuh:217-356).
//
// Geometry.  In seed-relative doubled coordinates y = 2 (x - s) a point p
// contributes the half-space a . y <= H with a = p - s and
// H = |a|^2 - (w_p - w_s) (the power bisector of faster_convex_cell.cuh:287-294
// scaled by 2|a|); box planes are a = -+e_k, H = 2 (s_k - lo_k), 2 (hi_k - s_k).
// A vertex is a triple (i, j, k) of plane slots ordered so that
// D = a_i . (a_j x a_k) > 0; its position is y = N / D with
// N = H_i (a_j x a_k) + H_j (a_k x a_i) + H_k (a_i x a_j).
// Plane q removes the vertex iff P = a_q . N - H_q D > 0 (strictly, as
// voronoi_convex_cell.cuh:256-260).
//
// Exactness.  Every quantity is a polynomial of depth < 16 in the rounded
// seed differences, so its FP32 error is at most gamma_16 (< 9.6e-7) times
// its permanent (the same expression over absolute values).  Vertex
// positions are kept with a rigorous error radius derived from those
// bounds; a removal test is decided in FP32 only when its value exceeds its
// error bound, and otherwise recomputed in FP64 from the FP32 inputs with the
// bound 1e-13 x permanent (gamma_24 < 2.7e-15); a test that FP64 cannot decide
// either is decided exactly by expansion arithmetic (pg_exact.hpp).
#include "pg_common.hpp"
#include "pg_exact.hpp"
#include "pg_kernels.hpp"

#include <float.h>

namespace pg {

constexpr float kG32 = 2e-6f;
constexpr float kTiny32 = 1e-30f;
constexpr double kG64 = 1e-13;
constexpr double kTiny64 = 1e-280;
constexpr float kSlack = 1e-6f;  // relative margin of the conservative prune tests

struct CellParams {
    int64_t n;
    const IndexHeader *hdr;
    const float4 *pts;      // Morton-sorted points (x, y, z, w)
    const uint32_t *perm;   // sorted -> original
    const uint32_t *inv;    // original -> sorted
    const Node *nodes;
    const uint32_t *dup_head;
    const uint32_t *dup_next;
    int64_t begin, end;
    const int32_t *guesses;
    const int32_t *guess_offsets;
    int guesses_sorted;
    int guess_stride;             // > 0: guesses are Morton-ordered rows of this many sorted ids
    const float *guess_d2;        // with guess_stride (may be null): squared distance of each row's last id
    int seed_window;              // Morton neighbours offered first (0: none)
    int32_t *rows;                // row pool (every tier bump-allocates its rows)
    uint64_t rows_elems;          // pool capacity
    uint32_t pool_chunk;          // pool entries a tier-0 wave takes at a time (0: exactly what a round needs)
    int64_t *row_offsets;         // [n] sorted cell -> row start in `rows`
    int32_t *counts;
    int32_t *status;
    unsigned long long *stats;
    uint32_t *counters;           // [0] tier-0 work, [1] tier-m queue length, [2] tier-m work,
                                  // [3] tier-2 queue length, [4] tier-2 work,
                                  // [5] cells flagged for FP64 tier 0, [6] FP64 tier-0 windows,
                                  // [7] removal tests recorded by tier 0 (taken by their FP32 sign),
                                  // [8] tier-1 queue length, [9] tier-1 work
    unsigned long long *pool_top; // row-pool allocation counter (in the scratch, zeroed per launch)
    uint32_t *mid_list;           // tier-m queue of sorted cell ids (room for every cell)
    uint32_t *big_list;           // tier-1 queue of sorted cell ids (room for every cell)
    uint32_t *huge_list;          // tier-2 queue of sorted cell ids (room for every cell)
    uint32_t queue_cap;           // end - begin
    uint32_t *exact_list;         // [end - begin] flags: cell for the FP64 tier-0 instance
    int32_t *promo;               // [promo_cap][kPromoPlanes] plane owners of the cells queued for tier m
    uint32_t promo_cap;           // queue slots that keep their owners
    uint32_t *tent;               // [tent_cap][5] tests tier 0 took by their FP32 sign: sorted cell id (or
                                  // kNoCell), the owners of the vertex's planes, the query point | sign << 31
    uint32_t tent_cap;
    uint32_t *cycles;             // optional [n]: per-cell cost counter (original order)
    // batch of clouds (null clouds: one cloud described by hdr)
    const IndexHeader *clouds;    // [n_clouds] per-cloud box and subtree root
    const int64_t *cloud_cells;   // [n_clouds + 1] sorted-cell range of each cloud
    const int64_t *quads;         // [n_clouds + 1] first work quad of each cloud within [begin, end)
    int n_clouds;
    const float *init_h;          // optional [n] (original order): predicted half-width of each cell
    float *extent;                // optional [n] (original order): half-width of each built cell
};

// A vertex is a triple of plane slots packed kBits bits apiece.
template <int kBits>
__device__ __forceinline__ uint32_t tri_pack(uint32_t a, uint32_t b, uint32_t c) {
    return a | (b << kBits) | (c << (2 * kBits));
}
template <int kBits>
__device__ __forceinline__ uint32_t tri_get(uint32_t t, int e) {
    return (t >> (kBits * e)) & ((1u << kBits) - 1u);
}
template <int kBits>
__device__ __forceinline__ bool tri_has_edge(uint32_t t, uint32_t x, uint32_t y) {
    const uint32_t a = tri_get<kBits>(t, 0), b = tri_get<kBits>(t, 1), c = tri_get<kBits>(t, 2);
    return (a == x && b == y) || (b == x && c == y) || (c == x && a == y);
}
__device__ __forceinline__ uint64_t bloom_bit(uint32_t id) { return 1ull << ((id * 2654435761u) >> 26); }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

}  // namespace pg

#include "pg_group_cell.hpp"

namespace pg {

// One thread per query (seed, three vertex planes, query plane): ids >= 0 are
// original point ids, ids < 0 box-plane codes -1 - (2 axis + side).  Writes
// (sign(P) sign(D), sign(D), sign(P)), or (2, 0, 0) when an expansion
// exceeded its capacity and (3, 0, 0) for an id out of range.
__global__ void exact_side_kernel(int64_t m, int64_t n, const IndexHeader *hdr, const float4 *pts,
                                  const uint32_t *inv, const int32_t *queries, int32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t *q = queries + 5 * i;
    bool ok = q[0] >= 0 && q[0] < n;
    xp::PlaneIn pl[4];
    for (int e = 0; e < 4; ++e) {
        const int32_t id = q[1 + e];
        ok = ok && id >= -6 && id < n;
        if (ok && id >= 0) {
            const float4 p = pts[inv[id]];
            pl[e] = xp::PlaneIn{p.x, p.y, p.z, p.w, -1};
        } else {
            pl[e] = xp::PlaneIn{0.f, 0.f, 0.f, 0.f, ok ? -1 - id : 0};
        }
    }
    int32_t *o = out + 3 * i;
    if (!ok) {
        o[0] = 3, o[1] = 0, o[2] = 0;
        return;
    }
    const float4 sp = pts[inv[q[0]]];
    const float s4[4] = {sp.x, sp.y, sp.z, sp.w};
    int sD = 0, sP = 0;
    if (!xp::exact_signs(pl, s4, hdr->box_lo, hdr->box_hi, sD, sP)) {
        o[0] = 2, o[1] = 0, o[2] = 0;
        return;
    }
    o[0] = sD * sP, o[1] = sD, o[2] = sP;
}

hipError_t exact_side(int64_t n, const void *index_ws, int64_t m, const int32_t *queries, int32_t *out,
                      hipStream_t stream) {
    if (m <= 0) return hipSuccess;
    const IndexLayout L = make_index_layout(n);
    const char *ws = (const char *)index_ws;
    hipLaunchKernelGGL(exact_side_kernel, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, stream, m, n,
                       (const IndexHeader *)(ws + L.header), (const float4 *)(ws + L.points),
                       (const uint32_t *)(ws + L.inv), queries, out);
    return hipGetLastError();
}

__global__ void cell_prep_kernel(uint32_t *counters) {
    if (threadIdx.x < 64) counters[threadIdx.x] = 0;  // work counters and the pool counter
}

// Every tier keeps its cells in LDS: the scratch holds the work counters,
// the row-pool counter, four per-cell lists (the FP64 tier-0 flags and the
// tier-m, tier-1 and tier-2 queues, each with room for every cell, so no
// queue can fill), the plane owners of the first cells queued for tier m
// (room for one cell in 32, and 64 more; a cell past them starts without) and
// the removal tests tier 0 took by their FP32 sign (room for one test per two
// cells, and 1024 more; a cell past them goes to the FP64 twin).
uint64_t promo_slots(int64_t cells) { return (uint64_t)cells / 32 + 64; }
uint64_t tent_slots(int64_t cells) { return (uint64_t)cells / 2 + 1024; }
uint64_t cell_scratch_bytes(int64_t cells) {
    return 256 + 16 * (uint64_t)cells + 4 * (uint64_t)kPromoPlanes * promo_slots(cells) + 20 * tent_slots(cells);
}

hipError_t launch_cells(const CellArgs &a, hipStream_t stream) {
    const IndexLayout L = make_index_layout(a.n);
    const char *ws = (const char *)a.index_ws;
    CellParams prm;
    const IndexHeader *hdr = (const IndexHeader *)(ws + L.header);
    prm.n = a.n;
    prm.hdr = hdr;
    prm.pts = (const float4 *)(ws + L.points);
    prm.perm = (const uint32_t *)(ws + L.perm);
    prm.inv = (const uint32_t *)(ws + L.inv);
    prm.nodes = (const Node *)(ws + L.nodes);
    prm.dup_head = (const uint32_t *)(ws + L.dup_head);
    prm.dup_next = (const uint32_t *)(ws + L.dup_next);
    prm.begin = a.begin;
    prm.end = a.end;
    prm.guesses = a.guesses;
    prm.guess_offsets = a.guess_offsets;
    prm.guesses_sorted = a.guesses_sorted;
    prm.guess_stride = a.guess_stride;
    prm.guess_d2 = a.guess_d2;
    prm.seed_window = a.seed_window;
    prm.rows = a.rows;
    prm.rows_elems = (uint64_t)a.rows_elems;
    prm.row_offsets = a.row_offsets;
    prm.counts = a.counts;
    prm.status = a.status;
    prm.stats = a.stats;
    prm.cycles = a.cycles;
    prm.clouds = a.clouds;
    prm.cloud_cells = a.cloud_cells;
    prm.quads = a.quads;
    prm.n_clouds = a.n_clouds;
    prm.init_h = a.init_h;
    prm.extent = a.extent;
    const int64_t count = a.end - a.begin;
    if (count <= 0) return hipSuccess;
    if (cell_scratch_bytes(count) > a.scratch_bytes || count > (int64_t)0xFFFFFFFF) return hipErrorInvalidValue;
    char *sc = (char *)a.scratch;
    prm.counters = (uint32_t *)sc;
    prm.pool_top = (unsigned long long *)(sc + 128);
    prm.queue_cap = (uint32_t)count;
    prm.big_list = (uint32_t *)(sc + 256);
    prm.huge_list = (uint32_t *)(sc + 256 + 4 * count);
    prm.exact_list = (uint32_t *)(sc + 256 + 8 * count);
    prm.mid_list = (uint32_t *)(sc + 256 + 12 * count);
    prm.promo = (int32_t *)(sc + 256 + 16 * count);
    prm.promo_cap = (uint32_t)promo_slots(count);
    prm.tent = (uint32_t *)(sc + 256 + 16 * count + 4 * kPromoPlanes * promo_slots(count));
    prm.tent_cap = (uint32_t)tent_slots(count);
    const int64_t cells_per_block = kT0Threads / kT0Lanes;
    const int64_t blocks_needed = (count + cells_per_block - 1) / cells_per_block;
    const int64_t max_blocks = (int64_t)a.waves * kWave / kT0Threads;
    const int64_t blocks = max_blocks < blocks_needed ? max_blocks : blocks_needed;
    // A tier-0 wave takes the pool in chunks, one device-wide atomic per
    // several rounds; what the waves of tier 0 leave unused is at most 1/16
    // of the pool, and as much again for its FP64 twin (same grid, same
    // chunks).  A pool that runs short is reported and grown, never silent.
    {
        const uint64_t waves0 = (uint64_t)(blocks > 0 ? blocks : 1) * (kT0Threads / kWave);
        const uint64_t per_wave = prm.rows_elems / (16 * waves0);
        prm.pool_chunk = per_wave >= 64 ? (uint32_t)(per_wave < 1024 ? per_wave : 1024) : 0u;
    }
    hipLaunchKernelGGL(cell_prep_kernel, dim3(1), dim3(64), 0, stream, prm.counters);
    if (hipError_t e = fill_u32(prm.exact_list, count, 0u, stream)) return e;
    hipLaunchKernelGGL(cell_group_kernel, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(kT0Threads), 0, stream, prm);
    {
        const uint64_t want = ((uint64_t)prm.tent_cap + 255) / 256;
        hipLaunchKernelGGL(cell_tentative_check_kernel, dim3((unsigned)(want < 2048 ? want : 2048)), dim3(256), 0,
                           stream, prm);
    }
    if (prm.clouds != nullptr)
        hipLaunchKernelGGL(cell_group_exact_batch_kernel, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(kT0Threads),
                           0, stream, prm);
    else
        hipLaunchKernelGGL(cell_group_exact_kernel, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(kT0Threads), 0,
                           stream, prm);
    // tier m keeps 12 waves per CU resident where tier 1 keeps 8
    hipLaunchKernelGGL(cell_group_mid_kernel, dim3((unsigned)(a.waves1 * 3 / 2 * kWave / kTmThreads)),
                       dim3(kTmThreads), 0, stream, prm);
    hipLaunchKernelGGL(cell_group_big_kernel, dim3((unsigned)(a.waves1 * kWave / kT1Threads)), dim3(kT1Threads), 0,
                       stream, prm);
    hipLaunchKernelGGL(cell_group_huge_kernel, dim3((unsigned)kTier2Waves), dim3(kT2Threads), 0, stream, prm);
    return hipGetLastError();
}

}  // namespace pg
