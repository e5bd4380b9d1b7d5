// Shared definitions of the paragram_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

// ------------------------------------------------------------------ limits
// Rows of every capacity tier (the tiers and their plane / vertex capacities
// are defined in pg_group_cell.hpp) are bump-allocated from one row pool, so
// a row takes exactly its length.  A cell of F faces has 2F-4 vertices.
constexpr int kTier1Waves = 1024;      // least persistent waves of the tier-1 launch (8 per CU where more)
constexpr int kTier2Waves = 256;       // persistent waves of the tier-2 launch
constexpr int kPoolPerCell = 32;       // default row-pool entries per cell (a Poisson cell has 15.5 faces)
constexpr int64_t kPoolSlack = 65536;  // plus this many entries per launch

// Default row-pool elements of a launch over m cells.  A launch whose rows
// do not fit reports it (kStatPoolShort, with the size needed in
// kStatPoolDemand) and the host runs it again with a pool of that size.
__host__ __device__ inline int64_t row_pool_elems(int64_t m) { return m * kPoolPerCell + kPoolSlack; }
constexpr int kStackCap = 64;          // LBVH depth is < 56 for N < 2^26
constexpr int kMaxLeaf = 32;
constexpr int kDefaultLeaf = 32;      // points per LBVH leaf (two 16-lane batches); measured fastest
constexpr int64_t kMaxPoints = (int64_t(1) << 26) - 1;  // leaf refs keep 26 bits of first index
constexpr int kWave = 64;

// Status codes of the reference (voronoi_convex_cell.cuh:91-100).
enum Status : int32_t {
    kSuccess = 0,
    kPlaneOverflow = 1,
    kVertexOverflow = 2,
    kEmptyCell = 3,
    kSecurityRadiusNotReached = 4,
    kInconsistentBoundary = 5,
};

// Device-side counters, accumulated with one atomic per wave.
enum StatSlot : int {
    kStatFp64Rechecks = 0,   // FP32 filter undecided, recomputed in FP64
    kStatUnresolved = 1,     // FP64 and expansion arithmetic both undecided (FP64 sign taken)
    kStatInconsistent = 2,   // orientation or boundary inconsistency
    kStatPromoted = 3,       // cells passed on to a larger-capacity tier
    kStatClips = 4,          // successful clips
    kStatCandidates = 5,     // candidates that reached the vertex tests
    kStatNodes = 6,          // internal nodes visited
    kStatCells = 7,          // cells processed
    kStatAsymmetric = 8,     // directed edges without a reverse edge
    kStatDropped = 9,        // row entries pointing at empty cells
    kStatLeaves = 10,        // leaves visited
    kStatSkipped = 11,       // candidates rejected by the support test
    kStatTruncated = 12,     // cells beyond the largest tier: status 1, empty row
    kStatPoolShort = 13,     // rows that did not fit the row pool (the launch must run again)
    kStatPoolDemand = 14,    // (max, not sum) row-pool elements a launch that ran short needed
    kStatEdges = 15,         // edge total E (written by the symmetry check)
    kStatExact = 16,         // removal tests decided by expansion arithmetic (FP64 undecided)
    kStatExactLog = 17,      // 17..23: original ids of the first cells with such a test, plus 1
    kStatBailTest = 24,      // FP32-only tier 0 gave a cell up: a removal test its filter could not decide
    kStatBailVertex = 25,    // ... a new vertex whose FP32 orientation was not safely positive
    kStatTentative = 26,     // removal tests tier 0 took by their FP32 sign (decided in FP64 after it)
    kStatTentativeWrong = 27,  // ... whose FP64 decision differed (or whose vertex lay on a predicted
                               // cube's plane): the cell was rebuilt by the FP64 twin
    kStatTentativeLog = 28,  // 28..31: original ids of the first such cells, plus 1
    kNumStats = 32,
};
constexpr int kTentativeLogSlots = 4;
// Removal tests one tier-0 cell may take by their FP32 sign (kept in LDS
// until the cell is written out); a cell with more goes to the FP64 twin.
constexpr int kTentMax = 8;
constexpr uint32_t kNoCell = 0xFFFFFFFFu;  // a record slot that holds no test
// Up to this many cells flagged for the FP64-capable tier 0 go to tier m's
// queue instead (tier m builds them beside its own; the twin's launch would
// last as long as its slowest cell for a handful of cells).
constexpr uint32_t kFewFlagged = 2048;
constexpr int kExactLogSlots = 7;
constexpr int kPromoPlanes = 48;  // plane owners kept for a cell promoted to tier m (tier 0's plane capacity)
// Statistics first summed per group in LDS: slots 0..11, then the two bail
// counters (LDS slots 12 and 13).
constexpr int kCtrBailTest = 12, kCtrBailVertex = 13;
constexpr int kNumCtr = 14;
__host__ __device__ constexpr int ctr_stat_slot(int k) {
    return k == kCtrBailTest ? kStatBailTest : (k == kCtrBailVertex ? kStatBailVertex : k);
}

// ------------------------------------------------------------------ LBVH
// Internal node of the Karras hierarchy in traversal form: the boxes of both
// children sit in the parent (one 64-byte fetch per visited node).  A child
// reference is an internal node index, or a leaf (bit 31 set) that names a
// contiguous range of the Morton-sorted points: first << 5 | (count - 1).
struct alignas(16) Node {
    float4 lo0;  // child 0: lo.xyz, max weight
    float4 hi0;  // child 0: hi.xyz, reference (bits)
    float4 lo1;
    float4 hi1;
};
static_assert(sizeof(Node) == 64, "node record is one 64-byte line");

__host__ __device__ inline uint32_t leaf_ref(uint32_t first, uint32_t count) {
    return 0x80000000u | (first << 5) | (count - 1u);
}
__host__ __device__ inline bool is_leaf(uint32_t ref) { return (ref & 0x80000000u) != 0; }
__host__ __device__ inline uint32_t leaf_first(uint32_t ref) { return (ref & 0x7fffffffu) >> 5; }
__host__ __device__ inline uint32_t leaf_count(uint32_t ref) { return (ref & 31u) + 1u; }

// Header at the start of the index workspace (device memory).
struct IndexHeader {
    float raw_lo[3], raw_hi[3];   // exact AABB of the points
    float box_lo[3], box_hi[3];   // clipping box: (min - 1) - 1e-5, (max + 1) + 1e-5 in FP32
    uint32_t root_ref;
    uint32_t n;
    uint32_t has_weights;
    uint32_t leaf_size;
    uint32_t has_duplicates;      // some points coincide in position and weight
    uint32_t knn_spill;           // kNN queries left to the deep-stack pass (see knn)
};

// Offsets of the arrays inside the index workspace.
struct IndexLayout {
    uint64_t header;     // IndexHeader
    uint64_t points;     // float4 [n]  Morton-sorted (x, y, z, w)
    uint64_t perm;       // uint32 [n]  sorted -> original
    uint64_t inv;        // uint32 [n]  original -> sorted
    uint64_t nodes;      // Node [max(n-1,1)]
    uint64_t keys;       // uint32 [n]  sort keys (ping)
    uint64_t keys_alt;   // uint32 [n]  (pong)
    uint64_t vals_alt;   // uint32 [n]  (after the build: the kNN pass's spilled queries)
    uint64_t hist;       // uint32 [256 * sort_blocks + 256]
    uint64_t children;   // uint32 [2*(n-1)]  Karras child codes
    uint64_t parents;    // uint32 [2n-1]     parent of internal (0..n-2) and leaf (n-1..2n-2)
    uint64_t ranges;     // uint2  [n-1]      covered range of internal nodes
    uint64_t boxes;      // float4 [2*(n-1)]  internal node AABB lo/w, hi
    uint64_t flags;      // uint32 [n-1]      refit arrival counters
    uint64_t reduce;     // float  [6 * reduce_blocks]
    uint64_t scan_tmp;   // uint32 block sums of the histogram scan
    uint64_t dup_head;   // uint32 [n]  first sorted index of the point's duplicate group
    uint64_t dup_next;   // uint32 [n]  next member of the group (0xFFFFFFFF at the end)
    uint64_t dup_table;  // uint32 [2n] hash table of the points in long equal-key runs (build only)
    uint64_t total;
    int sort_blocks;
    int reduce_blocks;
};

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;                 // items per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kReduceThreads = 256;
constexpr int kReduceBlocksMax = 1024;

__host__ __device__ inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

inline IndexLayout make_index_layout(int64_t n) {
    IndexLayout L{};
    const uint64_t nn = (uint64_t)(n > 0 ? n : 1);
    const uint64_t ni = nn > 1 ? nn - 1 : 1;
    L.sort_blocks = (int)((nn + kSortTile - 1) / kSortTile);
    L.reduce_blocks = (int)((nn + kReduceThreads * 8 - 1) / (kReduceThreads * 8));
    if (L.reduce_blocks > kReduceBlocksMax) L.reduce_blocks = kReduceBlocksMax;
    if (L.reduce_blocks < 1) L.reduce_blocks = 1;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        uint64_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.header = take(sizeof(IndexHeader));
    L.points = take(16 * nn);
    L.perm = take(4 * nn);
    L.inv = take(4 * nn);
    L.nodes = take(sizeof(Node) * ni);
    L.keys = take(4 * nn);
    L.keys_alt = take(4 * nn);
    L.vals_alt = take(4 * nn);
    L.hist = take(4 * (256ull * L.sort_blocks + 256));
    L.children = take(8 * ni);
    L.parents = take(4 * (2 * nn));
    L.ranges = take(8 * ni);
    L.boxes = take(32 * ni);
    L.flags = take(4 * ni);
    L.reduce = take(4 * 6 * (uint64_t)L.reduce_blocks);
    L.scan_tmp = take(4 * ((256ull * L.sort_blocks + 4095) / 4096 + 2));
    L.dup_head = take(4 * nn);
    L.dup_next = take(4 * nn);
    L.dup_table = take(8 * nn);
    L.total = o;
    return L;
}

// Batch of clouds in one index: the single-cloud layout over all points,
// then one IndexHeader per cloud (its AABB, clipping box and subtree root).
constexpr int kMaxClouds = 1 << 12;
struct BatchLayout {
    IndexLayout idx;
    uint64_t clouds;   // IndexHeader [n_clouds]
    uint64_t offsets;  // int64 [n_clouds + 1]  cloud ranges (input order = sorted order)
    uint64_t quads;    // int64 [n_clouds + 1]  work quads of the current cell range
    uint64_t total;
};
inline BatchLayout make_batch_layout(int64_t n, int n_clouds) {
    BatchLayout B{};
    const uint64_t nc = (uint64_t)(n_clouds > 0 ? n_clouds : 1);
    B.idx = make_index_layout(n);
    B.clouds = B.idx.total;
    B.offsets = align_up(B.clouds + sizeof(IndexHeader) * nc, 256);
    B.quads = align_up(B.offsets + 8 * (nc + 1), 256);
    B.total = align_up(B.quads + 8 * (nc + 1), 256);
    return B;
}
// Sort key of a batch: cloud id in the top cloud_bits bits, then the point's
// Morton code on a per-cloud cubic grid with bits_per_axis bits per axis.
inline int batch_cloud_bits(int n_clouds) {
    int b = 0;
    while ((1 << b) < n_clouds) ++b;
    return b;
}
inline int batch_axis_bits(int n_clouds) {
    const int m = (32 - batch_cloud_bits(n_clouds)) / 3;
    return m < 10 ? m : 10;
}

}  // namespace pg
