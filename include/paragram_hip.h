/*
 * paragram_hip.h -- C ABI of libparagram_hip.so, the gfx950 (MI355X) engine
 * behind paragram_amd.voronoi_diagram / paragram_amd.power_diagram.
 *
 * The reference exposes its engine as the pybind module paragram._backend._C
 * with two functions (paragram/csrc/ext.cpp:10-29, declared at
 * paragram/csrc/laguerre/laguerre.h:12-28):
 *     build_voronoi_ultra(points, knn, debug, initial_guesses,
 *                         initial_guesses_offsets, leaf_size)
 *     build_laguerre_voronoi_ultra(points, weights, ...)
 * each returning (adjacency int32[E], offsets int32[N+1], status int32[N]).
 * Those two calls are split here into the stages their C++ bodies run
 * (voronoi_ultra.cu:143-376, laguerre_ultra.cu:176-454), so that the host
 * (Python over ctypes) owns every allocation and the single device-to-host
 * read of E.
 *
 * Conventions
 *   - Every pointer argument except `stream` and host out-parameters is
 *     device memory of the current HIP device; sizes are element counts.
 *   - `stream` is a hipStream_t (0 = the null stream); all work is ordered on
 *     it and no call synchronises the host, except where stated.
 *   - The library allocates nothing: workspaces are sized by the *_bytes
 *     calls and passed in.
 *   - Return value 0 = success; otherwise pg_last_error() (thread-local)
 *     describes the failure.  Code 1 = invalid argument, 2 = HIP error.
 *   - N <= 67,108,863 (2^26 - 1) points per call.
 */
#ifndef PARAGRAM_HIP_H
#define PARAGRAM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_ABI_VERSION 2

/* Number of statistics counters written by the calls that take `stats`
 * (uint64 each, accumulated): 0 FP64 rechecks, 1 removal tests undecided
 * even by expansion arithmetic, 2 inconsistent, 3 cells promoted to a
 * larger-capacity tier, 4 clips, 5 candidates tested, 6 nodes visited,
 * 7 cells, 8 asymmetric edges, 9 entries to empty cells, 10 leaves visited,
 * 11 candidates rejected by the support test, 12 cells truncated (beyond the
 * largest tier: status 1 and an empty row), 13 rows that did not fit the row
 * pool, 14 (a maximum, not a sum) the pool elements such a launch needed,
 * 15 edge total E, 16 removal tests decided by expansion arithmetic,
 * 17..23 original ids (plus 1) of the first cells with such a test,
 * 24 cells the FP32-only tier gave up on at a removal test its filter could
 * not decide, 25 ... at a new vertex of uncertain orientation (those cells
 * are rebuilt by the FP64-capable twin), 26..31 reserved. */
#define PG_NUM_STATS 32

/* ABI version compiled into the library (== PG_ABI_VERSION). */
int pg_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *pg_last_error(void);

/* Bytes of the spatial-index workspace for n points.
 * Replaces the allocations of pwr_bvh.cu:228-270 (cudaMalloc of boxes,
 * nodes, primIDs, parents, counters) and the points4 tensor of
 * voronoi_ultra.cu:186. */
int pg_index_workspace_bytes(int64_t n, uint64_t *bytes);

/* Builds the LBVH over points (n x 3 float32, optional weights n float32):
 * exact AABB and FP32 clipping box, Morton sort, Karras hierarchy, AABB and
 * max-weight refit.  leaf_size < 0 selects the default.
 * Replaces pack_points_kernel / combine_points_weights_kernel
 * (voronoi_ultra.cu:30-37, laguerre_ultra.cu:31-41), build_cubql_bvh_and_knn
 * and build_pwr_bvh_and_knn (pwr_bvh.cu:217-339) and the root-box read of
 * voronoi_ultra.cu:234-238 (the box stays on the device). */
int pg_build_index(int64_t n, const float *points, const float *weights, int32_t leaf_size, void *workspace,
                   uint64_t workspace_bytes, void *stream);

/* Byte offsets within the index workspace of n points (host only, no device
 * work): offsets[0..5] = header, Morton-sorted float4 points (x, y, z, w),
 * perm (uint32 sorted -> original), inv (uint32 original -> sorted), nodes
 * (64-byte records: child 0 lo.xyz + max weight, hi.xyz + reference, then
 * child 1 the same; a reference with bit 31 set is a leaf,
 * first << 5 | (count - 1), else an internal node), total.  For inspecting
 * and testing the hierarchy. */
int pg_index_layout(int64_t n, uint64_t *offsets);

/* Leaf statistics of the index of pg_build_index (or of a batch):
 * out[0..3] = leaves, fewest points in a leaf, most, points in all leaves
 * (device memory, 4 uint64).
 * Replaces print_leaf_statistics's device pass (pwr_bvh.cu:341-453), printed
 * by the debug output of a power diagram (laguerre_ultra.cu:259-265). */
int pg_leaf_stats(int64_t n, const void *workspace, uint64_t *out, void *stream);

/* The k nearest neighbours (1 <= k <= 16) of every point over the index of
 * pg_build_index: ids[i*k + t] (original ids, ascending squared distance,
 * the point itself excluded, a short list padded with its last id or 0) and,
 * if d2 is non-null, the squared distances (FLT_MAX for padding).
 * The index workspace's sort scratch holds the queries whose search needs a
 * deep stack (the index itself is not changed).
 * Replaces cuBQL_knn_kernel (pwr_bvh.cu:159-215). */
int pg_knn(int64_t n, void *workspace, int32_t k, int32_t *ids, float *d2, void *stream);

/* Persistent waves used by pg_compute_cells on the current device. */
int pg_default_waves(int32_t *waves);

/* Bytes of cell scratch for launches over at most max_cells cells (work
 * counters, the row-pool counter, per cell the FP64 flag and the queues of
 * the three large-capacity tiers: 16 bytes, the 48 plane owners of up to
 * max_cells / 32 + 64 cells promoted from tier 0, and 20-byte records of up
 * to max_cells / 2 + 1024 removal tests that tier 0 took by their FP32 sign
 * for the FP64 check after it: 32 bytes a cell in all; the cells themselves
 * live in LDS).
 * Replaces ConvexCellMemory::required_bytes (voronoi_convex_cell.cuh:37-44). */
int pg_cell_scratch_bytes(int64_t max_cells, uint64_t *bytes);

/* Default int32 elements of the row pool of a launch over `cells` cells
 * (32 per cell + 65536). */
int pg_row_pool_elems(int64_t cells, uint64_t *elems);

/* Builds the cells of the Morton-sorted range [begin, end): each cell's
 * sorted neighbour row (original ids) goes to rows + row_offsets[i] (i the
 * sorted id), its row length to counts[orig] and its status to status[orig].
 * rows is a pool of rows_elems int32 from which every cell takes exactly its
 * row length.  A row that does not fit is counted in stats[13] with the
 * pool size the launch needs in stats[14] (its cell gets status 1 and an
 * empty row): the caller then runs the launch again with a pool that large.
 * No other capacity can fill: a cell that outgrows a tier is rebuilt by the
 * next (stats[3]), and only a cell beyond the largest (1023 planes, 1024
 * vertices) gets status 1 (stats[12]).
 * guesses / guess_offsets: optional CSR of candidate ids clipped first
 * (original ids, or sorted ids when guesses_sorted != 0); negative ids skip.
 * seed_window (0..16): without guesses, that many Morton-adjacent points are
 * offered first (replaces the kNN pre-clip of voronoi_ultra.cu:70-85;
 * changes speed only, never the result).
 * stats: PG_NUM_STATS counters, accumulated (stats[26]: removal tests tier 0
 * took by their FP32 sign, all decided in FP64 after it; stats[27]: those
 * whose sign was wrong (or whose vertex lay on a plane of a predicted cube),
 * their cells rebuilt by the FP64-capable instance;
 * stats[28..31]: original ids + 1 of the first such cells).
 * cell_cycles: optional (may be null) per-cell cost in GPU clock cycles,
 * original order; bit 31 set marks cells built by a large-capacity tier.
 * Replaces the chunk loop and compute_bvh_{voronoi,power_diagram}_kernel
 * (voronoi_ultra.cu:39-141,283-305; laguerre_ultra.cu:43-174,337-357). */
int pg_compute_cells(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *guesses,
                     const int32_t *guess_offsets, int32_t guesses_sorted, int32_t seed_window, int32_t *rows,
                     uint64_t rows_elems, int64_t *row_offsets, int32_t *counts, int32_t *status, uint64_t *stats,
                     void *scratch, uint64_t scratch_bytes, int32_t waves, uint32_t *cell_cycles, void *stream);

/* Largest cell range (end - begin) of pg_compute_cells_noalloc:
 * 1,571,837 cells (its static cell scratch holds 48 MiB, 32 bytes a cell). */
int pg_noalloc_max_cells(int64_t *cells);

/* pg_compute_cells_ex without the scratch arguments: the cell scratch is a
 * static 48 MiB array of the library on each GPU (held until the process
 * ends), so the call allocates nothing (no hipMalloc, hipFree or allocating
 * library routine); the caller still passes the inputs and outputs.  Ranges
 * of more than pg_noalloc_max_cells cells are refused (split them).  Calls
 * share the current GPU's array: calls that overlap on one GPU race.
 * No counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it). */
int pg_compute_cells_noalloc(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *guesses,
                             const int32_t *guess_offsets, int32_t guesses_sorted, int32_t guess_stride,
                             const float *guess_d2, int32_t seed_window, const float *init_half_width, float *extent,
                             int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts,
                             int32_t *status, uint64_t *stats, int32_t waves, uint32_t *cell_cycles, void *stream);

/* The library's own kNN pre-clip: row i (Morton-sorted order) of ids holds
 * the sorted indices of point i's k nearest neighbours (itself as padding);
 * n_clouds = 0 for the index of pg_build_index, else the batch's count.
 * kth_d2 (may be null): float [n], sorted order, the squared distance of the
 * row's last neighbour (FLT_MAX for a short row).  Pass both to
 * pg_compute_cells_ex / pg_compute_cells_batch with guess_stride = k (no
 * offsets): the cells read their rows in the order they are built, and skip
 * the traversal's candidates nearer than kth_d2 (already offered).
 * No counterpart in the reference (its kNN output, pwr_bvh.cu:159-215, is
 * indexed by original id and carries no distances). */
int pg_knn_guesses(int64_t n, void *workspace, int32_t n_clouds, int32_t k, int32_t *ids, float *kth_d2,
                   void *stream);

/* pg_compute_cells with guess_stride (0: guesses as in pg_compute_cells;
 * 1..64: rows of pg_knn_guesses with that stride, guess_offsets unused, and
 * guess_d2 its kth_d2 or null) and two optional per-cell arrays (original
 * order):
 * init_half_width (input, may be null): a predicted half-width h of each cell
 *   (max-norm distance of its vertices from the seed).  The cell then starts
 *   as the box cut by the cube of half-width h about the seed instead of the
 *   box; a cell whose final vertices touch the cube, or that the cube
 *   empties, is rebuilt from the box, so the output never depends on h
 *   (only the time does).
 * extent (output, may be null): the half-width of each built cell (0 for an
 *   empty cell), the training target of such predictions.
 * No counterpart in the reference; for the learned radius, which replaces the
 * kNN pre-clip radius of voronoi_ultra.cu:70-85 as the traversal's start. */
int pg_compute_cells_ex(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *guesses,
                        const int32_t *guess_offsets, int32_t guesses_sorted, int32_t guess_stride,
                        const float *guess_d2, int32_t seed_window, const float *init_half_width, float *extent,
                        int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts, int32_t *status,
                        uint64_t *stats, void *scratch, uint64_t scratch_bytes, int32_t waves,
                        uint32_t *cell_cycles, void *stream);

/* ---- Learned cell radius -------------------------------------------------
 * No counterpart in the reference (its traversal starts from the radius of a
 * kNN pre-clip, voronoi_ultra.cu:70-85).  A 16 -> 64 -> 64 -> 2 ReLU network
 * (5378 fp32 parameters: W1[16][64], b1, W2[64][64], b2, W3[64][2], b3;
 * BF16 products, FP32 accumulation) predicts log(h / d_8) and faces / 16 of
 * each cell from its 8 nearest neighbours; h feeds pg_compute_cells_ex.
 * Training and inference run on the GPU (inference on the matrix cores). */
int pg_mlp_param_count(void);

/* The 16 features per point (original order) from pg_knn's ids and d2 (k = 8)
 * over the index of pg_build_index; log_d8[i] = log of the 8th distance. */
int pg_mlp_features(int64_t n, const void *workspace, const int32_t *ids, const float *d2, float *features,
                    float *log_d8, void *stream);

/* pg_mlp_features over a batch index (pg_build_batch_index) with
 * pg_knn_batch's ids and d2: the hull features use each point's own cloud's
 * AABB. */
int pg_mlp_features_batch(int64_t n, const void *workspace, int32_t n_clouds, const int32_t *ids, const float *d2,
                          float *features, float *log_d8, void *stream);

/* out[i*2 + 0..1] = network(features[i*16 .. i*16 + 15]). */
int pg_mlp_forward(int64_t n, const float *features, const float *params, float *out, void *stream);

/* Elements of the float scratch of pg_mlp_train_step for a batch of n rows. */
int pg_mlp_train_scratch_elems(int64_t n, uint64_t *elems);

/* One training step on n rows: loss = mean over rows with mask > 0 (count
 * of them passed as `count`) of the squared error of both outputs against
 * targets[i*2 + 0..1]; grad[0..5377] receives the mean gradient and
 * grad[5378] the mean loss; with lr > 0, Adam (0.9, 0.999, 1e-8; step from 1)
 * updates params in place.  Gradients are reduced in a fixed order, so a step
 * is deterministic. */
int pg_mlp_train_step(int64_t n, const float *features, const float *targets, const float *mask, float count,
                      float *params, float *grad, float *adam_m, float *adam_v, float *scratch, float lr, int32_t step,
                      void *stream);

/* ---- Batches of clouds --------------------------------------------------
 * No counterpart in the reference, which builds one cloud per call
 * (voronoi_ultra.cu:143-376, laguerre_ultra.cu:176-454); for callers with many
 * clouds (e.g. a stream of LiDAR scans) that want one launch for all of them.
 * The clouds are contiguous ranges of one points array: cloud c is
 * cloud_offsets[c] .. cloud_offsets[c + 1] (host array of n_clouds + 1
 * int64, starting at 0 and ending at n, non-decreasing; 1 <= n_clouds <=
 * 4096).  Each cloud's cells, box (its own AABB +- 1) and neighbour ids are
 * those of a single-cloud call on that range; ids are global indices into
 * the points array (subtract the cloud's first index for cloud-local ids).
 * The CSR calls below (scan, symmetry check, emission) work unchanged on
 * the batch as one diagram over all n points. */

/* Bytes of the batch index workspace. */
int pg_batch_workspace_bytes(int64_t n, int32_t n_clouds, uint64_t *bytes);

/* Builds the batch index (per-cloud AABBs and boxes, one sort and hierarchy
 * over keys cloud id || per-cloud Morton code, each cloud's subtree root).
 * The cloud offsets (host array) are copied into the workspace before the
 * call returns. */
int pg_build_batch_index(int64_t n, int32_t n_clouds, const float *points, const float *weights,
                         const int64_t *cloud_offsets, int32_t leaf_size, void *workspace, uint64_t workspace_bytes,
                         void *stream);

/* Host helper: first work quad (four consecutive cells of one cloud) of each
 * cloud within the sorted cell range [begin, end); quads[n_clouds] = total. */
int pg_batch_quads(int32_t n_clouds, const int64_t *cloud_offsets, int64_t begin, int64_t end, int64_t *quads);

/* pg_compute_cells over the cells [begin, end) of a batch (any range of the
 * batch, so the chunk plan applies unchanged); the persistent waves take one
 * quad per round from a global counter.  cloud_offsets: the host array given
 * to pg_build_batch_index for this workspace (checked here; the work quads
 * are computed on the device from the workspace's copy, so nothing is copied
 * from the host).  guesses / guess_offsets: optional,
 * as for pg_compute_cells (original ids of the same cloud), or rows of
 * pg_knn_guesses (guess_stride = k, guess_d2 = its kth_d2 or null). */
int pg_compute_cells_batch(int64_t n, const void *workspace, int32_t n_clouds, const int64_t *cloud_offsets,
                           int64_t begin, int64_t end, const int32_t *guesses, const int32_t *guess_offsets,
                           int32_t guess_stride, const float *guess_d2, int32_t *rows, uint64_t rows_elems,
                           int64_t *row_offsets, int32_t *counts, int32_t *status, uint64_t *stats, void *scratch,
                           uint64_t scratch_bytes, int32_t waves, uint32_t *cell_cycles, void *stream);

/* Writes the cloud offsets (host array, checked as for pg_build_batch_index)
 * and the work quads of the whole range [0, n) into the batch workspace, and
 * waits for the copies (a set-up call).  The two entry points below then read
 * them from the workspace and launch kernels only, so a batch's index build,
 * cells and CSR can be captured in one graph and replayed for new points of
 * the same cloud sizes.  No counterpart in the reference. */
int pg_batch_set_offsets(int64_t n, int32_t n_clouds, const int64_t *cloud_offsets, void *workspace,
                         uint64_t workspace_bytes, void *stream);

/* pg_build_batch_index with the cloud offsets already in the workspace
 * (pg_batch_set_offsets). */
int pg_build_batch_index_resident(int64_t n, int32_t n_clouds, const float *points, const float *weights,
                                  int32_t leaf_size, void *workspace, uint64_t workspace_bytes, void *stream);

/* pg_compute_cells_batch over all n cells with the offsets and work quads
 * already in the workspace (pg_batch_set_offsets). */
int pg_compute_cells_batch_resident(int64_t n, const void *workspace, int32_t n_clouds, const int32_t *guesses,
                                    const int32_t *guess_offsets, int32_t guess_stride, const float *guess_d2,
                                    int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts,
                                    int32_t *status, uint64_t *stats, void *scratch, uint64_t scratch_bytes,
                                    int32_t waves, uint32_t *cell_cycles, void *stream);

/* pg_knn over a batch index: each point's neighbours come from its own cloud
 * (the batch's ids are global); a point alone in its cloud gets its own id as
 * padding (the cell kernels skip it). */
int pg_knn_batch(int64_t n, void *workspace, int32_t n_clouds, int32_t k, int32_t *ids, float *d2,
                 void *stream);

/* Exact signs of m removal predicates on the GPU, over the index of
 * pg_build_index (its points and FP32 box).  queries: m x 5 int32 (seed,
 * three planes of a vertex, query plane): ids >= 0 are original point ids,
 * ids < 0 box planes -1 - (2 axis + side) (side 0 = low).  out: m x 3 int32
 * (sign(P) sign(D), sign(D), sign(P)) with D = det[a_0; a_1; a_2] and
 * P = a_q . N - H_q D in the seed's doubled coordinates; (2, 0, 0) when an
 * expansion exceeded its capacity, (3, 0, 0) for an id out of range.  The
 * same device code decides the removal tests the FP64 filter cannot
 * (expansion arithmetic; Shewchuk, DCG 18, 1997).
 * Replaces the fast-math FP32 test of CCVertexIsInHalfSpace
 * (voronoi_convex_cell.cuh:256-260, faster_convex_cell.cuh:280-285) where
 * it is ambiguous. */
int pg_exact_side(int64_t n, const void *workspace, int64_t m, const int32_t *queries, int32_t *out, void *stream);

/* The same exact-sign code run on the host for one predicate (host memory;
 * no GPU needed): seed (x, y, z, w), planes 4 x (x, y, z, w) (three vertex
 * planes, then the query plane), box_codes[4] (-1: the plane of that point,
 * 0..5: box plane 2 axis + side of [box_lo, box_hi], point ignored); out as
 * pg_exact_side.  For checking the device code against rational arithmetic
 * on machines without a GPU. */
int pg_exact_side_host(const float *seed, const float *planes, const int32_t *box_codes, const float *box_lo,
                       const float *box_hi, int32_t *out);

/* The exact vertex position the FP64-capable cell tiers fall back to where
 * the FP64 bound on a vertex is loose (three nearly coaxial planes), on the
 * host for one vertex: seed and box as pg_exact_side_host, planes 3 x
 * (x, y, z, w), box_codes[3]; out[0..2] = the vertex y = 2 (x - seed) of the
 * three planes rounded to FP32, out[3] = a bound on the rounding error (max
 * norm), or -1 when the orientation D is not positive (no vertex).  The
 * reference computes vertices in FP32 only (voronoi_convex_cell.cuh:232-254);
 * this is for checking the device code against rational arithmetic. */
int pg_exact_vertex_host(const float *seed, const float *planes, const int32_t *box_codes, const float *box_lo,
                         const float *box_hi, float *out);

/* Bytes of temporary storage for pg_scan_counts over n values. */
int pg_scan_tmp_bytes(int64_t n, uint64_t *bytes);

/* offsets[0..n] = exclusive prefix sum of counts[0..n); offsets[n] = total.
 * Replaces torch::cat + torch::cumsum of voronoi_ultra.cu:347-350. */
int pg_scan_counts(int64_t n, const int32_t *counts, int32_t *offsets, void *tmp, void *stream);

/* out[i - begin] = counts[perm[i]] for sorted i in [begin, end). */
int pg_chunk_counts(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *counts,
                    int32_t *out, void *stream);

/* The cloud-local form of a batch's CSR (offsets [n+1] and adjacency over all
 * points, as the batch calls return them): local_adjacency[e] = adjacency[e]
 * minus the first point of its row's cloud, and local_offsets [n + n_clouds]
 * = each cloud's offsets (n_c + 1 of them, from 0) one cloud after another.
 * cloud_offsets: device int64 [n_clouds + 1], the batch's (0 .. n,
 * non-decreasing).  No counterpart in the reference (one cloud per call);
 * the per-cloud results of paragram_amd.batched_voronoi_diagram. */
int pg_batch_localize(int64_t n, int32_t n_clouds, const int64_t *cloud_offsets, const int32_t *offsets,
                      const int32_t *adjacency, int32_t *local_adjacency, int32_t *local_offsets, void *stream);

/* Copies the rows of sorted cells [begin, end) from rows + row_offsets[i]
 * to packed + local_offsets[i - begin]. */
int pg_pack_rows(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *counts,
                 const int32_t *local_offsets, const int32_t *rows, const int64_t *row_offsets, int32_t *packed,
                 void *stream);

/* Counts asymmetric edges and entries pointing at empty cells (stats[8],
 * stats[9]) and copies offsets[n] to stats[15].  Source rows are addressed
 * by sorted cell: src + src_offsets[i], or src + i * src_stride when
 * src_offsets is null. */
int pg_check_symmetry(int64_t n, const void *workspace, const int32_t *counts, const int32_t *src,
                      const int64_t *src_offsets, int64_t src_stride, const int32_t *offsets, uint64_t *stats,
                      void *stream);

/* Writes each original cell's row to adjacency + offsets[orig].
 * Replaces gather_adjacency_kernel (adjacency.cuh:7-43). */
int pg_emit_csr(int64_t n, const void *workspace, const int32_t *counts, const int32_t *offsets, const int32_t *src,
                const int64_t *src_offsets, int64_t src_stride, int32_t *adjacency, void *stream);

/* Union symmetrisation of an emitted CSR, one phase per call (0: count,
 * 1: new counts, 2: populate, 3: sort rows); the caller scans new_counts into
 * new_offsets between phases 1 and 2 and zeroes `extra` before phases 0 and 2.
 * Replaces count_symmetrized_adjacency_sizes and populate_symmetric_adjacency
 * (adjacency.cuh:78-159). */
int pg_symmetrize_union(int64_t n, const int32_t *counts, const int32_t *offsets, const int32_t *adjacency,
                        int32_t *keep, uint32_t *extra, int32_t *new_counts, const int32_t *new_offsets,
                        int32_t *new_adjacency, int32_t phase, void *stream);

/* p[0..n) = v (a fill kernel; used instead of hipMemsetAsync). */
int pg_fill_u32(uint32_t *p, int64_t n, uint32_t v, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PARAGRAM_HIP_H */
