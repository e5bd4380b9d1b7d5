// Host-side launchers shared between the translation units of the library.
#pragma once

#include "pg_common.hpp"

namespace pg {

// Exclusive scan of n uint32 values (in place allowed).  tmp holds
// scan_tmp_elems(n) uint32; total (device, may be null) receives the sum.
int64_t scan_tmp_elems(int64_t n);
hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, uint32_t *total,
                              hipStream_t stream);

hipError_t build_index(int64_t n, const float *points, const float *weights, int leaf_size, void *ws,
                       hipStream_t stream);
hipError_t build_batch_index(int64_t n, int n_clouds, const float *points, const float *weights,
                             const int64_t *cloud_offsets, int leaf_size, void *ws, hipStream_t stream);
hipError_t leaf_stats(int64_t n, const void *ws, unsigned long long *out, hipStream_t stream);
hipError_t knn(int64_t n, void *index_ws, int k, int32_t *ids, float *d2, hipStream_t stream,
               int n_clouds = 0, bool sorted_rows = false);
hipError_t mlp_features(int64_t n, const void *index_ws, const int32_t *ids, const float *d2, float *X,
                        float *logd8, hipStream_t stream, int n_clouds = 0);
hipError_t mlp_forward(int64_t n, const float *X, const float *params, float *Y, hipStream_t stream);
int mlp_train_blocks(int64_t n);
hipError_t mlp_train_step(int64_t n, const float *X, const float *T, const float *mask, float count, float *params,
                          float *grad, float *m, float *v, float *part, float lr, int step, hipStream_t stream);

struct CellArgs {
    int64_t n;
    const void *index_ws;
    int64_t begin, end;           // sorted cell range
    const int32_t *guesses;       // may be null
    const int32_t *guess_offsets; // indexed by original id (null with guesses == null)
    int guesses_sorted;           // guesses already hold sorted indices (kNN output)
    int guess_stride = 0;         // > 0: Morton-ordered rows of this many sorted ids (offsets unused)
    const float *guess_d2 = nullptr;  // with guess_stride: squared distance of each row's last id (sorted order)
    int seed_window;              // Morton neighbours offered before the traversal
    int32_t *rows;                // row pool: every built cell's row, bump-allocated
    int64_t rows_elems;           // capacity of the row pool
    int64_t *row_offsets;         // [n] sorted cell -> row start in rows
    int32_t *counts;              // [n] original order
    int32_t *status;              // [n] original order
    unsigned long long *stats;    // kNumStats counters (accumulated)
    void *scratch;
    uint64_t scratch_bytes;
    int waves;                    // persistent tier-0 waves
    int waves1;                   // persistent tier-1 waves
    uint32_t *cycles;             // optional per-cell cost counters
    const IndexHeader *clouds = nullptr;  // batch: per-cloud headers (null: one cloud)
    const int64_t *cloud_cells = nullptr; // batch: [n_clouds + 1] sorted-cell offsets
    const int64_t *quads = nullptr;       // batch: [n_clouds + 1] quad offsets within [begin, end)
    int n_clouds = 0;
    const float *init_h = nullptr;        // optional predicted half-width per cell (original order)
    float *extent = nullptr;              // optional built half-width per cell (original order)
};
hipError_t launch_cells(const CellArgs &a, hipStream_t stream);
hipError_t exact_side(int64_t n, const void *index_ws, int64_t m, const int32_t *queries, int32_t *out,
                      hipStream_t stream);
uint64_t cell_scratch_bytes(int64_t cells);

// CSR helpers.
hipError_t gather_counts_sorted(const int32_t *counts, const uint32_t *perm, int64_t begin, int64_t end,
                                uint32_t *out, hipStream_t stream);
hipError_t emit_rows(int64_t n, const uint32_t *inv, const int32_t *counts, const int32_t *offsets,
                     const int32_t *src, const int64_t *src_offsets, int64_t src_stride, int32_t *adjacency,
                     hipStream_t stream);
hipError_t check_symmetry(int64_t n, const uint32_t *perm, const uint32_t *inv, const int32_t *counts, const int32_t *src,
                          const int64_t *src_offsets, int64_t src_stride, const int32_t *offsets,
                          unsigned long long *stats, hipStream_t stream);
hipError_t batch_quads(const int64_t *off, int n_clouds, int64_t begin, int64_t end, int64_t *quads,
                       hipStream_t stream);
hipError_t batch_localize(int64_t n, int n_clouds, const int64_t *cloud_off, const int32_t *offsets,
                          const int32_t *adj, int32_t *local_adj, int32_t *local_offsets, hipStream_t stream);
hipError_t pack_rows(int64_t begin, int64_t end, const uint32_t *perm, const int32_t *counts,
                     const uint32_t *local_offsets, const int32_t *rows, const int64_t *row_offsets,
                     int32_t *packed, hipStream_t stream);
hipError_t symmetrize_union(int64_t n, const int32_t *counts, const int32_t *offsets, const int32_t *adj,
                            int32_t *keep, uint32_t *extra, int32_t *new_counts, const int32_t *new_offsets,
                            int32_t *new_adj, int phase, hipStream_t stream);
hipError_t fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t stream);

}  // namespace pg
