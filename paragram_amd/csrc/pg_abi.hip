// extern "C" entry points of libparagram_hip.so (declared in
// include/paragram_hip.h).  Argument checks and error reporting only; the
// work is in pg_index.hip, pg_cell.hip and pg_csr.hip.
#include "../../include/paragram_hip.h"
#include "pg_common.hpp"
#include "pg_exact.hpp"
#include "pg_kernels.hpp"

#include <mutex>
#include <stdio.h>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    return fail(2, std::string("Kernel failed: ") + what + ": " + hipGetErrorString(e));
}

inline hipStream_t as_stream(void *s) { return (hipStream_t)s; }

int check_n(int64_t n) {
    if (n < 0) return fail(1, "number of points must be non-negative");
    if (n > pg::kMaxPoints) return fail(1, "too many points for one call (limit 67108863)");
    return 0;
}

// Cell scratch of pg_compute_cells_noalloc: one array per GPU, held until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// the device array: its device address comes from hipGetSymbolAddress, once
// per device, and only that address is handed to kernels.
constexpr uint64_t kNoAllocScratchBytes = 48ull << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved on the first
// call on that device, with that device current).
int noalloc_scratch(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return check_hip(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(1, "device ordinal out of range for the static cell scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(2, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int pg_abi_version(void) { return PG_ABI_VERSION; }

const char *pg_last_error(void) { return g_last_error.c_str(); }

int pg_index_workspace_bytes(int64_t n, uint64_t *bytes) {
    if (int r = check_n(n)) return r;
    if (!bytes) return fail(1, "bytes is null");
    *bytes = pg::make_index_layout(n).total;
    return 0;
}

int pg_build_index(int64_t n, const float *points, const float *weights, int32_t leaf_size, void *workspace,
                   uint64_t workspace_bytes, void *stream) {
    if (int r = check_n(n)) return r;
    if (n == 0) return 0;
    if (!points || !workspace) return fail(1, "points and workspace must be non-null");
    if (workspace_bytes < pg::make_index_layout(n).total) return fail(1, "index workspace too small");
    int leaf = leaf_size < 0 ? pg::kDefaultLeaf : leaf_size;
    if (leaf < 1) leaf = 1;
    if (leaf > pg::kMaxLeaf) leaf = pg::kMaxLeaf;
    return check_hip(pg::build_index(n, points, weights, leaf, workspace, as_stream(stream)), "build_index");
}

int pg_index_layout(int64_t n, uint64_t *offsets) {
    if (int r = check_n(n)) return r;
    if (!offsets) return fail(1, "offsets is null");
    const pg::IndexLayout L = pg::make_index_layout(n);
    offsets[0] = L.header;
    offsets[1] = L.points;
    offsets[2] = L.perm;
    offsets[3] = L.inv;
    offsets[4] = L.nodes;
    offsets[5] = L.total;
    return 0;
}

int pg_leaf_stats(int64_t n, const void *workspace, uint64_t *out, void *stream) {
    if (int r = check_n(n)) return r;
    if (n == 0) return 0;
    if (!workspace || !out) return fail(1, "workspace and out must be non-null");
    return check_hip(pg::leaf_stats(n, workspace, (unsigned long long *)out, as_stream(stream)), "leaf_stats");
}

int pg_batch_workspace_bytes(int64_t n, int32_t n_clouds, uint64_t *bytes) {
    if (int r = check_n(n)) return r;
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (!bytes) return fail(1, "bytes is null");
    *bytes = pg::make_batch_layout(n, n_clouds).total;
    return 0;
}

namespace {
int check_cloud_offsets(int64_t n, int32_t n_clouds, const int64_t *off) {
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (!off) return fail(1, "cloud_offsets is null");
    if (off[0] != 0 || off[n_clouds] != n) return fail(1, "cloud_offsets must start at 0 and end at n");
    for (int32_t c = 0; c < n_clouds; ++c)
        if (off[c + 1] < off[c]) return fail(1, "cloud_offsets must be non-decreasing");
    return 0;
}
}  // namespace

int pg_build_batch_index(int64_t n, int32_t n_clouds, const float *points, const float *weights,
                         const int64_t *cloud_offsets, int32_t leaf_size, void *workspace, uint64_t workspace_bytes,
                         void *stream) {
    if (int r = check_n(n)) return r;
    if (int r = check_cloud_offsets(n, n_clouds, cloud_offsets)) return r;
    if (n == 0) return 0;
    if (!points || !workspace) return fail(1, "points and workspace must be non-null");
    const pg::BatchLayout B = pg::make_batch_layout(n, n_clouds);
    if (workspace_bytes < B.total) return fail(1, "batch workspace too small");
    int leaf = leaf_size < 0 ? pg::kDefaultLeaf : leaf_size;
    if (leaf < 1) leaf = 1;
    if (leaf > pg::kMaxLeaf) leaf = pg::kMaxLeaf;
    int64_t *dev_off = (int64_t *)((char *)workspace + B.offsets);
    // the host array may go when this returns: the copy completes first
    hipError_t e = hipMemcpyAsync(dev_off, cloud_offsets, 8 * ((size_t)n_clouds + 1), hipMemcpyHostToDevice,
                                  as_stream(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(as_stream(stream));
    if (e != hipSuccess) return check_hip(e, "build_batch_index");
    return check_hip(pg::build_batch_index(n, n_clouds, points, weights, dev_off, leaf, workspace, as_stream(stream)),
                     "build_batch_index");
}

int pg_batch_quads(int32_t n_clouds, const int64_t *cloud_offsets, int64_t begin, int64_t end, int64_t *quads) {
    if (!cloud_offsets || !quads) return fail(1, "cloud_offsets and quads must be non-null");
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (begin < 0 || end < begin || end > cloud_offsets[n_clouds]) return fail(1, "cell range out of bounds");
    int64_t q = 0;
    for (int32_t c = 0; c < n_clouds; ++c) {
        quads[c] = q;
        const int64_t lo = cloud_offsets[c] > begin ? cloud_offsets[c] : begin;
        const int64_t hi = cloud_offsets[c + 1] < end ? cloud_offsets[c + 1] : end;
        if (hi > lo) q += (hi - lo + 3) / 4;
    }
    quads[n_clouds] = q;
    return 0;
}

int pg_knn(int64_t n, void *workspace, int32_t k, int32_t *ids, float *d2, void *stream) {
    if (int r = check_n(n)) return r;
    if (k < 1 || k > 16) return fail(1, "k must be in [1, 16]");
    if (n == 0) return 0;
    if (!workspace || !ids) return fail(1, "workspace and ids must be non-null");
    return check_hip(pg::knn(n, workspace, k, ids, d2, as_stream(stream)), "knn");
}

int pg_mlp_param_count(void) { return 5378; }

int pg_mlp_features(int64_t n, const void *workspace, const int32_t *ids, const float *d2, float *features,
                    float *log_d8, void *stream) {
    if (int r = check_n(n)) return r;
    if (n == 0) return 0;
    if (!workspace || !ids || !d2 || !features || !log_d8)
        return fail(1, "workspace, ids, d2, features and log_d8 must be non-null");
    return check_hip(pg::mlp_features(n, workspace, ids, d2, features, log_d8, as_stream(stream)), "mlp_features");
}

int pg_mlp_features_batch(int64_t n, const void *workspace, int32_t n_clouds, const int32_t *ids, const float *d2,
                          float *features, float *log_d8, void *stream) {
    if (int r = check_n(n)) return r;
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (n == 0) return 0;
    if (!workspace || !ids || !d2 || !features || !log_d8)
        return fail(1, "workspace, ids, d2, features and log_d8 must be non-null");
    return check_hip(pg::mlp_features(n, workspace, ids, d2, features, log_d8, as_stream(stream), n_clouds),
                     "mlp_features_batch");
}

int pg_mlp_forward(int64_t n, const float *features, const float *params, float *out, void *stream) {
    if (n < 0) return fail(1, "number of rows must be non-negative");
    if (n == 0) return 0;
    if (!features || !params || !out) return fail(1, "features, params and out must be non-null");
    return check_hip(pg::mlp_forward(n, features, params, out, as_stream(stream)), "mlp_forward");
}

int pg_mlp_train_scratch_elems(int64_t n, uint64_t *elems) {
    if (n < 0 || !elems) return fail(1, "invalid argument");
    *elems = (uint64_t)pg::mlp_train_blocks(n) * (5378 + 1);
    return 0;
}

int pg_mlp_train_step(int64_t n, const float *features, const float *targets, const float *mask, float count,
                      float *params, float *grad, float *adam_m, float *adam_v, float *scratch, float lr, int32_t step,
                      void *stream) {
    if (n <= 0) return fail(1, "a training batch needs at least one row");
    if (!features || !targets || !mask || !params || !grad || !adam_m || !adam_v || !scratch)
        return fail(1, "all arrays must be non-null");
    if (step < 1) return fail(1, "step counts from 1");
    return check_hip(pg::mlp_train_step(n, features, targets, mask, count, params, grad, adam_m, adam_v, scratch, lr,
                                        step, as_stream(stream)),
                     "mlp_train_step");
}

int pg_knn_batch(int64_t n, void *workspace, int32_t n_clouds, int32_t k, int32_t *ids, float *d2,
                 void *stream) {
    if (int r = check_n(n)) return r;
    if (k < 1 || k > 16) return fail(1, "k must be in [1, 16]");
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (n == 0) return 0;
    if (!workspace || !ids) return fail(1, "workspace and ids must be non-null");
    return check_hip(pg::knn(n, workspace, k, ids, d2, as_stream(stream), n_clouds), "knn_batch");
}

int pg_knn_guesses(int64_t n, void *workspace, int32_t n_clouds, int32_t k, int32_t *ids, float *kth_d2,
                   void *stream) {
    if (int r = check_n(n)) return r;
    if (k < 1 || k > 16) return fail(1, "k must be in [1, 16]");
    if (n_clouds < 0 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [0, 4096]");
    if (n == 0) return 0;
    if (!workspace || !ids) return fail(1, "workspace and ids must be non-null");
    return check_hip(pg::knn(n, workspace, k, ids, kth_d2, as_stream(stream), n_clouds, true), "knn_guesses");
}

int pg_default_waves(int32_t *waves) {
    if (!waves) return fail(1, "waves is null");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return check_hip(e, "hipGetDevice");
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return check_hip(e, "hipDeviceGetAttribute");
    *waves = cus * 16;
    return 0;
}

int pg_cell_scratch_bytes(int64_t max_cells, uint64_t *bytes) {
    if (max_cells < 0) return fail(1, "max_cells must be non-negative");
    if (max_cells > pg::kMaxPoints) return fail(1, "too many cells for one launch (limit 67108863)");
    if (!bytes) return fail(1, "bytes is null");
    *bytes = pg::cell_scratch_bytes(max_cells);
    return 0;
}

int pg_row_pool_elems(int64_t cells, uint64_t *elems) {
    if (cells < 0 || !elems) return fail(1, "invalid argument");
    *elems = (uint64_t)pg::row_pool_elems(cells);
    return 0;
}

int pg_compute_cells(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *guesses,
                     const int32_t *guess_offsets, int32_t guesses_sorted, int32_t seed_window, int32_t *rows,
                     uint64_t rows_elems, int64_t *row_offsets, int32_t *counts, int32_t *status, uint64_t *stats,
                     void *scratch, uint64_t scratch_bytes, int32_t waves, uint32_t *cell_cycles, void *stream) {
    return pg_compute_cells_ex(n, workspace, begin, end, guesses, guess_offsets, guesses_sorted, 0, nullptr,
                               seed_window, nullptr, nullptr, rows, rows_elems, row_offsets, counts, status, stats,
                               scratch, scratch_bytes, waves, cell_cycles, stream);
}

int pg_compute_cells_ex(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *guesses,
                        const int32_t *guess_offsets, int32_t guesses_sorted, int32_t guess_stride,
                        const float *guess_d2, int32_t seed_window, const float *init_half_width, float *extent,
                        int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts, int32_t *status,
                        uint64_t *stats, void *scratch, uint64_t scratch_bytes, int32_t waves,
                        uint32_t *cell_cycles, void *stream) {
    if (int r = check_n(n)) return r;
    if (begin < 0 || end > n || begin > end) return fail(1, "cell range out of bounds");
    if (begin == end) return 0;
    if (!workspace || !rows || !row_offsets || !counts || !status || !stats || !scratch)
        return fail(1, "workspace, rows, row_offsets, counts, status, stats and scratch must be non-null");
    if (guess_stride < 0 || guess_stride > 64) return fail(1, "guess_stride must be in [0, 64]");
    if (guess_stride == 0 && (guesses == nullptr) != (guess_offsets == nullptr))
        return fail(1, "initial_guesses and initial_guesses_offsets must be provided together");
    if (guess_stride > 0 && guesses == nullptr) return fail(1, "guess_stride needs guesses");
    if (waves < 1) return fail(1, "waves must be positive");
    const int w0 = (waves + 3) / 4 * 4;
    if (scratch_bytes < pg::cell_scratch_bytes(end - begin)) return fail(1, "cell scratch too small");
    pg::CellArgs a;
    a.guess_stride = guess_stride;
    a.guess_d2 = guess_stride > 0 ? guess_d2 : nullptr;
    a.n = n;
    a.index_ws = workspace;
    a.begin = begin;
    a.end = end;
    a.guesses = guesses;
    a.guess_offsets = guess_offsets;
    a.guesses_sorted = guesses_sorted;
    a.seed_window = seed_window < 0 ? 0 : (seed_window > 16 ? 16 : seed_window);
    a.rows = rows;
    a.rows_elems = (int64_t)rows_elems;
    a.row_offsets = row_offsets;
    a.counts = counts;
    a.status = status;
    a.stats = (unsigned long long *)stats;
    a.scratch = scratch;
    a.scratch_bytes = scratch_bytes;
    a.waves = w0;
    a.waves1 = w0 / 2 > pg::kTier1Waves ? w0 / 2 : pg::kTier1Waves;  // 8 per CU: what tier 1 keeps resident
    a.cycles = cell_cycles;
    a.init_h = init_half_width;
    a.extent = extent;
    return check_hip(pg::launch_cells(a, as_stream(stream)), "compute_cells");
}

namespace {
// Cells [begin, end) of a batch whose cloud offsets and work quads are in the
// workspace already (written by the caller of this helper).
int cells_batch(int64_t n, const void *workspace, int32_t n_clouds, int64_t begin, int64_t end,
                const int32_t *guesses, const int32_t *guess_offsets, int32_t guess_stride, const float *guess_d2,
                int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts, int32_t *status,
                uint64_t *stats, void *scratch, uint64_t scratch_bytes, int32_t waves, uint32_t *cell_cycles,
                void *stream, const char *what) {
    const int w0 = (waves + 3) / 4 * 4;
    const pg::BatchLayout B = pg::make_batch_layout(n, n_clouds);
    const char *ws = (const char *)workspace;
    pg::CellArgs a;
    a.n = n;
    a.index_ws = workspace;
    a.begin = begin;
    a.end = end;
    a.guess_stride = guess_stride;
    a.guess_d2 = guess_stride > 0 ? guess_d2 : nullptr;
    a.guesses = guesses;
    a.guess_offsets = guess_offsets;
    a.guesses_sorted = 0;
    a.seed_window = 0;
    a.rows = rows;
    a.rows_elems = (int64_t)rows_elems;
    a.row_offsets = row_offsets;
    a.counts = counts;
    a.status = status;
    a.stats = (unsigned long long *)stats;
    a.scratch = scratch;
    a.scratch_bytes = scratch_bytes;
    a.waves = w0;
    a.waves1 = w0 / 2 > pg::kTier1Waves ? w0 / 2 : pg::kTier1Waves;  // 8 per CU: what tier 1 keeps resident
    a.cycles = cell_cycles;
    a.clouds = (const pg::IndexHeader *)(ws + B.clouds);
    a.cloud_cells = (const int64_t *)(ws + B.offsets);
    a.quads = (const int64_t *)(ws + B.quads);
    a.n_clouds = n_clouds;
    return check_hip(pg::launch_cells(a, as_stream(stream)), what);
}

int check_batch_cells(int64_t n, int64_t begin, int64_t end, const void *workspace, int32_t *rows,
                      int64_t *row_offsets, int32_t *counts, int32_t *status, uint64_t *stats, void *scratch,
                      uint64_t scratch_bytes, const int32_t *guesses, const int32_t *guess_offsets,
                      int32_t guess_stride, int32_t waves) {
    if (begin < 0 || end > n || begin > end) return fail(1, "cell range out of bounds");
    if (!workspace || !rows || !row_offsets || !counts || !status || !stats || !scratch)
        return fail(1, "workspace, rows, row_offsets, counts, status, stats and scratch must be non-null");
    if (waves < 1) return fail(1, "waves must be positive");
    if (scratch_bytes < pg::cell_scratch_bytes(end - begin)) return fail(1, "cell scratch too small");
    if (guess_stride < 0 || guess_stride > 64) return fail(1, "guess_stride must be in [0, 64]");
    if (guess_stride == 0 && (guesses == nullptr) != (guess_offsets == nullptr))
        return fail(1, "initial_guesses and initial_guesses_offsets must be provided together");
    if (guess_stride > 0 && guesses == nullptr) return fail(1, "guess_stride needs guesses");
    return 0;
}
}  // namespace

int pg_compute_cells_batch(int64_t n, const void *workspace, int32_t n_clouds, const int64_t *cloud_offsets,
                           int64_t begin, int64_t end, const int32_t *guesses, const int32_t *guess_offsets,
                           int32_t guess_stride, const float *guess_d2, int32_t *rows, uint64_t rows_elems,
                           int64_t *row_offsets, int32_t *counts, int32_t *status, uint64_t *stats, void *scratch,
                           uint64_t scratch_bytes, int32_t waves, uint32_t *cell_cycles, void *stream) {
    if (int r = check_n(n)) return r;
    if (int r = check_cloud_offsets(n, n_clouds, cloud_offsets)) return r;
    if (begin == end && begin >= 0 && end <= n) return 0;
    if (int r = check_batch_cells(n, begin, end, workspace, rows, row_offsets, counts, status, stats, scratch,
                                  scratch_bytes, guesses, guess_offsets, guess_stride, waves))
        return r;
    // the work quads from the cloud offsets that pg_build_batch_index left in
    // the workspace (checked above against the host copy), on the stream
    const pg::BatchLayout B = pg::make_batch_layout(n, n_clouds);
    char *ws = (char *)const_cast<void *>(workspace);
    hipError_t e = pg::batch_quads((const int64_t *)(ws + B.offsets), n_clouds, begin, end, (int64_t *)(ws + B.quads),
                                   as_stream(stream));
    if (e != hipSuccess) return check_hip(e, "compute_cells_batch");
    return cells_batch(n, workspace, n_clouds, begin, end, guesses, guess_offsets, guess_stride, guess_d2, rows,
                       rows_elems, row_offsets, counts, status, stats, scratch, scratch_bytes, waves, cell_cycles,
                       stream, "compute_cells_batch");
}

int pg_batch_set_offsets(int64_t n, int32_t n_clouds, const int64_t *cloud_offsets, void *workspace,
                         uint64_t workspace_bytes, void *stream) {
    if (int r = check_n(n)) return r;
    if (int r = check_cloud_offsets(n, n_clouds, cloud_offsets)) return r;
    if (!workspace) return fail(1, "workspace is null");
    const pg::BatchLayout B = pg::make_batch_layout(n, n_clouds);
    if (workspace_bytes < B.total) return fail(1, "batch workspace too small");
    std::vector<int64_t> quads((size_t)n_clouds + 1);
    if (int r = pg_batch_quads(n_clouds, cloud_offsets, 0, n, quads.data())) return r;
    char *ws = (char *)workspace;
    hipError_t e = hipMemcpyAsync(ws + B.offsets, cloud_offsets, 8 * ((size_t)n_clouds + 1), hipMemcpyHostToDevice,
                                  as_stream(stream));
    if (e == hipSuccess)
        e = hipMemcpyAsync(ws + B.quads, quads.data(), 8 * quads.size(), hipMemcpyHostToDevice, as_stream(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(as_stream(stream));  // a set-up call: the host arrays may go
    return check_hip(e, "batch_set_offsets");
}

int pg_build_batch_index_resident(int64_t n, int32_t n_clouds, const float *points, const float *weights,
                                  int32_t leaf_size, void *workspace, uint64_t workspace_bytes, void *stream) {
    if (int r = check_n(n)) return r;
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (n == 0) return 0;
    if (!points || !workspace) return fail(1, "points and workspace must be non-null");
    const pg::BatchLayout B = pg::make_batch_layout(n, n_clouds);
    if (workspace_bytes < B.total) return fail(1, "batch workspace too small");
    int leaf = leaf_size < 0 ? pg::kDefaultLeaf : leaf_size;
    if (leaf < 1) leaf = 1;
    if (leaf > pg::kMaxLeaf) leaf = pg::kMaxLeaf;
    const int64_t *dev_off = (const int64_t *)((char *)workspace + B.offsets);
    return check_hip(pg::build_batch_index(n, n_clouds, points, weights, dev_off, leaf, workspace, as_stream(stream)),
                     "build_batch_index_resident");
}

int pg_compute_cells_batch_resident(int64_t n, const void *workspace, int32_t n_clouds, const int32_t *guesses,
                                    const int32_t *guess_offsets, int32_t guess_stride, const float *guess_d2,
                                    int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts,
                                    int32_t *status, uint64_t *stats, void *scratch, uint64_t scratch_bytes,
                                    int32_t waves, uint32_t *cell_cycles, void *stream) {
    if (int r = check_n(n)) return r;
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (n == 0) return 0;
    if (int r = check_batch_cells(n, 0, n, workspace, rows, row_offsets, counts, status, stats, scratch,
                                  scratch_bytes, guesses, guess_offsets, guess_stride, waves))
        return r;
    return cells_batch(n, workspace, n_clouds, 0, n, guesses, guess_offsets, guess_stride, guess_d2, rows, rows_elems,
                       row_offsets, counts, status, stats, scratch, scratch_bytes, waves, cell_cycles, stream,
                       "compute_cells_batch_resident");
}

int pg_exact_side(int64_t n, const void *workspace, int64_t m, const int32_t *queries, int32_t *out, void *stream) {
    if (int r = check_n(n)) return r;
    if (m < 0) return fail(1, "number of queries must be non-negative");
    if (m == 0) return 0;
    if (n == 0 || !workspace || !queries || !out) return fail(1, "workspace, queries and out must be non-null");
    return check_hip(pg::exact_side(n, workspace, m, queries, out, as_stream(stream)), "exact_side");
}

int pg_exact_side_host(const float *seed, const float *planes, const int32_t *box_codes, const float *box_lo,
                       const float *box_hi, int32_t *out) {
    if (!seed || !planes || !box_codes || !box_lo || !box_hi || !out) return fail(1, "invalid argument");
    pg::xp::PlaneIn pl[4];
    for (int e = 0; e < 4; ++e) {
        if (box_codes[e] < -1 || box_codes[e] > 5) return fail(1, "box codes must be -1 (point) or 0..5");
        pl[e] = pg::xp::PlaneIn{planes[4 * e], planes[4 * e + 1], planes[4 * e + 2], planes[4 * e + 3], box_codes[e]};
    }
    int sD = 0, sP = 0;
    if (!pg::xp::exact_signs(pl, seed, box_lo, box_hi, sD, sP)) {
        out[0] = 2, out[1] = 0, out[2] = 0;
        return 0;
    }
    out[0] = sD * sP, out[1] = sD, out[2] = sP;
    return 0;
}

int pg_exact_vertex_host(const float *seed, const float *planes, const int32_t *box_codes, const float *box_lo,
                         const float *box_hi, float *out) {
    if (!seed || !planes || !box_codes || !box_lo || !box_hi || !out) return fail(1, "invalid argument");
    pg::xp::PlaneIn pl[3];
    for (int e = 0; e < 3; ++e) {
        if (box_codes[e] < -1 || box_codes[e] > 5) return fail(1, "box codes must be -1 (point) or 0..5");
        pl[e] = pg::xp::PlaneIn{planes[4 * e], planes[4 * e + 1], planes[4 * e + 2], planes[4 * e + 3], box_codes[e]};
    }
    float y[3], w;
    if (!pg::xp::exact_position(pl, seed, box_lo, box_hi, y, w)) {
        out[0] = out[1] = out[2] = 0.0f;
        out[3] = -1.0f;
        return 0;
    }
    out[0] = y[0], out[1] = y[1], out[2] = y[2], out[3] = w;
    return 0;
}

int pg_scan_tmp_bytes(int64_t n, uint64_t *bytes) {
    if (n < 0 || !bytes) return fail(1, "invalid argument");
    *bytes = 4 * (uint64_t)pg::scan_tmp_elems(n);
    return 0;
}

int pg_scan_counts(int64_t n, const int32_t *counts, int32_t *offsets, void *tmp, void *stream) {
    if (n < 0 || !offsets || (n > 0 && (!counts || !tmp))) return fail(1, "invalid argument");
    if (n == 0) return check_hip(pg::fill_u32((uint32_t *)offsets, 1, 0u, as_stream(stream)), "scan");
    return check_hip(pg::exclusive_scan_u32((const uint32_t *)counts, (uint32_t *)offsets, n, (uint32_t *)tmp,
                                            (uint32_t *)offsets + n, as_stream(stream)),
                     "scan");
}

int pg_chunk_counts(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *counts,
                    int32_t *out, void *stream) {
    if (begin < 0 || end > n || begin > end || !workspace || !counts || !out) return fail(1, "invalid argument");
    const pg::IndexLayout L = pg::make_index_layout(n);
    const uint32_t *perm = (const uint32_t *)((const char *)workspace + L.perm);
    return check_hip(pg::gather_counts_sorted(counts, perm, begin, end, (uint32_t *)out, as_stream(stream)),
                     "chunk_counts");
}

int pg_batch_localize(int64_t n, int32_t n_clouds, const int64_t *cloud_offsets, const int32_t *offsets,
                      const int32_t *adjacency, int32_t *local_adjacency, int32_t *local_offsets, void *stream) {
    if (int r = check_n(n)) return r;
    if (n_clouds < 1 || n_clouds > pg::kMaxClouds) return fail(1, "number of clouds must be in [1, 4096]");
    if (!cloud_offsets || !offsets || !local_offsets || (n > 0 && (!adjacency || !local_adjacency)))
        return fail(1, "invalid argument");
    return check_hip(pg::batch_localize(n, n_clouds, cloud_offsets, offsets, adjacency, local_adjacency,
                                        local_offsets, as_stream(stream)),
                     "batch_localize");
}

int pg_pack_rows(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *counts,
                 const int32_t *local_offsets, const int32_t *rows, const int64_t *row_offsets, int32_t *packed,
                 void *stream) {
    if (begin < 0 || end > n || begin > end || !workspace || !counts || !local_offsets || !rows || !row_offsets ||
        !packed)
        return fail(1, "invalid argument");
    const pg::IndexLayout L = pg::make_index_layout(n);
    const uint32_t *perm = (const uint32_t *)((const char *)workspace + L.perm);
    return check_hip(pg::pack_rows(begin, end, perm, counts, (const uint32_t *)local_offsets, rows, row_offsets,
                                   packed, as_stream(stream)),
                     "pack_rows");
}

int pg_check_symmetry(int64_t n, const void *workspace, const int32_t *counts, const int32_t *src,
                      const int64_t *src_offsets, int64_t src_stride, const int32_t *offsets, uint64_t *stats,
                      void *stream) {
    if (n <= 0) return 0;
    if (!workspace || !counts || !src || !offsets || !stats) return fail(1, "invalid argument");
    if (!src_offsets && src_stride <= 0) return fail(1, "src_stride must be positive without src_offsets");
    const pg::IndexLayout L = pg::make_index_layout(n);
    const uint32_t *perm = (const uint32_t *)((const char *)workspace + L.perm);
    const uint32_t *inv = (const uint32_t *)((const char *)workspace + L.inv);
    return check_hip(pg::check_symmetry(n, perm, inv, counts, src, src_offsets, src_stride, offsets,
                                        (unsigned long long *)stats, as_stream(stream)),
                     "check_symmetry");
}

int pg_emit_csr(int64_t n, const void *workspace, const int32_t *counts, const int32_t *offsets, const int32_t *src,
                const int64_t *src_offsets, int64_t src_stride, int32_t *adjacency, void *stream) {
    if (n <= 0) return 0;
    if (!workspace || !counts || !offsets || !src || !adjacency) return fail(1, "invalid argument");
    if (!src_offsets && src_stride <= 0) return fail(1, "src_stride must be positive without src_offsets");
    const pg::IndexLayout L = pg::make_index_layout(n);
    const uint32_t *inv = (const uint32_t *)((const char *)workspace + L.inv);
    return check_hip(pg::emit_rows(n, inv, counts, offsets, src, src_offsets, src_stride, adjacency,
                                   as_stream(stream)),
                     "emit_csr");
}

int pg_symmetrize_union(int64_t n, const int32_t *counts, const int32_t *offsets, const int32_t *adjacency,
                        int32_t *keep, uint32_t *extra, int32_t *new_counts, const int32_t *new_offsets,
                        int32_t *new_adjacency, int32_t phase, void *stream) {
    if (n <= 0) return 0;
    if (phase < 0 || phase > 3) return fail(1, "phase must be 0..3");
    return check_hip(pg::symmetrize_union(n, counts, offsets, adjacency, keep, extra, new_counts, new_offsets,
                                          new_adjacency, phase, as_stream(stream)),
                     "symmetrize_union");
}

int pg_fill_u32(uint32_t *p, int64_t n, uint32_t v, void *stream) {
    if (n < 0 || (n > 0 && !p)) return fail(1, "invalid argument");
    return check_hip(pg::fill_u32(p, n, v, as_stream(stream)), "fill");
}

// The most cells one launch of pg_compute_cells_noalloc covers: the largest
// range whose cell scratch fits the static array.
int64_t noalloc_max_cells() {
    int64_t c = (int64_t)(kNoAllocScratchBytes / 32);
    while (c > 0 && pg::cell_scratch_bytes(c) > kNoAllocScratchBytes) --c;
    while (pg::cell_scratch_bytes(c + 1) <= kNoAllocScratchBytes) ++c;
    return c;
}

int pg_noalloc_max_cells(int64_t *cells) {
    if (!cells) return fail(1, "cells is null");
    *cells = noalloc_max_cells();
    return 0;
}

int pg_compute_cells_noalloc(int64_t n, const void *workspace, int64_t begin, int64_t end, const int32_t *guesses,
                             const int32_t *guess_offsets, int32_t guesses_sorted, int32_t guess_stride,
                             const float *guess_d2, int32_t seed_window, const float *init_half_width, float *extent,
                             int32_t *rows, uint64_t rows_elems, int64_t *row_offsets, int32_t *counts,
                             int32_t *status, uint64_t *stats, int32_t waves, uint32_t *cell_cycles, void *stream) {
    if (int r = check_n(n)) return r;
    if (begin < 0 || end > n || begin > end) return fail(1, "cell range out of bounds");
    if (begin == end) return 0;
    const uint64_t need = pg::cell_scratch_bytes(end - begin);
    if (need > kNoAllocScratchBytes)
        return fail(1, "cell range too large for the static cell scratch (at most " +
                           std::to_string(noalloc_max_cells()) + " cells per launch; split the range)");
    void *ptr = nullptr;
    if (int r = noalloc_scratch(&ptr)) return r;
    return pg_compute_cells_ex(n, workspace, begin, end, guesses, guess_offsets, guesses_sorted, guess_stride,
                               guess_d2, seed_window, init_half_width, extent, rows, rows_elems, row_offsets, counts,
                               status, stats, ptr, kNoAllocScratchBytes, waves, cell_cycles, stream);
}

}  // extern "C"
