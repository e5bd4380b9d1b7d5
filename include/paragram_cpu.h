/*
 * paragram_cpu.h -- C ABI of libparagram_cpu.so: the diagram built on host
 * cores (C++/OpenMP) with the clipping algorithm of the gfx950 kernels
 * (filtered FP32 predicates, FP64 recheck, exact expansion-arithmetic
 * fallback; kNN pre-clip, depth-first hierarchy traversal with the octant
 * and per-vertex ball prunes).  A reported CPU baseline beside the GPU path,
 * and a host stand-in for the engine behind the reference's seam
 * (paragram._backend._C, paragram/_backend.py:138-195) on machines without a
 * GPU; the product path (paragram_amd.voronoi_diagram / power_diagram)
 * never calls it.  Host memory throughout.
 */
#ifndef PARAGRAM_CPU_H
#define PARAGRAM_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGC_ABI_VERSION 1

int pgc_abi_version(void);

/* Message of the last failing call on this thread. */
const char *pgc_last_error(void);

/* Builds the diagram of n points (n x 3 float32; weights n float32 or null)
 * with leaf_size points per hierarchy leaf (< 1: 32) on nthreads threads
 * (<= 0: the OpenMP default); *handle receives the result (pgc_free).
 * Replaces build_voronoi_ultra / build_laguerre_voronoi_ultra
 * (voronoi_ultra.cu:143-376, laguerre_ultra.cu:176-454) on the host. */
int pgc_build(int64_t n, const float *points, const float *weights, int32_t leaf_size, int32_t nthreads,
              void **handle);

/* Adjacency entries E of a result (-1 for a null handle). */
int64_t pgc_result_edges(const void *handle);

/* Copies the result: adjacency int32[E], offsets int32[n+1], status int32[n]
 * (rows sorted ascending, original ids; the Diagram of paragram/_api.py:9-13). */
int pgc_result_copy(const void *handle, int32_t *adjacency, int32_t *offsets, int32_t *status);

/* out[0..8] = FP64 rechecks, tests decided by expansion arithmetic, tests
 * left undecided, clips, candidates tested, hierarchy nodes expanded, and
 * wall times in nanoseconds: the whole build, the index (sort, hierarchy,
 * duplicate groups), the cells (kNN pre-clip included). */
int pgc_result_stats(const void *handle, int64_t *out);

void pgc_free(void *handle);

#ifdef __cplusplus
}
#endif

#endif /* PARAGRAM_CPU_H */
