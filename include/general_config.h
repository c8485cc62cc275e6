/* Compile-time settings read by the reference's callers (include/general_config.h:1-87): the
 * same names with the same meaning, so a caller that reads them compiles unchanged.
 *
 * Format parameters used by the conversions (sparse_conversion.h) and the test routines
 * (spmv_testroutine.h): SELL_SIGMA, CHUNK_SIZE (the SELL chunk height C), BSR_BlockDimRow.
 * Benchmark loop (spmv_benchmark.h): MIN_ITER / MAX_ITER iterations, TIME_LIMIT seconds.
 * Output files of the callers: MAT_FEATURES, MAT_PERFORMANCE.
 * Host-side scheduling names (KERNEL_FLAG, SCHE_MODE, SCHEDULE_STRATEGY, OMP_ROWS_SIZE) are
 * accepted for the callers' sake; the GPU kernels behind SpOps_* do not depend on them.
 * Every value can be overridden with -D. */
#ifndef SPOPS_DROPIN_GENERAL_CONFIG_H
#define SPOPS_DROPIN_GENERAL_CONFIG_H

#ifndef MAT_TILE_SIZE
#define MAT_TILE_SIZE 256
#endif
#ifndef MAX_DIAG_NUM
#define MAX_DIAG_NUM 10240
#endif
#ifndef MAX_ITER
#define MAX_ITER 1000
#endif
#ifndef MIN_ITER
#define MIN_ITER 20
#endif
#ifndef TIME_LIMIT
#define TIME_LIMIT 20.0
#endif
#ifndef NUM_FORMATS
#define NUM_FORMATS 7
#endif

#ifndef SELL_SIGMA
#define SELL_SIGMA 16384
#endif
#ifndef CHUNK_SIZE
#define CHUNK_SIZE 4
#endif
#ifndef NTRATIO
#define NTRATIO (0.6)
#endif
#ifndef CSR5_SIGMA
#define CSR5_SIGMA 16
#endif
#ifndef BSR_BlockDimRow
#define BSR_BlockDimRow 16
#endif

#ifndef OMP_ROWS_SIZE
#define OMP_ROWS_SIZE 64
#endif
/* 0 serial, 1 OpenMP, 2 row-balanced OpenMP in the reference; every flag runs the GPU kernel here */
#ifndef KERNEL_FLAG
#define KERNEL_FLAG 1
#endif

/* SpGEMM hash / vector parameters of the reference's CPU kernels (kept for callers that print them) */
#ifndef HASH_SCAL
#define HASH_SCAL 107
#endif
#ifndef MIN_HT_S
#define MIN_HT_S 8
#endif
#ifndef MIN_HT_N
#define MIN_HT_N 8
#endif
#ifndef SMALL_THRESHOLD
#define SMALL_THRESHOLD 100
#endif
#ifndef VEC_LENGTH
#define VEC_LENGTH 8
#endif
#ifndef VEC_LENGTH_BIT
#define VEC_LENGTH_BIT 3
#endif
#ifndef VEC_LENGTH_LONG
#define VEC_LENGTH_LONG 4
#endif
#ifndef VEC_LENGTH_LONG_BIT
#define VEC_LENGTH_LONG_BIT 2
#endif

#ifndef SCHEDULE_STRATEGY
#if defined(DYN)
#define SCHEDULE_STRATEGY dynamic
#elif defined(STATIC)
#define SCHEDULE_STRATEGY static, CHUNK_SIZE
#else
#define SCHEDULE_STRATEGY static
#endif
#endif
/* set_omp_schedule mode: 0 static, 1 static with chunk, 2 dynamic, 3 guided */
#ifndef SCHE_MODE
#define SCHE_MODE 1
#endif

#ifndef MAT_FEATURES
#define MAT_FEATURES "./features/mat_features.txt"
#endif
#ifndef MAT_PERFORMANCE
#define MAT_PERFORMANCE "./performance/mat_perf.txt"
#endif

#endif
