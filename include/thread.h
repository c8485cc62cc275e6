/* Host thread helpers with the reference's names and meaning (include/thread.h:1-23,
 * utils/thread.cpp:1-70): the callers size their own OpenMP loops with them.  The GPU kernels
 * behind SpOps_* and LeSpGEMM_VLength do not use the host thread count.
 *
 *   Le_get_core_num()       processors OpenMP can use (1 without OpenMP)
 *   Le_set_thread_num(n)    remember n and make it OpenMP's thread count
 *   Le_get_thread_num()     the remembered count, or Le_get_core_num() before any set
 *   Le_get_thread_id()      OpenMP thread number of the caller (0 outside a parallel region)
 *   set_omp_schedule(m, c)  runtime schedule: 0 static, 1 static with chunk c, 2 dynamic, 3 guided */
#ifndef SPOPS_DROPIN_THREAD_H
#define SPOPS_DROPIN_THREAD_H

#include <cstdio>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace spops_dropin {
inline int &thread_count_setting() {
  static int n = 0;                 // 0: not set
  return n;
}
}  // namespace spops_dropin

inline int Le_get_core_num() {
#ifdef _OPENMP
  return omp_get_num_procs();
#else
  return 1;
#endif
}

inline void Le_set_thread_num(const int thread_num) {
#ifdef _OPENMP
  spops_dropin::thread_count_setting() = thread_num;
  if (thread_num > 0) omp_set_num_threads(thread_num);
#else
  (void)thread_num;
  spops_dropin::thread_count_setting() = 1;
#endif
}

inline int Le_get_thread_num() {
#ifdef _OPENMP
  const int n = spops_dropin::thread_count_setting();
  return n > 0 ? n : Le_get_core_num();
#else
  return 1;
#endif
}

inline int Le_get_thread_id() {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

inline void set_omp_schedule(int sche_mode, int chunk_size = 0) {
#ifdef _OPENMP
  static const char *const names[] = {"Static", "StaticConst", "Dynamic", "Guided"};
  omp_sched_t kind = omp_sched_static;
  int chunk = 0;
  if (sche_mode == 1) chunk = chunk_size;
  else if (sche_mode == 2) kind = omp_sched_dynamic, chunk = chunk_size;
  else if (sche_mode == 3) kind = omp_sched_guided, chunk = chunk_size;
  else sche_mode = 0;
  omp_set_schedule(kind, chunk);
  std::printf("===  OMP %s schedule strategy  === \n", names[sche_mode]);
#else
  (void)sche_mode;
  (void)chunk_size;
#endif
}

#endif
