/* Platform description read by the reference's callers (include/plat_config.h:1-27).  The
 * callers use CPU_SOCKET * CPU_CORES_PER_SOC * CPU_HYPER_THREAD as their default host thread
 * count and CPU_MAX_FREQUENCY for the reference's cycle-counter timer; the drop-in's timer
 * (timer.h) reads a monotonic clock instead, so the frequency is informational.
 *
 * CPU_L2CACHE_SIZE and SIMD_WIDTH keep the reference's values on purpose: the kernel-3
 * dense/sparse classification of LeSpGEMM_VLength budgets MIXED_L2_FRACTION of
 * CPU_L2CACHE_SIZE (spgemm.h), and the SELL/BSR conversions align widths to
 * SIMD_WIDTH / 8 / sizeof(value) (sparse_conversion.h); other values would change the
 * layouts and the kernel-3 flags that the reference's outputs are compared on.
 * Every value can be overridden with -D. */
#ifndef SPOPS_DROPIN_PLAT_CONFIG_H
#define SPOPS_DROPIN_PLAT_CONFIG_H

#ifndef CPU_SOCKET
#define CPU_SOCKET 1
#endif
#ifndef CPU_CORES_PER_SOC
#define CPU_CORES_PER_SOC 16
#endif
#ifndef CPU_HYPER_THREAD
#define CPU_HYPER_THREAD 1
#endif
#ifndef NUMA_REGIONS
#define NUMA_REGIONS 1
#endif
#ifndef CPU_MAX_FREQUENCY
#define CPU_MAX_FREQUENCY 3.0e9
#endif

#ifndef CPU_L3CACHE_SIZE
#define CPU_L3CACHE_SIZE 536870912
#endif
#ifndef CPU_L2CACHE_SIZE
#define CPU_L2CACHE_SIZE 67108864
#endif
#ifndef CPU_L1DCACHE_SIZE
#define CPU_L1DCACHE_SIZE 4194304
#endif
#ifndef CACHE_LINE
#define CACHE_LINE 64
#endif
#ifndef MAIN_MEM_SIZE
#define MAIN_MEM_SIZE 503
#endif

#ifndef SIMD_WIDTH
#define SIMD_WIDTH 256
#endif

#endif
