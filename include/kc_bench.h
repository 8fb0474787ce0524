/*
 * kc_bench.h — device microbenchmarks (libkc_bench.so) that measure the ceilings the
 * engine's kernels are reported against: the rate of random atomics on an array the size
 * of a count table, random 64-byte reads, and streaming copy bandwidth.  They have no
 * counterpart in the reference; bench.py uses them beside rocprof's counters.
 */
#ifndef KC_BENCH_H
#define KC_BENCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KC_BENCH_ATOMIC_ADD_U32 0 /* non-returning 32-bit atomicAdd at a random word */
#define KC_BENCH_ATOMIC_OR_U64 1  /* non-returning 64-bit atomicOr at a random word */
#define KC_BENCH_READ_64B 2       /* plain 8-byte load at a random word (one 64 B access) */
#define KC_BENCH_COPY 3           /* copy array_bytes to a second buffer (n_ops unused) */
#define KC_BENCH_ATOMIC_ADD_U32_WG 4  /* as 0 at workgroup scope (executed in the XCD's L2) */
#define KC_BENCH_ATOMIC_ADD_U32_RET 5 /* as 0 with the returned value used */

/* Best of `reps` timed launches (after one warm-up) of n_ops operations on a fresh
 * array of array_bytes; *ms_out receives the time of one launch.  Returns 0, or -1 (bad
 * argument), -2 (allocation failed), -3 (HIP error). */
int kc_bench_device(int32_t device, int kind, uint64_t array_bytes, uint64_t n_ops, int reps, double* ms_out);

#ifdef __cplusplus
}
#endif

#endif
