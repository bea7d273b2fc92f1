/*
 * floors.h -- latency floors of the GPU at the colour-pass shape (a
 * measurement library for the benchmark, liblats_measure.so; not part of the
 * solver's C ABI and no reference counterpart).  Errors are reported through
 * lats_last_error() of libatssim_hip.so, which this library links.
 */
#ifndef LATS_MEASURE_FLOORS_H
#define LATS_MEASURE_FLOORS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mean microseconds per launch of one graph of `launches` kernels of
 * `threads` threads (256 per block) that do nothing (EMPTY) or gather 4
 * positions each through random indices and store their sum (GATHER: the
 * dependent ELL -> position chain of a colour pass, without its math); or
 * mean microseconds per barrier of one cooperative kernel of threads / 256
 * co-resident blocks crossing `launches` grid-wide barriers (GRID_BARRIER,
 * bounded waits: LATS_ERR_TIMEOUT if one gives up); or mean microseconds
 * per launch of a graph of kernels that each read the same buffer of
 * `threads` float4s, block b the same lines in every launch (REREAD: under
 * the FETCH_SIZE counter it shows whether a launch finds the previous
 * launch's lines in L2); or mean microseconds per phase of one cooperative
 * kernel of threads / 256 co-resident blocks in which block b starts phase q
 * once its two neighbours published q - 1 (HANDOFF_*: neighbours one block
 * apart, on other XCDs, with agent-scope acquire / release fences around the
 * hand-off or with relaxed device-scope accesses only; HANDOFF_XCD:
 * neighbours eight blocks apart, on the same XCD, relaxed; bounded waits). */
#define LATS_FLOOR_EMPTY 0
#define LATS_FLOOR_GATHER 1
#define LATS_FLOOR_GRID_BARRIER 2
#define LATS_FLOOR_REREAD 3
#define LATS_FLOOR_HANDOFF_FENCED 4
#define LATS_FLOOR_HANDOFF_RELAXED 5
#define LATS_FLOOR_HANDOFF_XCD 6
int lats_time_floor(int kind, uint32_t threads, uint32_t launches, float* us);

#ifdef __cplusplus
}
#endif
#endif /* LATS_MEASURE_FLOORS_H */
