/*
 * pcvl_hip.h -- C ABI of the MI355X simulation core for Perceval's
 * Naive / SLOS / CliffordClifford2017 backends.
 *
 * The reference binds its native arithmetic (the closed exqalibur library)
 * through pybind11; each entry point below replaces one of those bindings and
 * cites the reference call site it serves.  Conventions:
 *   - every function returns 0 on success or a negative status
 *     (-1 bad argument, -2 HIP error, -3 allocation failure, -4 out of range);
 *     pcvl_last_error() returns the text of the calling thread's last failure;
 *   - no exception crosses the boundary;
 *   - complex numbers are interleaved (re, im) doubles;
 *   - unitaries are m x m, row-major, indexed [output mode][input mode], as
 *     _umat[ok, ik] in perceval/backends/_naive.py:65;
 *   - d_ pointers are device memory owned by the caller, h_ pointers are host
 *     memory owned by the caller; work is enqueued on `stream` (0 = default)
 *     unless the function says it synchronises.
 */
#ifndef PCVL_HIP_H
#define PCVL_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Text of the calling thread's last error ("" if none). */
const char *pcvl_last_error(void);

/* ABI version of this library (major * 100 + minor). */
int pcvl_abi_version(void);

/* ---------------------------------------------------------------------------
 * Permanents -- replaces xq.permanent_cx, called once per output state by
 * NaiveBackend._compute_permanent (perceval/backends/_naive.py:70-71) and by
 * the generic all_prob/prob_distribution loops
 * (perceval/backends/_abstract_backends.py:173-208).
 *
 * Batch layout: matrix b has order h_n[b] (0..28) and occupies h_n[b]^2
 * complex elements, row-major, starting at complex element h_offsets[b] of
 * d_mats.  d_out receives one complex permanent per matrix.
 * ------------------------------------------------------------------------- */
typedef struct pcvl_perm_plan_s *pcvl_perm_plan;

/* Build a reusable launch plan (work queue, partial-sum slots) for a batch
 * shape.  max_chunk_log2 <= 0 selects the default lane chunk (2^13 steps). */
int pcvl_perm_plan_create(const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                          int max_chunk_log2, pcvl_perm_plan *out);

/* Enqueue the permanents of the batch on `stream`.  Allocates nothing, so it
 * may be captured in a HIP graph and replayed; replays of one plan must not
 * overlap in time. */
int pcvl_perm_plan_execute(pcvl_perm_plan plan, const double *d_mats, double *d_out,
                           hipStream_t stream);

int pcvl_perm_plan_destroy(pcvl_perm_plan plan);

/* Work items, persistent grid size and the common order (0 if mixed). */
int pcvl_perm_plan_info(pcvl_perm_plan plan, int64_t *nitems, int *grid, int *single_n);

/* One-shot: plan, execute and synchronise `stream`. */
int pcvl_perm_batch(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets,
                    int64_t count, double *d_out, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PCVL_HIP_H */
