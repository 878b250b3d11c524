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


/* ---------------------------------------------------------------------------
 * Fock-space indexing -- replaces xq.FSArray(m, n) iteration order and
 * FSArray.find (perceval/utils/states.py:255-298, perceval/backends/_slos.py:190).
 * States are ranked in descending lexicographic order of the occupation
 * vector, as pinned by tests/utils/test_statevector.py:430-438.
 * ------------------------------------------------------------------------- */

/* Number of n-photon states over m modes, C(m+n-1, n); -1 if above 2^63. */
int64_t pcvl_fock_count(int m, int n);

/* d_out[i] = rank of occupation row i (count rows of m int32) among states
 * with the same photon number; -1 for a row with a negative entry or more
 * than nmax photons.  Synchronises `stream`. */
int pcvl_fock_rank(const int32_t *d_occ, int64_t count, int m, int nmax, int64_t *d_out,
                   hipStream_t stream);

/* Row i of d_occ (m int32) = the n-photon state of rank d_rank[i]; a row of
 * -1 for a rank outside [0, C(m+n-1, n)).  Synchronises `stream`. */
int pcvl_fock_unrank(const int64_t *d_rank, int64_t count, int m, int n, int32_t *d_occ,
                     hipStream_t stream);

/* ---------------------------------------------------------------------------
 * SLOS -- replaces _Path._decompose/_Path.compute -> FSMap.compute_slos_layer
 * and xq.all_prob_normalize_output (perceval/backends/_slos.py:61-102,
 * 195-214).  A plan holds the photon-addition path of one input state
 * (_slos.py:61-86) and the rank tables; executing it for a unitary computes
 * the final layer over all C(m+n-1, n) output states in FSArray order:
 *   PCVL_SLOS_COEFS: complex c(s), amplitude = c(s) * sqrt(prod s! / prod t!)
 *                    (_slos.py:192);
 *   PCVL_SLOS_PROBS: |c(s)|^2 * prod s! / prod t!  (_slos.py:211-213).
 * ------------------------------------------------------------------------- */
#define PCVL_SLOS_COEFS 0
#define PCVL_SLOS_PROBS 1

typedef struct pcvl_slos_plan_s *pcvl_slos_plan;

/* m in [1, 64], at most 40 photons, every layer below 2^40 states. */
int pcvl_slos_plan_create(int m, const int32_t *h_in_occ, pcvl_slos_plan *out);

/* Photon number, output count, workspace bytes needed by execute, and the
 * input mode of each added photon (h_path: n int32, may be NULL). */
int pcvl_slos_plan_info(pcvl_slos_plan plan, int *n, int64_t *out_count, size_t *work_bytes,
                        int32_t *h_path);

/* d_U: m x m complex unitary (device); d_out: out_count complex (COEFS) or
 * doubles (PROBS); d_work: caller-owned device workspace of at least the
 * size reported by pcvl_slos_plan_info.  Enqueued on `stream`; allocates
 * nothing. */
int pcvl_slos_plan_execute(pcvl_slos_plan plan, const double *d_U, int mode, double *d_out,
                           void *d_work, size_t work_bytes, hipStream_t stream);

int pcvl_slos_plan_destroy(pcvl_slos_plan plan);

/* ---------------------------------------------------------------------------
 * Exact boson sampling -- replaces xq.Clifford2017.samples
 * (perceval/backends/_clifford2017.py:52-57): algorithm A of Clifford &
 * Clifford 2017, count samples drawn in parallel.  Sample s uses the
 * counter-based random stream (seed, first_sample + s), so a call is fully
 * determined by its arguments.  d_out: count rows of m int32 occupation
 * counts.  At most 24 photons and 256 modes.
 * ------------------------------------------------------------------------- */

/* Device workspace needed by pcvl_cc_samples for n photons and count samples. */
int pcvl_cc_workspace_bytes(int n, int64_t count, size_t *bytes);

/* Enqueued on `stream`, which is synchronised before returning. */
int pcvl_cc_samples(const double *d_U, int m, const int32_t *h_in_occ, int64_t count, uint64_t seed,
                    uint64_t first_sample, int32_t *d_out, void *d_work, size_t work_bytes,
                    hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PCVL_HIP_H */
