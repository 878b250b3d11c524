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
 *     unless the function says it synchronises;
 *   - a plan is immutable after creation except through its own setters, so
 *     several threads may execute one plan at once on different streams with
 *     different output and workspace buffers (the permanent plans' queue
 *     cursor and partial sums are plan state: one execution of a permanent
 *     plan at a time); creating, setting and destroying a plan are not
 *     concurrent with its use.
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
 * Batch layout: matrix b has order h_n[b] (0..32) and occupies h_n[b]^2
 * complex elements, row-major, starting at complex element h_offsets[b] of
 * d_mats.  d_out receives one complex permanent per matrix.  Orders 9..28
 * share one persistent launch; orders 29..32 take a launch of their own.
 * ------------------------------------------------------------------------- */
typedef struct pcvl_perm_plan_s *pcvl_perm_plan;

/* Build a reusable launch plan (work queue, partial-sum slots) for a batch
 * shape.  max_chunk_log2 <= 0 selects the default lane chunk (2^11 steps);
 * the matrices that make up the last round of the queue get chunks 4x
 * shorter. */
int pcvl_perm_plan_create(const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                          int max_chunk_log2, pcvl_perm_plan *out);

/* As pcvl_perm_plan_create, with the kernel variant of single-order batches:
 * 0 picks by order, 1 keeps Gray-code row 1 in VGPRs, 2 reads it from LDS
 * (fewer VGPRs, more resident waves); 3 and 4 are 1 and 2 with each row flip
 * fused into the preceding term.  Mixed-order batches ignore it.  Used by the
 * autotuner. */
int pcvl_perm_plan_create_ex(const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                             int max_chunk_log2, int variant, pcvl_perm_plan *out);

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

/* pcvl_perm_batch without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The plan
 * (work items, queue cursor, partial sums) goes into a 16 MiB scratch array
 * that each GPU holds from module load until the process ends
 * (pcvl_perm_noalloc_max_bytes); a batch whose plan needs more returns -4
 * before anything runs.  Calls that overlap on one GPU share that array and
 * race.  Synchronises `stream`; results equal pcvl_perm_batch's bit for bit. */
int pcvl_perm_batch_noalloc(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                            double *d_out, hipStream_t stream);
size_t pcvl_perm_noalloc_max_bytes(void);


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

/* Ranks (FSArray order) of the n-photon, m-mode states that the masks keep
 * (AStrongSimulationBackend.set_mask / xq.FSMask / the masked xq.FSArray,
 * perceval/backends/_abstract_backends.py:103-146): h_cond holds nmask rows
 * of m conditions, -1 for any count, else a count matched exactly, or reached
 * at least on the modes where h_at_least (m flags, may be NULL) is non-zero.
 * mask_n is the photon count the masks were set for (set_mask's n): with
 * mask_n <= n a state is kept when it matches a row; with mask_n > n, when
 * the mask_n - n photons still to come can complete a row (no exact count
 * exceeded, the photons missing from exact counts and floors at most
 * mask_n - n, and exactly mask_n - n for a row with no free or at-least mode
 * to take the rest), as the reference's simulator needs for the separated parts of
 * a noisy input (perceval/simulators/simulator.py:412-470).  d_out must hold
 * C(m+n-1, n) entries; *d_count (device) receives the number of states kept.
 * At most 16 masks and 2^34 states.  Synchronises `stream`. */
int pcvl_fock_mask_select(int m, int n, int mask_n, const int32_t *h_cond, int nmask, const int32_t *h_at_least,
                          int64_t *d_out, int64_t *d_count, hipStream_t stream);

/* d_dst[k] = d_src[d_idx[k]] (float64), e.g. the probabilities of a mask's
 * states. */
int pcvl_gather_f64(const double *d_src, const int64_t *d_idx, int64_t count, double *d_dst, hipStream_t stream);

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

/* m in [1, 256], at most 40 photons, every layer below 2^40 states (the rank
 * tables of the per-state kernel, 16 m (n + 1) bytes, within 160 KiB). */
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

/* A plan restricted to a herald mask: the masked FSArray layers of
 * SLOSBackend._deploy (perceval/backends/_slos.py:161-164; the mask of
 * set_mask, perceval/backends/_abstract_backends.py:103-137).  h_cond holds
 * one condition per mode: -1 for any count, c >= 0 for exactly c photons.
 * Intermediate layers hold only the states that can still match (herald
 * counts at most c, and the free modes at most n - sum c photons), so the
 * states outside the mask are never computed.  execute writes the matching
 * output states only, in FSArray order (out_count of pcvl_slos_plan_info:
 * C(m_free + n - sum c - 1, n - sum c), or 0 when the heralds need more than
 * n photons).  Runs the block kernel (8); needs m >= 2, at least two free
 * modes and 1..30 photons.  Summation runs in herald-modes-first order, so
 * results match the unmasked plan to rounding, not to the bit. */
int pcvl_slos_plan_create_masked(int m, const int32_t *h_in_occ, const int32_t *h_cond, pcvl_slos_plan *out);
/* The general masks of set_mask(masks, n, at_least_modes): nrows condition
 * rows (h_conds, nrows x m, -1 any count, c >= 0 a count), a state matching
 * any row; h_at_least (m flags, or NULL) makes a mode's count a floor (at
 * least c).  mask_n is set_mask's n (0: the input's photon count n; else at
 * least n).  Layer k keeps the states whose constrained-mode counts can still
 * reach some row with the mask_n - k photons to come (tests/utils/test_mask.py:
 * 33-45, partial match), so with mask_n > n the output holds the states that
 * photons from elsewhere can complete (the selection rule of
 * pcvl_fock_mask_select).  The final layer may then hold several herald
 * configurations; execute writes the kept states in FSArray order (a final
 * reordering pass, whose buffer pcvl_slos_plan_info's work_bytes includes).
 * Needs at least two modes that no row constrains, 1..30 photons, at most 64
 * rows and 10^6 herald configurations. */
int pcvl_slos_plan_create_masks(int m, const int32_t *h_in_occ, int nrows, const int32_t *h_conds,
                                const int32_t *h_at_least, int mask_n, pcvl_slos_plan *out);

/* Several input states in one plan: the _Path tree of
 * SLOSBackend.preprocess([inputs]) (perceval/backends/_slos.py:37-86,170-185),
 * so layers on a shared photon-addition prefix are computed once.  h_inputs
 * holds count rows of m occupations.  Outputs are concatenated in input
 * order: input i occupies elements [h_out_offsets[i], h_out_offsets[i+1]) of
 * d_out (complex for COEFS, double for PROBS, FSArray order each).  The
 * workspace holds one layer per depth below the largest photon number. */
typedef struct pcvl_slos_tree_s *pcvl_slos_tree;
int pcvl_slos_tree_create(int m, const int32_t *h_inputs, int count, pcvl_slos_tree *out);
/* h_out_offsets receives count + 1 entries; any output pointer may be NULL. */
int pcvl_slos_tree_info(pcvl_slos_tree plan, int *nmax, int64_t *h_out_offsets, size_t *work_bytes,
                        int64_t *layer_launches);
int pcvl_slos_tree_execute(pcvl_slos_tree plan, const double *d_U, int mode, void *d_out, void *d_work,
                           size_t work_bytes, hipStream_t stream);
int pcvl_slos_tree_destroy(pcvl_slos_tree plan);

/* Layer kernel used by execute (the same parents in the same summation
 * order, so every kernel gives bit-identical results):
 *   0  auto: 8 where its layout was built (m >= 2, 1..30 photons, every layer
 *      below 2^32 states), else 6 where it fits, else 4 where it fits, else 1;
 *   1  one thread per output state, 64-bit ranks (any plan);
 *   4  one thread per output state, the unranking walk and the parent gather
 *      in two phases (parent layers below 2^26 states, LDS below 64 KiB);
 *   6  runs of three consecutive states per thread, successor steps instead
 *      of walks, gather transposed through LDS (as 4, plus LDS below 64 KiB);
 *   8  prefix/suffix blocks: one scalar walk per block of states sharing
 *      their first m - w modes, the last w modes from a table per suffix
 *      photon count (pcvl_slos_plan_set_block_width).
 * Kernels not applicable to the plan give -1.  For comparisons. */
int pcvl_slos_plan_set_kernel(pcvl_slos_plan plan, int kernel);

/* Suffix width w of kernel 8 (0: the default, min(16, m - 1, max(12, m / 2))
 * narrowed until its tables stay under 512 MiB; else 1 <= w <= min(m - 1,
 * 16)).  Rebuilds
 * the plan's block tables (synchronises the device).  A tuning knob: the
 * results do not depend on w.  No counterpart in the reference. */
int pcvl_slos_plan_set_block_width(pcvl_slos_plan plan, int w);
/* Current suffix width of kernel 8, 0 when the plan has no block layout. */
int pcvl_slos_plan_block_width(pcvl_slos_plan plan);

/* ---------------------------------------------------------------------------
 * Exact boson sampling -- replaces xq.Clifford2017.samples
 * (perceval/backends/_clifford2017.py:52-57): algorithm A of Clifford &
 * Clifford 2017, count samples drawn in parallel.  Sample s uses the
 * counter-based random stream (seed, first_sample + s), so a call is fully
 * determined by its arguments.  d_out: count rows of m int32 occupation
 * counts.  At most 30 photons and 256 modes.
 * ------------------------------------------------------------------------- */

/* Device workspace needed by pcvl_cc_samples for n photons and count samples. */
int pcvl_cc_workspace_bytes(int n, int64_t count, size_t *bytes);

/* Enqueued on `stream`, which is synchronised before returning. */
int pcvl_cc_samples(const double *d_U, int m, const int32_t *h_in_occ, int64_t count, uint64_t seed,
                    uint64_t first_sample, int32_t *d_out, void *d_work, size_t work_bytes,
                    hipStream_t stream);

/* Samples of several input states in one call, for SamplesProvider.prepare
 * (perceval/simulators/noisy_sampling_simulator.py:55-74), which fills a
 * sample pool per input state: h_inputs holds ninputs rows of m occupations,
 * h_counts the samples wanted from each.  Rows of d_out (sum of counts, m
 * int32 each) are grouped by input in input order.  Sample s of input i is
 * global sample row0_i + s and uses the counter stream (seed, first_sample +
 * row0_i + s): the same draws as pcvl_cc_samples on input i with first_sample
 * + row0_i.  Inputs of equal photon number share one launch per photon step.
 * Allocates its workspace stream-ordered and synchronises `stream`. */
int pcvl_cc_samples_multi(const double *d_U, int m, const int32_t *h_inputs, const int64_t *h_counts, int ninputs,
                          uint64_t seed, uint64_t first_sample, int32_t *d_out, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Phase gradients -- no counterpart in the reference, which differentiates
 * numerically by repeated SLOSBackend.set_circuit/all_prob calls
 * (benchmark/benchmark_QML-DE-solver.py:140-142; perceval/backends/_slos.py:130-141).
 * For each of nsets phase vectors (d_phases: nsets x nparams) of one mesh:
 *   loss = sum over targets of P(s), P(s) = |perm(U[s, in])|^2 / (prod s! prod t!),
 *   grad = d loss / d phases  (nsets x nparams),
 *   probs = P(s) per target (nsets x ntargets; d_probs may be NULL).
 * The mesh program lists nops components in circuit order (U = L_K ... L_1):
 *   h_kind[p] = 0: phase shifter on mode h_mode[p], phase = phases[h_param[p]],
 *                  or h_const[8p] when h_param[p] = -1;
 *   h_kind[p] = 1: fixed 2x2 unitary on modes (h_mode[p], h_mode[p]+1),
 *                  h_const[8p..8p+7] row-major interleaved complex.
 * h_targets: ntargets rows of m int32.  m <= 32, 1..8 photons, <= 512 targets
 * (run in blocks of 64 whose cofactors sit in LDS); -4 when the program,
 * index tables and target norms would take more than 160 KiB of LDS (a
 * caller then splits the targets over several plans and adds the results).
 * Synchronises `stream`.
 * ------------------------------------------------------------------------- */
int pcvl_mesh_adjoint(const int32_t *h_kind, const int32_t *h_mode, const int32_t *h_param,
                      const double *h_const, int nops, int m, const double *d_phases, int nparams,
                      int nsets, const int32_t *h_in_occ, const int32_t *h_targets, int ntargets,
                      double *d_loss, double *d_grad, double *d_probs, hipStream_t stream);

/* The same as a plan: pcvl_mesh_plan_create validates and uploads the mesh
 * program, the input state and the targets once (allocates and synchronises);
 * pcvl_mesh_plan_execute then only enqueues the kernel for nsets phase vectors
 * (no allocation, no synchronisation; graph-capturable), for training loops
 * that evaluate the same circuit and targets at new phases every step.
 * pcvl_mesh_adjoint is create + execute + synchronise + destroy. */
typedef struct pcvl_mesh_plan_s *pcvl_mesh_plan;
int pcvl_mesh_plan_create(const int32_t *h_kind, const int32_t *h_mode, const int32_t *h_param,
                          const double *h_const, int nops, int m, int nparams, const int32_t *h_in_occ,
                          const int32_t *h_targets, int ntargets, pcvl_mesh_plan *out);
int pcvl_mesh_plan_execute(pcvl_mesh_plan plan, const double *d_phases, int nsets, double *d_loss,
                           double *d_grad, double *d_probs, hipStream_t stream);
/* Weights w_s of the plan's targets (ntargets doubles, finite; NULL: all 1):
 * execute then returns loss = sum_s w_s P(s) and its gradient, the objective
 * of a diagonal observable (e.g. sum over every output state of lambda_s
 * P(s), as benchmark/benchmark_QML-DE-solver.py:81-84 evaluates); probs stay
 * the unweighted P(s).  Copies synchronously; not concurrent with an
 * execution of the plan. */
int pcvl_mesh_plan_set_weights(pcvl_mesh_plan plan, const double *h_weights);
int pcvl_mesh_plan_destroy(pcvl_mesh_plan plan);

/* ---------------------------------------------------------------------------
 * Learned Metropolis independence sampler -- no counterpart in the reference
 * (whose sampler is xq.Clifford2017, perceval/backends/_clifford2017.py:52-57).
 * A proposal network (bf16 weights, fp32 arithmetic) proposes output states
 * as non-decreasing mode sequences; FP64 permanents give the exact target
 * probabilities, so the chain's stationary law is the boson-sampling law.
 * All buffers are caller-owned device memory; matrices are row-major.
 * ------------------------------------------------------------------------- */

/* Y[i][j] = relu?(sum_k A(i,k) B(k,j) + bias[j]) (+ Y[i][j] when accumulate),
 * A(i,k) = d_A[i*sai + k*sak], B(k,j) = d_B[k*sbk + j*sbj]; B is bf16 when
 * b_is_bf16, else fp32.  d_bias may be NULL. */
int pcvl_mlp_gemm(const float *d_A, int64_t sai, int64_t sak, const void *d_B, int b_is_bf16,
                  int64_t sbk, int64_t sbj, float *d_Y, int64_t ldy, int M, int N, int K,
                  const float *d_bias, int relu, int accumulate, hipStream_t stream);

/* dH[e] = 0 where H[e] <= 0 (ReLU backward, in place). */
int pcvl_mlp_relu_bwd(float *d_dH, const float *d_H, int64_t count, hipStream_t stream);

/* d_db[j] = sum_i dY[i][j] for an M x N matrix. */
int pcvl_mlp_colsum(const float *d_dY, int M, int N, float *d_db, hipStream_t stream);

/* Weight and bias gradients over R rows: d_dW (Mo x No) = dY^T X and d_db
 * (Mo) = colsum(dY), dY R x Mo (row stride ldd), X R x No (row stride ldx).
 * Rows are split over blocks in a fixed partition and the partials summed
 * in order (deterministic).  d_dW or d_db may be NULL; d_work holds
 * pcvl_mlp_wgrad_workspace(R, Mo, No) floats. */
int pcvl_mlp_wgrad_workspace(int R, int Mo, int No, int64_t *floats);
int pcvl_mlp_wgrad(const float *d_dY, int64_t ldd, const float *d_X, int64_t ldx, int R, int Mo, int No,
                   float *d_dW, float *d_db, float *d_work, int64_t work_floats, hipStream_t stream);

/* Masked softmax cross-entropy per row (modes below d_lo[i] masked):
 * loss[i] = -log q(target[i]); d_dlogits (may be NULL) = (q - onehot) * scale. */
int pcvl_mlp_ce(const float *d_logits, int M, int m, const int32_t *d_target, const int32_t *d_lo,
                float scale, float *d_loss, float *d_dlogits, hipStream_t stream);

/* One Adam step on fp32 master weights (step counts from 1); refreshes the
 * bf16 copy when d_wbf16 is not NULL. */
int pcvl_mlp_adam(float *d_w, const float *d_g, float *d_m1, float *d_m2, uint16_t *d_wbf16,
                  int64_t count, float lr, float beta1, float beta2, float eps, int step,
                  hipStream_t stream);

/* bf16 (round to nearest even) copy of fp32 weights. */
int pcvl_mlp_to_bf16(const float *d_w, uint16_t *d_wbf16, int64_t count, hipStream_t stream);

/* Inputs of step p for B sequences (B x n int32): counts of the first p modes
 * and a one-hot of p (d_x: B x (m+n) fp32); d_lo[b] = seq[b][p-1] (0 at p=0). */
int pcvl_mis_features(const int32_t *d_seq, int B, int n, int m, int p, float *d_x, int32_t *d_lo,
                      hipStream_t stream);

/* Teacher-forcing inputs of every (sequence, step) pair, row b*n+p, with
 * d_lo and the target mode seq[b][p]. */
int pcvl_mis_features_all(const int32_t *d_seq, int B, int n, int m, float *d_x, int32_t *d_lo,
                          int32_t *d_target, hipStream_t stream);

/* Draws step p of every sequence from its masked softmax (uniform stream
 * (seed, first + b), counter p) and accumulates log q in d_logq. */
int pcvl_mis_sample_step(const float *d_logits, int B, int m, int n, int p, const int32_t *d_lo,
                         uint64_t seed, uint64_t first, int32_t *d_seq, double *d_logq,
                         hipStream_t stream);

/* All n steps of B proposals in one launch: the same network (W1 H x (m+n),
 * W2 m x H, bf16; biases fp32) and the same uniform (seed, first + b, p) for
 * step p of proposal b as the stepwise path, with the first layer's
 * pre-activation updated in place instead of recomputed.  Writes the
 * sequences (B x n) and the exact log q of each.  m <= 64, 8 <= H <= 256,
 * H % 8 == 0. */
int pcvl_mis_sample_fused(const uint16_t *d_W1, const float *d_b1, const uint16_t *d_W2, const float *d_b2,
                          int B, int m, int n, int H, uint64_t seed, uint64_t first, int32_t *d_seq,
                          double *d_logq, hipStream_t stream);

/* n x n submatrices U[seq[b], cols] of B proposals (complex, packed). */
int pcvl_mis_submatrices(const double *d_U, int m, const int32_t *d_cols, int n, const int32_t *d_seq,
                         int64_t B, double *d_out, hipStream_t stream);

/* Permanents of U[seq[b], cols] for B sorted mode sequences (B x n int32, n <= 28),
 * with repeated rows summed by multiplicity (a group of k equal rows costs
 * k+1 Gray terms instead of 2^k).  d_out: B complex. */
int pcvl_mis_permanents(const double *d_U, int m, const int32_t *d_cols, int n, const int32_t *d_seq,
                        int64_t B, double *d_out, hipStream_t stream);

/* log p of B proposals from their permanents (complex) and sequences. */
int pcvl_mis_logp(const double *d_perms, const int32_t *d_seq, int n, int64_t B, double log_prod_in,
                  double *d_logp, hipStream_t stream);

/* Accept/reject scan over B proposals with log weights log p - log q,
 * starting from a state of log weight logw0; d_state[t] = proposal index held
 * after step t (-1: the initial state); *d_accepted = accepted moves.  Step t
 * is taken iff lw_t - cur >= 0 or log u_t < lw_t - cur, u_t the uniform of
 * stream (seed, first + t).  B <= 2^36; 16 B bytes of workspace come from the
 * library's stream-ordered pool. */
int pcvl_mis_chain(const double *d_logw, int64_t B, double logw0, uint64_t seed, uint64_t first,
                   int64_t *d_state, int64_t *d_accepted, hipStream_t stream);

/* Occupation counts (B x m int32) of sorted mode sequences (B x n int32). */
int pcvl_mis_seq_to_occ(const int32_t *d_seq, int64_t B, int n, int m, int32_t *d_occ, hipStream_t stream);

/* Sorted mode sequences (B x n) of occupation vectors (B x m); a row whose
 * photon number differs from n becomes -1s. */
int pcvl_mis_occ_to_seq(const int32_t *d_occ, int64_t B, int m, int n, int32_t *d_seq, hipStream_t stream);

/* Chain output: d_out[t] = proposal d_state[t] of d_prop, or d_x0 for -1. */
int pcvl_mis_gather(const int32_t *d_prop, const int64_t *d_state, int64_t B, int n, const int32_t *d_x0,
                    int32_t *d_out, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * HIP graphs of plan executions.  No counterpart in the reference (its
 * arithmetic has no device launches); for callers that replay a fixed
 * workload -- the same plan on the same device buffers -- with one launch
 * call instead of host work per kernel.  Capture runs on a private stream;
 * the graph reads and writes the buffers given here at every launch, so they
 * must outlive it.
 * ------------------------------------------------------------------------- */
typedef struct pcvl_graph_s *pcvl_graph;

/* `repeats` (1..1024) executions of a permanent plan on fixed buffers. */
int pcvl_perm_graph_create(pcvl_perm_plan plan, const double *d_mats, double *d_out, int repeats,
                           pcvl_graph *out);
/* `repeats` executions of a SLOS plan (arguments as pcvl_slos_plan_execute). */
int pcvl_slos_graph_create(pcvl_slos_plan plan, const double *d_U, int mode, double *d_out, void *d_work,
                           size_t work_bytes, int repeats, pcvl_graph *out);
/* Enqueue the whole graph on `stream`. */
int pcvl_graph_launch(pcvl_graph graph, hipStream_t stream);
int pcvl_graph_destroy(pcvl_graph graph);

/* ---------------------------------------------------------------------------
 * Approximate permanents (NaiveApprox backend).
 * ------------------------------------------------------------------------- */

/* Gurvits estimates of a batch of matrices (same packing as the permanent
 * plans, orders 0..48), replacing xq.estimate_permanent_cx(M, iterations)
 * (perceval/backends/_naive_approx.py:52-62).  d_est receives one complex
 * mean per matrix and d_err its 99% error bound (normal approximation,
 * 2.5758 * sample standard deviation / sqrt(iters)).  Sample s of matrix b
 * takes its signs from the counter-based stream (seed, stream0 + b, s).
 * Copies the small item table to the device and synchronises `stream`. */
int pcvl_gurvits_batch(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                       int64_t iters, uint64_t seed, uint64_t stream0, double *d_est, double *d_err,
                       hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Measurement helper (no counterpart in the reference).
 * ------------------------------------------------------------------------- */

/* Enqueue `blocks` x 256 lanes of pure FP64 FMA chains (`iters` rounds of 16
 * chains per lane); *flops receives the flops issued (FMA = 2).  d_out holds
 * blocks*256 doubles.  Timing it gives the attainable vector FP64 rate. */
int pcvl_probe_fp64_fma(int blocks, int64_t iters, double *d_out, int64_t *flops, hipStream_t stream);

/* Measurement helper, no counterpart in the reference: whether the FP64
 * matrix cores run beside the vector ALU.  blocks x 256 threads, iters loop
 * trips; mode 0: four v_mfma_f64_16x16x4f64 accumulators per wave, 1: 16
 * v_fma_f64 chains per lane, 2: both in the same loop.  d_out: blocks*256
 * doubles; *flops_mfma / *flops_valu receive each unit's flop count. */
int pcvl_probe_fp64_mix(int blocks, int64_t iters, int mode, double *d_out, int64_t *flops_mfma,
                        int64_t *flops_valu, hipStream_t stream);

/* Measurement helper, no counterpart in the reference: the read rate of a
 * device buffer of `bytes` bytes read `reps` times over by blocks x 256
 * threads (16-byte loads).  d_out: blocks*256 doubles. */
int pcvl_probe_read(const double *d_buf, uint64_t bytes, int reps, int blocks, double *d_out, hipStream_t stream);

/* Measurement helper, no counterpart in the reference: the XCD that runs
 * each of `blocks` workgroups (its XCC_ID hardware register), into
 * d_out[0..blocks).  The persistent SLOS launches assume XCD = blockIdx % 8. */
int pcvl_probe_xcc(int blocks, int32_t *d_out, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PCVL_HIP_H */
