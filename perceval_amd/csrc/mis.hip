// Metropolis independence sampler (MIS) for boson sampling: target
// probabilities of proposed states by FP64 permanents, and the accept/reject
// scan that keeps the chain exact.
//
// No counterpart in the reference (its sampler is Clifford2017,
// perceval/backends/_clifford2017.py:52-57).  Proposals are non-decreasing
// mode sequences (one per Fock state) drawn from the proposal network with
// log-probability log q.  For a proposal s with input modes cols:
//     log p(s) = log |perm(U[s, cols])|^2 - log prod_j s_j! - log prod_i t_i!,
// and the chain moves from x to the proposal y with probability
// min(1, p(y) q(x) / (p(x) q(y))); with exact p and q its stationary law is
// exactly p.
#include "common.h"
#include "../../include/pcvl_hip.h"

namespace pcvl {

// Packs the n x n submatrix U[seq[b], cols] of every proposal b (row r =
// output photon r of the sorted sequence, column c = input photon c).
__global__ void mis_submatrix_kernel(const cd *__restrict__ U, int m, const int32_t *__restrict__ cols, int n,
                                     const int32_t *__restrict__ seq, int64_t B, cd *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = (int64_t)n * n;
    if (e >= B * nn) return;
    const int64_t b = e / nn;
    const int rc = (int)(e % nn);
    const int r = rc / n, c = rc % n;
    const int j = seq[b * n + r];
    out[e] = (j >= 0 && j < m) ? U[(size_t)j * m + cols[c]] : cd{0.0, 0.0};   // invalid rows give perm 0
}

// log p of every proposal from its permanent; repeated modes of the sorted
// sequence give prod s_j!.
__global__ void mis_logp_kernel(const cd *__restrict__ perms, const int32_t *__restrict__ seq, int n, int64_t B,
                                double log_prod_in, double *__restrict__ logp)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const cd a = perms[b];
    double lf = 0.0;
    int run = 1;
    for (int r = 1; r < n; r++) {
        if (seq[b * n + r] == seq[b * n + r - 1]) {
            run++;
            lf += log((double)run);
        } else {
            run = 1;
        }
    }
    const double a2 = a.r * a.r + a.i * a.i;
    logp[b] = (a2 > 0.0 ? log(a2) : -INFINITY) - lf - log_prod_in;
}

// Sequential accept/reject over B proposals (one lane; the scan is O(B) and
// memory-light).  state[t] = index of the proposal the chain holds after step
// t, or -1 for the initial state.  Uniform of step t: stream (seed, first + t),
// counter 0.  Returns the number of accepted moves in *accepted.
__global__ void mis_chain_kernel(const double *__restrict__ logw, int64_t B, double logw0, uint64_t seed,
                                 uint64_t first, int64_t *__restrict__ state, int64_t *__restrict__ accepted)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double cur = logw0;
    int64_t idx = -1, acc = 0;
    for (int64_t t = 0; t < B; t++) {
        const double lw = logw[t];
        const double d = lw - cur;
        bool take = d >= 0.0;
        if (!take) {
            const double u = uniform01(seed, first + (uint64_t)t, 0);
            take = log(u) < d;
        }
        if (take) {
            cur = lw;
            idx = t;
            acc++;
        }
        state[t] = idx;
    }
    *accepted = acc;
}

// Occupation counts of sorted sequences (B x n) -> (B x m).
__global__ void mis_seq_to_occ_kernel(const int32_t *__restrict__ seq, int64_t B, int n, int m,
                                      int32_t *__restrict__ occ)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int32_t *o = occ + b * m;
    for (int j = 0; j < m; j++) o[j] = 0;
    for (int r = 0; r < n; r++) {
        const int j = seq[b * n + r];
        if (j >= 0 && j < m) o[j] += 1;
    }
}

// Sorted mode sequences of occupation vectors (B x m) -> (B x n); rows whose
// photon number differs from n are filled with -1.
__global__ void mis_occ_to_seq_kernel(const int32_t *__restrict__ occ, int64_t B, int m, int n,
                                      int32_t *__restrict__ seq)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int tot = 0;
    for (int j = 0; j < m; j++) tot += occ[b * m + j];
    int r = 0;
    if (tot == n) {
        for (int j = 0; j < m; j++)
            for (int v = 0; v < occ[b * m + j]; v++) seq[b * n + r++] = j;
    }
    for (; r < n; r++) seq[b * n + r] = -1;
}

// Chain output: sequence held after step t (proposal state[t], or the
// initial sequence x0 when state[t] = -1).
__global__ void mis_gather_kernel(const int32_t *__restrict__ prop, const int64_t *__restrict__ state, int64_t B,
                                  int n, const int32_t *__restrict__ x0, int32_t *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= B * n) return;
    const int64_t t = e / n;
    const int r = (int)(e % n);
    const int64_t k = state[t];
    out[e] = k < 0 ? x0[r] : prop[k * n + r];
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_mis_seq_to_occ(const int32_t *d_seq, int64_t B, int n, int m, int32_t *d_occ, hipStream_t stream)
{
    if (B <= 0) return OK;
    hipLaunchKernelGGL(mis_seq_to_occ_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream, d_seq, B, n, m,
                       d_occ);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_occ_to_seq(const int32_t *d_occ, int64_t B, int m, int n, int32_t *d_seq, hipStream_t stream)
{
    if (B <= 0) return OK;
    hipLaunchKernelGGL(mis_occ_to_seq_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream, d_occ, B, m, n,
                       d_seq);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_gather(const int32_t *d_prop, const int64_t *d_state, int64_t B, int n, const int32_t *d_x0,
                               int32_t *d_out, hipStream_t stream)
{
    if (B <= 0) return OK;
    const int64_t tot = B * n;
    hipLaunchKernelGGL(mis_gather_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, d_prop, d_state,
                       B, n, d_x0, d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_submatrices(const double *d_U, int m, const int32_t *d_cols, int n, const int32_t *d_seq,
                                    int64_t B, double *d_out, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && n >= 1 && B >= 0, "bad sizes");
    if (B == 0) return OK;
    const int64_t tot = B * n * n;
    hipLaunchKernelGGL(mis_submatrix_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream,
                       reinterpret_cast<const cd *>(d_U), m, d_cols, n, d_seq, B, reinterpret_cast<cd *>(d_out));
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_logp(const double *d_perms, const int32_t *d_seq, int n, int64_t B, double log_prod_in,
                             double *d_logp, hipStream_t stream)
{
    PCVL_ARG_CHECK(n >= 1 && B >= 0, "bad sizes");
    if (B == 0) return OK;
    hipLaunchKernelGGL(mis_logp_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream,
                       reinterpret_cast<const cd *>(d_perms), d_seq, n, B, log_prod_in, d_logp);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_chain(const double *d_logw, int64_t B, double logw0, uint64_t seed, uint64_t first,
                              int64_t *d_state, int64_t *d_accepted, hipStream_t stream)
{
    PCVL_ARG_CHECK(B >= 0, "bad size");
    if (B == 0) return OK;
    hipLaunchKernelGGL(mis_chain_kernel, dim3(1), dim3(64), 0, stream, d_logw, B, logw0, seed, first, d_state,
                       d_accepted);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}
