// Exact boson sampling: algorithm A of Clifford & Clifford (2017), batched
// over samples.  Replaces xq.Clifford2017.samples
// (perceval/backends/_clifford2017.py:52-57; docs/source/backends.rst:108-116).
//
// For an input with photons in modes cols[0..n-1] let A = U[:, cols] (m x n).
// A sample permutes the columns of A uniformly, then for k = 1..n draws the
// output mode r_k with probability proportional to
//     w_i = | sum_{l<k} A[i, perm[l]] * perm(B^{(-l)}) |^2,
//     B = A[r_1..r_{k-1}, perm[0..k-1]]  ((k-1) x k),
// where B^{(-l)} drops column l.  All k sub-permanents come from one Glynn
// pass over the rows of B: per Gray-code term, the k column sums c_l are
// shared and prod_{l' != l} c_l' is the product of a prefix and a suffix.
//
// Mapping: one launch per photon step k, specialised on k so that each step
// gets its own register allocation; within a launch one wavefront owns one
// sample (persistent over samples).  The 2^{k-2} Gray-code terms are split
// over the 32 lane pairs; partial sums of the k sub-permanents are reduced
// across the wave, then lane i computes the weight of candidate row i and a
// wave prefix sum draws r_k.  Per-sample state (permuted columns, drawn modes)
// lives in a caller-provided workspace.  Random numbers
// are the counter-based uniforms of common.h: stream = sample index, counters
// 0..n-2 for the column permutation (Fisher-Yates, i = n-1 down to 1) and
// n-1+k-1 for the k-th photon, so a run is fixed by its seed.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <vector>

namespace pcvl {

constexpr int CC_MAX_N = 30;
constexpr int CC_MAX_M = 256;   // multiple of 64

// Per-step shared state, sized for step K: the LDS a step takes (and with it
// the occupancy the compiler plans its registers for) does not grow with the
// largest photon count the library accepts.
template <int K>
struct CCShared {
    cd B[(K > 1 ? K - 1 : 1) * K];       // B[i*k + l] = 2 * U[r_i, cols[perm[l]]]
    cd sub[K];
    int pc[K];                           // cols[perm[l]]: input mode of permuted column l
    int r[K];
};

// Partial sums of the k sub-permanents over this lane pair's chunk of
// Gray-code terms (B holds 2*A rows; rows = k-1 >= 1).  The k columns are cut
// into halves L = [0, h) and H = [h, k) with h = k/2; the even lane of a pair
// owns L, the odd lane owns H, each padded to hmax = k - h columns (a padding
// column has column sum 1 and never flips).  Per term each lane forms the
// prefix products of its half and the half total T; the partners swap T, and
// prod_{l' != l} c_l' = prefix(l) * suffix(l) * T_partner comes out of one
// backward pass.  acc[j] is this lane's partial for column c0 + j.  With
// PLOG2 > 0 the 2^{rows-1} terms are cut into 2^PLOG2 parts and the wave
// takes part `part` only (terms part * 2^{rows-1-PLOG2} onwards).
template <int K, int PLOG2 = 0>
__device__ __forceinline__ void subperm_pair(const cd *__restrict__ B, int lane, cd *acc, uint32_t part = 0)
{
    constexpr int rows = K - 1;
    constexpr int h = K / 2;
    constexpr int hmax = K - h;
    constexpr int total_log2 = rows - 1 - PLOG2;          // 2^{rows-1-PLOG2} terms in the part
    static_assert(total_log2 >= 0, "more parts than Gray-code terms");
    constexpr int c = total_log2 > 5 ? total_log2 - 5 : 0; // 32 pairs
    constexpr int active_pairs = total_log2 > 5 ? 32 : (1 << total_log2);
    const int half = lane & 1;
    const int pair = lane >> 1;
    const int c0 = half ? h : 0;
    const int ncol = half ? hmax : h;    // columns this lane really owns
#pragma unroll
    for (int j = 0; j < hmax; j++) acc[j] = {0.0, 0.0};
    if (pair >= active_pairs) return;
    const uint32_t steps = 1u << c;
    const uint32_t base = (part << total_log2) + ((uint32_t)pair << c);
    const uint32_t g0 = base ^ (base >> 1);

    // For odd K the even lane's last column is padding: it reads a valid
    // column of B (column h) but scales it by 0 and starts at 1, so the
    // code has no lane-divergent branches.
    const double pad = (hmax - 1 < ncol) ? 1.0 : 0.0;
    cd cs[hmax];
#pragma unroll
    for (int j = 0; j < hmax; j++) {
        const double f = (j == hmax - 1) ? pad : 1.0;
        const cd b = B[c0 + j];
        cs[j] = {fma(0.5 * f, b.r, 1.0 - f), 0.5 * f * b.i};
    }
    for (int i = 1; i < rows; i++) {
        const double d = ((g0 >> (i - 1)) & 1) ? -0.5 : 0.5;
#pragma unroll
        for (int j = 0; j < hmax; j++) {
            const double dj = (j == hmax - 1) ? d * pad : d;
            const cd b = B[i * K + c0 + j];
            cs[j].r = fma(dj, b.r, cs[j].r);
            cs[j].i = fma(dj, b.i, cs[j].i);
        }
    }
    for (uint32_t u = 0; u < steps; u++) {
        const uint32_t t = base + u;
        if (u) {
            const int b = __builtin_ctz(t);
            const double s = (((t >> b) ^ (t >> (b + 1))) & 1) ? -1.0 : 1.0;
            const double sl = s * pad;
            const cd *row = B + (b + 1) * K + c0;
#pragma unroll
            for (int j = 0; j < hmax; j++) {
                const cd a = row[j];
                const double sj = (j == hmax - 1) ? sl : s;
                cs[j].r = fma(sj, a.r, cs[j].r);
                cs[j].i = fma(sj, a.i, cs[j].i);
            }
        }
        cd P[hmax];
        P[0] = {1.0, 0.0};
        if constexpr (hmax > 1) P[1] = cs[0];
#pragma unroll
        for (int j = 2; j < hmax; j++) P[j] = cmul(P[j - 1], cs[j - 1]);
        const cd T = (hmax == 1) ? cs[0] : cmul(P[hmax - 1], cs[hmax - 1]);
        cd Tp;
        Tp.r = __shfl_xor(T.r, 1, 64);
        Tp.i = __shfl_xor(T.i, 1, 64);
        cd S = (t & 1) ? cd{-Tp.r, -Tp.i} : Tp;
#pragma unroll
        for (int j = hmax - 1; j >= 1; j--) {
            cfma_acc(acc[j], P[j], S);
            S = cmul(S, cs[j]);
        }
        acc[0].r += S.r;
        acc[0].i += S.i;
        // Keep the next column-sum update below this term's products (no
        // overlap of two generations of cs; see perm_glynn.hip).
#pragma unroll
        for (int j = 0; j < hmax; j++)
            asm volatile("" : "+v"(cs[j].r), "+v"(cs[j].i) : "v"(acc[0].r), "v"(acc[0].i));
    }
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the lanes of one parity (the 32 pairs).
__device__ __forceinline__ double pair_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 2; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int K>
__device__ __forceinline__ void subperms(CCShared<K> &sh, int lane)
{
    constexpr int h = K / 2;
    constexpr int hmax = K - h;
    cd acc[hmax];
    subperm_pair<K>(sh.B, lane, acc);
    constexpr int rows = K - 1;
    const double scale = ldexp(1.0, -(rows - 1));
#pragma unroll
    for (int j = 0; j < hmax; j++) {
        const double re = pair_sum_d(acc[j].r), im = pair_sum_d(acc[j].i);
        if (lane == 0 && j < h) sh.sub[j] = {re * scale, im * scale};
        if (lane == 1) sh.sub[h + j] = {re * scale, im * scale};
    }
}

// Column permutation of each sample (Fisher-Yates, counters 0..n-2):
// pc[s*n + l] = input mode of permuted column l.
__global__ void cc_init_kernel(const int32_t *__restrict__ cols, int n, int64_t count, uint64_t seed, uint64_t first,
                               int32_t *__restrict__ pc)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    int perm[CC_MAX_N];
    for (int l = 0; l < n; l++) perm[l] = l;
    uint64_t ctr = 0;
    for (int i = n - 1; i >= 1; i--) {
        const double u = uniform01(seed, first + (uint64_t)s, ctr++);
        int j = (int)(u * (i + 1));
        if (j > i) j = i;
        const int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    for (int l = 0; l < n; l++) pc[s * n + l] = cols[perm[l]];
}

// Several input states with the same photon number: sample s of the batch
// is global sample gidx[s] (its counter stream) of input inp[s], whose
// photon columns are cols[inp[s] * n ...].
__global__ void cc_init_multi_kernel(const int32_t *__restrict__ cols, const int32_t *__restrict__ inp,
                                     const int64_t *__restrict__ gidx, int n, int64_t count, uint64_t seed,
                                     uint64_t first, int32_t *__restrict__ pc)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    const int32_t *c = cols + (size_t)inp[s] * n;
    int perm[CC_MAX_N];
    for (int l = 0; l < n; l++) perm[l] = l;
    uint64_t ctr = 0;
    const uint64_t stream = first + (uint64_t)gidx[s];
    for (int i = n - 1; i >= 1; i--) {
        const double u = uniform01(seed, stream, ctr++);
        int j = (int)(u * (i + 1));
        if (j > i) j = i;
        const int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    for (int l = 0; l < n; l++) pc[s * n + l] = c[perm[l]];
}

// Output rows of a multi-input batch: sample s goes to row gidx[s].
__global__ void cc_finish_multi_kernel(const int32_t *__restrict__ r_all, const int64_t *__restrict__ gidx, int m,
                                       int n, int64_t count, int32_t *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count * m) return;
    const int64_t s = e / m;
    const int i = (int)(e % m);
    int v = 0;
    for (int k = 0; k < n; k++) v += (r_all[s * n + k] == i) ? 1 : 0;
    out[gidx[s] * m + i] = v;
}

// Draw of photon K of sample s from the sub-permanents in sh.sub: the weight
// of every candidate row i (lane i of each 64-row chunk), then an inclusive
// scan over rows in order picks r_k.
template <int K>
__device__ __forceinline__ void cc_draw(const CCShared<K> &sh, const cd *__restrict__ U, int m, int n, int64_t s,
                                        uint64_t seed, uint64_t stream, int lane, int32_t *__restrict__ r_all)
{
    constexpr int CH = CC_MAX_M / 64;
    double wv[CH];
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        const int i = c * 64 + lane;
        double w = 0.0;
        if (c * 64 < m && i < m) {
            cd a = {0.0, 0.0};
#pragma unroll
            for (int l = 0; l < K; l++) cfma_acc(a, U[(size_t)i * m + sh.pc[l]], sh.sub[l]);
            w = a.r * a.r + a.i * a.i;
        }
        wv[c] = w;
        total += wave_sum_d(w);
    }
    const double target = uniform01(seed, stream, (uint64_t)(n - 1) + (uint64_t)(K - 1)) * total;
    int pick = m - 1;
    double base = 0.0;
    bool found = false;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        if (!found && c * 64 < m) {
            const int i = c * 64 + lane;
            double incl = wv[c];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const double v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            const bool hit = (i < m) && (target < base + incl);
            const unsigned long long mask = __ballot(hit);
            if (mask) {
                pick = c * 64 + (int)__builtin_ctzll(mask);
                found = true;
            }
            base += __shfl(incl, 63, 64);
        }
    }
    if (lane == 0) r_all[s * n + (K - 1)] = pick;
}

// Step k (= K) of every sample: one wavefront per sample, persistent.
// r[s*n + i] holds the output mode of photon i (i < k-1 already drawn).
template <int K>
__global__ void __launch_bounds__(64) cc_step_kernel(const cd *__restrict__ U, int m, int n, int64_t count,
                                                     uint64_t seed, uint64_t first, const int32_t *__restrict__ pc_all,
                                                     int32_t *__restrict__ r_all, const int64_t *__restrict__ gidx)
{
    __shared__ CCShared<K> sh;
    const int lane = threadIdx.x;
    for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
        const uint64_t stream = first + (uint64_t)(gidx ? gidx[s] : s);
        if (lane < K) sh.pc[lane] = pc_all[s * n + lane];
        if (lane < K - 1) sh.r[lane] = r_all[s * n + lane];
        __syncthreads();
        if constexpr (K == 1) {
            if (lane == 0) sh.sub[0] = {1.0, 0.0};
        } else {
            // B = 2 * A[r_1..r_{k-1}, perm[0..k-1]] with A = U[:, cols]
            for (int e = lane; e < (K - 1) * K; e += 64) {
                const int i = e / K, l = e % K;
                const cd a = U[(size_t)sh.r[i] * m + sh.pc[l]];
                sh.B[e] = {2.0 * a.r, 2.0 * a.i};
            }
            __syncthreads();
            subperms<K>(sh, lane);
        }
        __syncthreads();
        cc_draw<K>(sh, U, m, n, s, seed, stream, lane, r_all);
        __syncthreads();
    }
}

// Steps above CC_SPLIT: the 2^{K-2} Gray-code terms of a sample are cut into
// 2^{K - CC_SPLIT} parts of 2^{CC_SPLIT - 2} terms, one wave per (sample,
// part), so that even one sample fills the device.  cc_part_kernel writes
// each part's pair-summed partial sub-permanents; cc_split_draw_kernel sums a
// sample's parts in part order, scales, and draws.  The cut depends on K
// only, so a sample's draws do not depend on how many samples a call holds.
constexpr int CC_SPLIT = 20;
constexpr size_t CC_PART_BYTES = size_t(256) << 20;   // partial sums of one chunk of samples

constexpr int cc_parts_log2(int k) { return k > CC_SPLIT ? k - CC_SPLIT : 0; }

// Samples of one chunk [s0, s0 + nchunk): item t = (sample s0 + t / P, part t % P).
template <int K>
__global__ void __launch_bounds__(64) cc_part_kernel(const cd *__restrict__ U, int m, int n, int64_t s0,
                                                     int64_t nchunk, const int32_t *__restrict__ pc_all,
                                                     const int32_t *__restrict__ r_all, cd *__restrict__ partials)
{
    constexpr int PLOG2 = cc_parts_log2(K);
    constexpr int64_t P = int64_t(1) << PLOG2;
    constexpr int h = K / 2;
    constexpr int hmax = K - h;
    __shared__ CCShared<K> sh;
    const int lane = threadIdx.x;
    for (int64_t t = blockIdx.x; t < nchunk * P; t += gridDim.x) {
        const int64_t sl = t >> PLOG2;
        const uint32_t part = (uint32_t)(t & (P - 1));
        const int64_t s = s0 + sl;
        if (lane < K) sh.pc[lane] = pc_all[s * n + lane];
        if (lane < K - 1) sh.r[lane] = r_all[s * n + lane];
        __syncthreads();
        for (int e = lane; e < (K - 1) * K; e += 64) {
            const int i = e / K, l = e % K;
            const cd a = U[(size_t)sh.r[i] * m + sh.pc[l]];
            sh.B[e] = {2.0 * a.r, 2.0 * a.i};
        }
        __syncthreads();
        cd acc[hmax];
        subperm_pair<K, PLOG2>(sh.B, lane, acc, part);
        cd *dst = partials + t * K;
#pragma unroll
        for (int j = 0; j < hmax; j++) {
            const double re = pair_sum_d(acc[j].r), im = pair_sum_d(acc[j].i);
            if (lane == 0 && j < h) dst[j] = {re, im};
            if (lane == 1) dst[h + j] = {re, im};
        }
        __syncthreads();
    }
}

template <int K>
__global__ void __launch_bounds__(64) cc_split_draw_kernel(const cd *__restrict__ U, int m, int n, int64_t s0,
                                                           int64_t nchunk, uint64_t seed, uint64_t first,
                                                           const int32_t *__restrict__ pc_all,
                                                           int32_t *__restrict__ r_all,
                                                           const int64_t *__restrict__ gidx,
                                                           const cd *__restrict__ partials)
{
    constexpr int64_t P = int64_t(1) << cc_parts_log2(K);
    __shared__ CCShared<K> sh;
    const int lane = threadIdx.x;
    const double scale = ldexp(1.0, -(K - 2));
    for (int64_t sl = blockIdx.x; sl < nchunk; sl += gridDim.x) {
        const int64_t s = s0 + sl;
        const uint64_t stream = first + (uint64_t)(gidx ? gidx[s] : s);
        if (lane < K) {
            sh.pc[lane] = pc_all[s * n + lane];
            cd v = {0.0, 0.0};
            const cd *src = partials + sl * P * K + lane;
            for (int64_t p = 0; p < P; p++) {   // in part order
                const cd x = src[p * K];
                v.r += x.r;
                v.i += x.i;
            }
            sh.sub[lane] = {v.r * scale, v.i * scale};
        }
        __syncthreads();
        cc_draw<K>(sh, U, m, n, s, seed, stream, lane, r_all);
        __syncthreads();
    }
}

// Occupation vector of each sample from its drawn modes.
__global__ void cc_finish_kernel(const int32_t *__restrict__ r_all, int m, int n, int64_t count,
                                 int32_t *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count * m) return;
    const int64_t s = e / m;
    const int i = (int)(e % m);
    int v = 0;
    for (int k = 0; k < n; k++) v += (r_all[s * n + k] == i) ? 1 : 0;
    out[e] = v;
}

// Samples per chunk of the split steps, and the partial sums a chunk needs.
static int64_t cc_chunk(int n)
{
    const int64_t per_sample = (int64_t(1) << cc_parts_log2(n)) * n * (int64_t)sizeof(cd);
    return std::max<int64_t>(1, (int64_t)CC_PART_BYTES / per_sample);
}

static size_t cc_partial_bytes(int n, int64_t count)
{
    if (n <= CC_SPLIT) return 0;
    return (size_t)std::min(count, cc_chunk(n)) * ((size_t)1 << cc_parts_log2(n)) * (size_t)n * sizeof(cd);
}

template <int K>
static int launch_step(int k, const cd *U, int m, int n, int64_t count, uint64_t seed, uint64_t first,
                       const int32_t *pc, int32_t *r, cd *partials, int cus, hipStream_t st,
                       const int64_t *gidx = nullptr)
{
    if constexpr (K > CC_MAX_N) {
        set_error("internal: photon step out of range");
        return ERR_ARG;
    } else {
        if (k != K) return launch_step<K + 1>(k, U, m, n, count, seed, first, pc, r, partials, cus, st, gidx);
        if constexpr (K <= CC_SPLIT) {
            int per_cu = 0;
            PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(&cc_step_kernel<K>), 64, 0));
            if (per_cu < 1) per_cu = 1;
            const int grid = (int)std::min<int64_t>(count, (int64_t)cus * per_cu);
            hipLaunchKernelGGL((cc_step_kernel<K>), dim3(grid), dim3(64), 0, st, U, m, n, count, seed, first, pc, r,
                               gidx);
            PCVL_HIP_CHECK(hipGetLastError());
        } else {
            constexpr int64_t P = int64_t(1) << cc_parts_log2(K);
            int per_cu_p = 0, per_cu_d = 0;
            PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu_p, reinterpret_cast<const void *>(&cc_part_kernel<K>), 64, 0));
            PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu_d, reinterpret_cast<const void *>(&cc_split_draw_kernel<K>), 64, 0));
            per_cu_p = std::max(per_cu_p, 1);
            per_cu_d = std::max(per_cu_d, 1);
            // chunks sized for the photon count n of the call (the largest
            // partial buffer of its steps), so the workspace holds any step's
            const int64_t chunk = cc_chunk(n);
            for (int64_t s0 = 0; s0 < count; s0 += chunk) {
                const int64_t nchunk = std::min(chunk, count - s0);
                const int gp = (int)std::min<int64_t>(nchunk * P, (int64_t)cus * per_cu_p);
                hipLaunchKernelGGL((cc_part_kernel<K>), dim3(gp), dim3(64), 0, st, U, m, n, s0, nchunk, pc, r,
                                   partials);
                PCVL_HIP_CHECK(hipGetLastError());
                const int gd = (int)std::min<int64_t>(nchunk, (int64_t)cus * per_cu_d);
                hipLaunchKernelGGL((cc_split_draw_kernel<K>), dim3(gd), dim3(64), 0, st, U, m, n, s0, nchunk, seed,
                                   first, pc, r, gidx, partials);
                PCVL_HIP_CHECK(hipGetLastError());
            }
        }
        return OK;
    }
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_cc_workspace_bytes(int n, int64_t count, size_t *bytes)
{
    PCVL_ARG_CHECK(bytes != nullptr, "bytes is null");
    PCVL_ARG_CHECK(n >= 0 && n <= CC_MAX_N && count >= 0, "n outside [0, 30] or count < 0");
    // pc (count*n) | r (count*n) | cols (CC_MAX_N), then from a 256-byte
    // boundary the partial sums of the split steps (more than 20 photons)
    const size_t head = sizeof(int32_t) * ((size_t)2 * (size_t)count * (size_t)std::max(n, 1) + (size_t)CC_MAX_N);
    const size_t part = cc_partial_bytes(n, count);
    *bytes = part ? ((head + 255) & ~size_t(255)) + part : head;
    return OK;
}

extern "C" int pcvl_cc_samples(const double *d_U, int m, const int32_t *h_in_occ, int64_t count, uint64_t seed,
                               uint64_t first_sample, int32_t *d_out, void *d_work, size_t work_bytes,
                               hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= CC_MAX_M, "mode count outside [1, 256]");
    PCVL_ARG_CHECK(h_in_occ != nullptr, "input state is null");
    PCVL_ARG_CHECK(count >= 0, "count < 0");
    std::vector<int32_t> cols;
    for (int i = 0; i < m; i++) {
        PCVL_ARG_CHECK(h_in_occ[i] >= 0, "negative occupation");
        for (int v = 0; v < h_in_occ[i]; v++) cols.push_back(i);
    }
    PCVL_ARG_CHECK((int)cols.size() <= CC_MAX_N, "more than 30 photons");
    if (count == 0) return OK;
    PCVL_ARG_CHECK(d_U && d_out, "null device buffer");
    const int n = (int)cols.size();
    size_t need = 0;
    pcvl_cc_workspace_bytes(n, count, &need);
    PCVL_ARG_CHECK(d_work != nullptr && work_bytes >= need, "workspace missing or smaller than pcvl_cc_workspace_bytes");
    if (n == 0) {
        const int64_t tot = count * m;
        hipLaunchKernelGGL(cc_finish_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream,
                           static_cast<const int32_t *>(d_work), m, 0, count, d_out);
        PCVL_HIP_CHECK(hipGetLastError());
        return OK;
    }
    // workspace: pc (count*n) | r (count*n) | cols (CC_MAX_N) | partial sums
    int32_t *pc = static_cast<int32_t *>(d_work);
    int32_t *r = pc + (size_t)count * n;
    int32_t *d_cols = r + (size_t)count * n;
    const size_t head = sizeof(int32_t) * ((size_t)2 * (size_t)count * (size_t)n + (size_t)CC_MAX_N);
    cd *partials = reinterpret_cast<cd *>(static_cast<char *>(d_work) + ((head + 255) & ~size_t(255)));
    PCVL_HIP_CHECK(hipMemcpyAsync(d_cols, cols.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream));
    const cd *U = reinterpret_cast<const cd *>(d_U);
    int dev = 0, cus = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(cc_init_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, d_cols, n, count,
                       seed, first_sample, pc);
    PCVL_HIP_CHECK(hipGetLastError());
    for (int k = 1; k <= n; k++) {
        int rc = launch_step<1>(k, U, m, n, count, seed, first_sample, pc, r, partials, cus, stream);
        if (rc != OK) return rc;
    }
    const int64_t tot = count * m;
    hipLaunchKernelGGL(cc_finish_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, r, m, n, count,
                       d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    // the host column list copied above must outlive the copy
    PCVL_HIP_CHECK(hipStreamSynchronize(stream));
    return OK;
}

extern "C" int pcvl_cc_samples_multi(const double *d_U, int m, const int32_t *h_inputs, const int64_t *h_counts,
                                     int ninputs, uint64_t seed, uint64_t first_sample, int32_t *d_out,
                                     hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= CC_MAX_M, "mode count outside [1, 256]");
    PCVL_ARG_CHECK(ninputs >= 0 && (ninputs == 0 || (h_inputs && h_counts)), "null input table");
    // photon columns of each input, its photon number and its first output row
    std::vector<std::vector<int32_t>> cols((size_t)ninputs);
    std::vector<int64_t> row0((size_t)ninputs + 1, 0);
    for (int i = 0; i < ninputs; i++) {
        PCVL_ARG_CHECK(h_counts[i] >= 0, "negative sample count");
        for (int q = 0; q < m; q++) {
            const int32_t v = h_inputs[(size_t)i * m + q];
            PCVL_ARG_CHECK(v >= 0, "negative occupation");
            for (int t = 0; t < v; t++) cols[(size_t)i].push_back(q);
        }
        PCVL_ARG_CHECK(cols[(size_t)i].size() <= (size_t)CC_MAX_N, "more than 30 photons");
        row0[(size_t)i + 1] = row0[(size_t)i] + h_counts[i];
    }
    const int64_t total = row0[(size_t)ninputs];
    if (total == 0) return OK;
    PCVL_ARG_CHECK(d_U && d_out, "null device buffer");
    // group the samples by photon number; each group runs one launch sequence
    std::vector<int> ns;
    for (int i = 0; i < ninputs; i++)
        if (h_counts[i] > 0 && std::find(ns.begin(), ns.end(), (int)cols[(size_t)i].size()) == ns.end())
            ns.push_back((int)cols[(size_t)i].size());
    int64_t gmax = 0;
    for (int n : ns) {
        int64_t c = 0;
        for (int i = 0; i < ninputs; i++) if ((int)cols[(size_t)i].size() == n) c += h_counts[i];
        gmax = std::max(gmax, c);
    }
    int nmax = 1;
    for (int n : ns) nmax = std::max(nmax, n);
    // device block: pc | r (gmax * nmax each) | inp | gidx (gmax) | cols (ninputs * nmax) | partial sums
    const size_t b_pr = sizeof(int32_t) * (size_t)gmax * nmax;
    const size_t b_inp = sizeof(int32_t) * (size_t)gmax;
    const size_t b_gidx = sizeof(int64_t) * (size_t)gmax;
    const size_t b_cols = sizeof(int32_t) * (size_t)ninputs * nmax;
    size_t b_part = 0;
    for (int n : ns) b_part = std::max(b_part, cc_partial_bytes(n, gmax));
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    char *blk = nullptr;
    PCVL_HIP_CHECK(ws_alloc(reinterpret_cast<void **>(&blk),
                                  2 * al(b_pr) + al(b_inp) + al(b_gidx) + al(b_cols) + al(b_part), stream));
    int32_t *pc = reinterpret_cast<int32_t *>(blk);
    int32_t *r = reinterpret_cast<int32_t *>(blk + al(b_pr));
    int32_t *d_inp = reinterpret_cast<int32_t *>(blk + 2 * al(b_pr));
    int64_t *d_gidx = reinterpret_cast<int64_t *>(blk + 2 * al(b_pr) + al(b_inp));
    int32_t *d_cols = reinterpret_cast<int32_t *>(blk + 2 * al(b_pr) + al(b_inp) + al(b_gidx));
    cd *partials = reinterpret_cast<cd *>(blk + 2 * al(b_pr) + al(b_inp) + al(b_gidx) + al(b_cols));
    const cd *U = reinterpret_cast<const cd *>(d_U);
    int dev = 0, cus = 0;
    int rc = OK;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) { set_error(std::string("device query: ") + hipGetErrorString(e)); rc = ERR_HIP; }
    for (int n : ns) {
        if (rc != OK) break;
        std::vector<int32_t> inp, ctab((size_t)ninputs * nmax, 0);
        std::vector<int64_t> gidx;
        for (int i = 0; i < ninputs; i++) {
            if ((int)cols[(size_t)i].size() != n) continue;
            for (int l = 0; l < n; l++) ctab[(size_t)i * n + l] = cols[(size_t)i][(size_t)l];
            for (int64_t s = 0; s < h_counts[i]; s++) { inp.push_back(i); gidx.push_back(row0[(size_t)i] + s); }
        }
        const int64_t count = (int64_t)gidx.size();
        if (count == 0) continue;
        e = hipMemcpyAsync(d_inp, inp.data(), sizeof(int32_t) * inp.size(), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_gidx, gidx.data(), sizeof(int64_t) * gidx.size(), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_cols, ctab.data(), sizeof(int32_t) * ctab.size(), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) { set_error(std::string("hipMemcpyAsync(cc multi): ") + hipGetErrorString(e)); rc = ERR_HIP; break; }
        if (n > 0) {
            hipLaunchKernelGGL(cc_init_multi_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, d_cols,
                               d_inp, d_gidx, n, count, seed, first_sample, pc);
            e = hipGetLastError();
            if (e != hipSuccess) { set_error(std::string("cc_init_multi: ") + hipGetErrorString(e)); rc = ERR_HIP; break; }
            for (int k = 1; k <= n && rc == OK; k++)
                rc = launch_step<1>(k, U, m, n, count, seed, first_sample, pc, r, partials, cus, stream, d_gidx);
            if (rc != OK) break;
        }
        const int64_t tot = count * m;
        hipLaunchKernelGGL(cc_finish_multi_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, r, d_gidx,
                           m, n, count, d_out);
        e = hipGetLastError();
        if (e != hipSuccess) { set_error(std::string("cc_finish_multi: ") + hipGetErrorString(e)); rc = ERR_HIP; break; }
        // the host tables of this group must outlive their copies
        e = hipStreamSynchronize(stream);
        if (e != hipSuccess) { set_error(std::string("cc multi: ") + hipGetErrorString(e)); rc = ERR_HIP; break; }
    }
    hipError_t se = hipStreamSynchronize(stream);
    hipError_t fe = hipFreeAsync(blk, stream);
    if (rc == OK && se != hipSuccess) { set_error(std::string("cc multi: ") + hipGetErrorString(se)); rc = ERR_HIP; }
    if (rc == OK && fe != hipSuccess) { set_error(std::string("hipFreeAsync(cc multi): ") + hipGetErrorString(fe)); rc = ERR_HIP; }
    return rc;
}
