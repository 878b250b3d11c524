// Gurvits' randomised permanent estimate, batched.
//
// Replaces xq.estimate_permanent_cx(M, iterations) as called from
// NaiveApproxBackend._compute_permanent and prob_amplitude_with_error
// (perceval/backends/_naive_approx.py:52-62; the estimator is Gurvits &
// Samorodnitsky 2002, docs/source/references.bib:1135).  For x uniform in
// {+1,-1}^n,
//     X(x) = (prod_j x_j) prod_i (sum_j a_ij x_j),   E[X] = perm(A),
// since every term of the expanded product that is not a permutation carries
// an odd power of some x_j.  The estimate is the mean of `iters` samples; its
// 99% error bound is the normal one, 2.5758 * sqrt((E|X|^2 - |mean|^2) / iters).
//
// Sample s of matrix b takes its signs from the 64 bits rand64(seed,
// stream0 + b, s) (bit j -> x_j = -1), so the oracle draws the same samples
// and a batch can be split across calls.  One wavefront per matrix: A^T is
// staged in LDS, zero-padded to NP x NP (rows i >= n keep y_i = 1), and lane l
// takes samples l, l + 64, ...; per sample the column sums y = A x are built
// column by column (2 FMAs per entry, column reads broadcast from LDS), then
// multiplied as four independent chains.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <vector>

namespace pcvl {

constexpr int GURVITS_MAX_N = 48;

struct GurvitsItem {
    int64_t off;   // complex offset of the matrix
    int32_t n;
    int32_t pad;
};

template <int NP>
__device__ __forceinline__ cd gurvits_prod(const cd *y)
{
    cd p0 = y[0], p1 = y[1], p2 = y[2], p3 = y[3];
#pragma unroll
    for (int i = 4; i < NP; i++) {
        if ((i & 3) == 0) p0 = cmul(p0, y[i]);
        else if ((i & 3) == 1) p1 = cmul(p1, y[i]);
        else if ((i & 3) == 2) p2 = cmul(p2, y[i]);
        else p3 = cmul(p3, y[i]);
    }
    return cmul(cmul(p0, p1), cmul(p2, p3));
}

template <int NP>
__global__ void __launch_bounds__(64) gurvits_kernel(const cd *__restrict__ mats, const GurvitsItem *__restrict__ items,
                                                     int64_t count, int64_t iters, uint64_t seed, uint64_t stream0,
                                                     cd *__restrict__ est, double *__restrict__ err)
{
    __shared__ cd sAt[NP * NP];
    const int lane = threadIdx.x;
    for (int64_t b = blockIdx.x; b < count; b += gridDim.x) {
        const int n = __builtin_amdgcn_readfirstlane(items[b].n);
        const int64_t off = items[b].off;
        for (int e = lane; e < NP * NP; e += 64) {
            const int j = e / NP, i = e % NP;
            sAt[e] = (i < n && j < n) ? mats[off + (int64_t)i * n + j] : cd{0.0, 0.0};
        }
        __syncthreads();
        const uint64_t mask = n >= 64 ? ~0ull : ((1ull << n) - 1);
        cd sum = {0.0, 0.0};
        double sq = 0.0;
        for (int64_t s = lane; s < iters; s += 64) {
            const uint64_t bits = rand64(seed, stream0 + (uint64_t)b, (uint64_t)s);
            cd y[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) y[i] = {i < n ? 0.0 : 1.0, 0.0};
            for (int j = 0; j < n; j++) {
                const double x = ((bits >> j) & 1) ? -1.0 : 1.0;
                const cd *col = sAt + j * NP;
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    y[i].r = fma(x, col[i].r, y[i].r);
                    y[i].i = fma(x, col[i].i, y[i].i);
                }
            }
            cd X = gurvits_prod<NP>(y);
            if (__builtin_popcountll(bits & mask) & 1) X = {-X.r, -X.i};
            sum.r += X.r;
            sum.i += X.i;
            sq = fma(X.r, X.r, fma(X.i, X.i, sq));
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sum.r += __shfl_xor(sum.r, o, 64);
            sum.i += __shfl_xor(sum.i, o, 64);
            sq += __shfl_xor(sq, o, 64);
        }
        if (lane == 0) {
            const double inv = 1.0 / (double)iters;
            const cd mean = {sum.r * inv, sum.i * inv};
            const double var = sq * inv - (mean.r * mean.r + mean.i * mean.i);
            est[b] = mean;
            err[b] = 2.5758293035489004 * sqrt((var > 0.0 ? var : 0.0) * inv);
        }
        __syncthreads();
    }
}

template <int NP>
static int launch_gurvits(const cd *mats, const GurvitsItem *items, int64_t count, int64_t iters, uint64_t seed,
                          uint64_t stream0, cd *est, double *err, hipStream_t st)
{
    int dev = 0, cus = 0, per_cu = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&gurvits_kernel<NP>),
                                                                64, 0));
    if (per_cu < 1) per_cu = 1;
    const int grid = (int)std::min<int64_t>(count, (int64_t)cus * per_cu);
    hipLaunchKernelGGL((gurvits_kernel<NP>), dim3(grid), dim3(64), 0, st, mats, items, count, iters, seed, stream0, est,
                       err);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_gurvits_batch(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                                  int64_t iters, uint64_t seed, uint64_t stream0, double *d_est, double *d_err,
                                  hipStream_t stream)
{
    PCVL_ARG_CHECK(count >= 0 && iters >= 1, "count < 0 or iterations < 1");
    if (count == 0) return OK;
    PCVL_ARG_CHECK(d_mats && h_n && h_offsets && d_est && d_err, "null buffer");
    std::vector<GurvitsItem> items((size_t)count);
    int nmax = 0;
    for (int64_t b = 0; b < count; b++) {
        PCVL_ARG_CHECK(h_n[b] >= 0 && h_n[b] <= GURVITS_MAX_N, "matrix order outside [0, 48]");
        PCVL_ARG_CHECK(h_offsets[b] >= 0, "negative matrix offset");
        items[(size_t)b] = GurvitsItem{h_offsets[b], h_n[b], 0};
        nmax = std::max(nmax, (int)h_n[b]);
    }
    GurvitsItem *d_items = nullptr;
    PCVL_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&d_items), sizeof(GurvitsItem) * items.size(), stream));
    hipError_t e = hipMemcpyAsync(d_items, items.data(), sizeof(GurvitsItem) * items.size(), hipMemcpyHostToDevice,
                                  stream);
    int rc = OK;
    if (e != hipSuccess) {
        set_error(std::string("hipMemcpyAsync(gurvits items): ") + hipGetErrorString(e));
        rc = ERR_HIP;
    } else {
        const cd *mats = reinterpret_cast<const cd *>(d_mats);
        cd *est = reinterpret_cast<cd *>(d_est);
        if (nmax <= 8) rc = launch_gurvits<8>(mats, d_items, count, iters, seed, stream0, est, d_err, stream);
        else if (nmax <= 16) rc = launch_gurvits<16>(mats, d_items, count, iters, seed, stream0, est, d_err, stream);
        else if (nmax <= 24) rc = launch_gurvits<24>(mats, d_items, count, iters, seed, stream0, est, d_err, stream);
        else if (nmax <= 32) rc = launch_gurvits<32>(mats, d_items, count, iters, seed, stream0, est, d_err, stream);
        else if (nmax <= 40) rc = launch_gurvits<40>(mats, d_items, count, iters, seed, stream0, est, d_err, stream);
        else rc = launch_gurvits<48>(mats, d_items, count, iters, seed, stream0, est, d_err, stream);
    }
    // the table must outlive the kernel: wait for it, then release it
    hipError_t se = hipStreamSynchronize(stream);
    hipError_t fe = hipFreeAsync(d_items, stream);
    if (rc == OK && se != hipSuccess) { set_error(std::string("gurvits: ") + hipGetErrorString(se)); rc = ERR_HIP; }
    if (rc == OK && fe != hipSuccess) { set_error(std::string("hipFree(gurvits items): ") + hipGetErrorString(fe)); rc = ERR_HIP; }
    return rc;
}
