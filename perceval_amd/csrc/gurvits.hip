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
// and a batch can be split across calls.  A matrix's samples are split into
// parts so that a small batch still fills the device; one wavefront per
// (matrix, part) stages A in LDS (zero-padded to NP x NP), lane l takes
// samples l, l + 64, ... of the part, forms y_i = sum_j x_j a_ij row by row
// (2 FMAs per broadcast LDS read) and multiplies the y_i into four chains by
// i mod 4; a finalize pass sums each matrix's parts in order.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <vector>

namespace pcvl {

constexpr int GURVITS_MAX_N = 48;
constexpr int GURVITS_WAVES = 4;   // waves per workgroup of the matrix-core kernel

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

// One wavefront per (matrix, part): part q of a matrix takes samples
// [q * per, min(iters, (q + 1) * per)), lane l of them l, l + 64, ...; it
// writes (sum Re X, sum Im X, sum |X|^2) to its slot, and the finalize kernel
// sums a matrix's parts in order (deterministic for a given split).
template <int NP>
__global__ void __launch_bounds__(64) gurvits_kernel(const cd *__restrict__ mats, const GurvitsItem *__restrict__ items,
                                                     int64_t count, int parts, int64_t iters, int64_t per,
                                                     uint64_t seed, uint64_t stream0, double *__restrict__ partial)
{
    __shared__ cd sA[NP * NP];
    const int lane = threadIdx.x;
    const int64_t nwork = count * parts;
    for (int64_t wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const int64_t b = wi / parts;
        const int64_t q = wi % parts;
        const int n = __builtin_amdgcn_readfirstlane(items[b].n);
        const int64_t off = items[b].off;
        for (int e = lane; e < NP * NP; e += 64) {
            const int i = e / NP, j = e % NP;
            sA[e] = (i < n && j < n) ? mats[off + (int64_t)i * n + j] : cd{0.0, 0.0};
        }
        __syncthreads();
        const uint64_t mask = n >= 64 ? ~0ull : ((1ull << n) - 1);
        const int64_t s_lo = q * per;
        const int64_t s_hi = (s_lo + per < iters) ? s_lo + per : iters;
        cd sum = {0.0, 0.0};
        double sq = 0.0;
        for (int64_t s = s_lo + lane; s < s_hi; s += 64) {
            const uint64_t bits = rand64(seed, stream0 + (uint64_t)b, (uint64_t)s);
            double x[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) x[j] = ((bits >> j) & 1) ? -1.0 : 1.0;
            cd p0 = {1.0, 0.0}, p1 = {1.0, 0.0}, p2 = {1.0, 0.0}, p3 = {1.0, 0.0};
            for (int i = 0; i < n; i++) {
                const cd *row = sA + i * NP;
                cd y = {0.0, 0.0};
#pragma unroll
                for (int j = 0; j < NP; j++) {   // zero padding beyond n adds nothing
                    const cd a = row[j];
                    y.r = fma(x[j], a.r, y.r);
                    y.i = fma(x[j], a.i, y.i);
                }
                const int c = i & 3;
                if (c == 0) p0 = cmul(p0, y);
                else if (c == 1) p1 = cmul(p1, y);
                else if (c == 2) p2 = cmul(p2, y);
                else p3 = cmul(p3, y);
            }
            cd X = cmul(cmul(p0, p1), cmul(p2, p3));
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
            partial[3 * wi] = sum.r;
            partial[3 * wi + 1] = sum.i;
            partial[3 * wi + 2] = sq;
        }
        __syncthreads();
    }
}

// The same estimate with y = A x formed on the FP64 matrix cores, 16 samples
// at a time: Y (NP x 16) = A (NP x NP) X (NP x 16, entries +-1), real and
// imaginary parts as two products, by v_mfma_f64_16x16x4f64 over 16-row
// blocks rb and 4-column blocks kb.  Operand layouts of that instruction
// (checked on the device): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]
// of the 16 x 4 and 4 x 16 operands, and D[l / 16 + 4 v][l % 16] in register v
// of the 16 x 16 result.  So lane l works on sample s0 + l % 16, takes
// x_j = bit j of that sample's signs for j = 4 kb + l / 16, and ends holding
// rows l / 16 + 4 v of every row block of y; the product over i < n is taken
// over those rows, then across the four lanes of the sample (lane ^ 16,
// lane ^ 32).  Lanes 0-15 accumulate.  A's fragments stay in registers up to
// NP = 32 and are read from LDS above.
typedef double gv_d4 __attribute__((ext_vector_type(4)));

template <int NP>
__global__ void __launch_bounds__(64 * GURVITS_WAVES, NP <= 32 ? 3 : 1) gurvits_mfma_kernel(const cd *__restrict__ mats,
                                                          const GurvitsItem *__restrict__ items, int64_t count,
                                                          int parts, int64_t iters, int64_t per, uint64_t seed,
                                                          uint64_t stream0, double *__restrict__ partial)
{
    constexpr int RB = NP / 16, KB = NP / 4;
    // A is read from LDS at each product: the registers it would take cost a
    // wave per SIMD, and the GURVITS_WAVES waves of a workgroup share one copy
    constexpr bool AREG = false;
    __shared__ cd sA[NP * NP];
    __shared__ double sred[GURVITS_WAVES][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, quad = lane >> 4;
    const int64_t nwork = count * parts;
    for (int64_t wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const int64_t b = wi / parts;
        const int64_t q = wi % parts;
        const int n = __builtin_amdgcn_readfirstlane(items[b].n);
        const int64_t off = items[b].off;
        for (int e = threadIdx.x; e < NP * NP; e += 64 * GURVITS_WAVES) {
            const int i = e / NP, j = e % NP;
            sA[e] = (i < n && j < n) ? mats[off + (int64_t)i * n + j] : cd{0.0, 0.0};
        }
        __syncthreads();
        double ar[AREG ? RB * KB : 1], ai[AREG ? RB * KB : 1];
        if constexpr (AREG) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
#pragma unroll
                for (int kb = 0; kb < KB; kb++) {
                    const cd v = sA[(16 * rb + col) * NP + 4 * kb + quad];
                    ar[rb * KB + kb] = v.r;
                    ai[rb * KB + kb] = v.i;
                }
        }
        const uint64_t mask = n >= 64 ? ~0ull : ((1ull << n) - 1);
        const int64_t s_lo = q * per;
        const int64_t s_hi = (s_lo + per < iters) ? s_lo + per : iters;
        cd sum = {0.0, 0.0};
        double sq = 0.0;
        for (int64_t s0 = s_lo + 16 * wave; s0 < s_hi; s0 += 16 * GURVITS_WAVES) {
            const int64_t s = s0 + col;
            const bool valid = s < s_hi;
            const uint64_t bits = valid ? rand64(seed, stream0 + (uint64_t)b, (uint64_t)s) : 0ull;
            gv_d4 yr[RB], yi[RB];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                yr[rb] = gv_d4{0.0, 0.0, 0.0, 0.0};
                yi[rb] = gv_d4{0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                const double x = ((bits >> (4 * kb + quad)) & 1) ? -1.0 : 1.0;   // columns j >= n are zero in A
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    double a_r, a_i;
                    if constexpr (AREG) {
                        a_r = ar[rb * KB + kb];
                        a_i = ai[rb * KB + kb];
                    } else {
                        const cd v = sA[(16 * rb + col) * NP + 4 * kb + quad];
                        a_r = v.r;
                        a_i = v.i;
                    }
                    yr[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, x, yr[rb], 0, 0, 0);
                    yi[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, x, yi[rb], 0, 0, 0);
                }
            }
            // this lane's rows quad + 4 v of each block, then the sample's four lanes
            cd p = {1.0, 0.0};
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (16 * rb + quad + 4 * v < n) p = cmul(p, cd{yr[rb][v], yi[rb][v]});
            cd o = {__shfl_xor(p.r, 16, 64), __shfl_xor(p.i, 16, 64)};
            p = cmul(p, o);
            o = {__shfl_xor(p.r, 32, 64), __shfl_xor(p.i, 32, 64)};
            p = cmul(p, o);
            if (__builtin_popcountll(bits & mask) & 1) p = {-p.r, -p.i};
            if (valid && quad == 0) {
                sum.r += p.r;
                sum.i += p.i;
                sq = fma(p.r, p.r, fma(p.i, p.i, sq));
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sum.r += __shfl_xor(sum.r, o, 64);
            sum.i += __shfl_xor(sum.i, o, 64);
            sq += __shfl_xor(sq, o, 64);
        }
        if (lane == 0) {
            sred[wave][0] = sum.r;
            sred[wave][1] = sum.i;
            sred[wave][2] = sq;
        }
        __syncthreads();
        if (threadIdx.x == 0) {   // the waves' sums in wave order
            double r = 0.0, im = 0.0, q2 = 0.0;
            for (int w = 0; w < GURVITS_WAVES; w++) {
                r += sred[w][0];
                im += sred[w][1];
                q2 += sred[w][2];
            }
            partial[3 * wi] = r;
            partial[3 * wi + 1] = im;
            partial[3 * wi + 2] = q2;
        }
        __syncthreads();
    }
}

__global__ void gurvits_finalize_kernel(const double *__restrict__ partial, int64_t count, int parts, int64_t iters,
                                        cd *__restrict__ est, double *__restrict__ err)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    double sr = 0.0, si = 0.0, sq = 0.0;
    for (int q = 0; q < parts; q++) {
        const double *pp = partial + 3 * (b * parts + q);
        sr += pp[0];
        si += pp[1];
        sq += pp[2];
    }
    const double inv = 1.0 / (double)iters;
    const cd mean = {sr * inv, si * inv};
    const double var = sq * inv - (mean.r * mean.r + mean.i * mean.i);
    est[b] = mean;
    err[b] = 2.5758293035489004 * sqrt((var > 0.0 ? var : 0.0) * inv);
}

template <int NP>
static int launch_gurvits(const cd *mats, const GurvitsItem *items, int64_t count, int parts, int64_t iters,
                          int64_t per, uint64_t seed, uint64_t stream0, double *partial, hipStream_t st)
{
    int dev = 0, cus = 0, per_cu = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // orders above 8: y = A x on the matrix cores (NP a multiple of 16)
    constexpr bool MF = NP >= 16 && NP % 16 == 0;
    const void *kp = MF ? reinterpret_cast<const void *>(&gurvits_mfma_kernel<MF ? NP : 16>)
                        : reinterpret_cast<const void *>(&gurvits_kernel<NP>);
    const int threads = MF ? 64 * GURVITS_WAVES : 64;
    PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, threads, 0));
    if (per_cu < 1) per_cu = 1;
    const int grid = (int)std::min<int64_t>(count * parts, (int64_t)cus * per_cu);
    if constexpr (MF)
        hipLaunchKernelGGL((gurvits_mfma_kernel<NP>), dim3(grid), dim3(threads), 0, st, mats, items, count, parts,
                           iters, per, seed, stream0, partial);
    else
        hipLaunchKernelGGL((gurvits_kernel<NP>), dim3(grid), dim3(64), 0, st, mats, items, count, parts, iters, per,
                           seed, stream0, partial);
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
    // samples of a matrix are split into parts of at least 64 * 16 samples so
    // that the batch fills the device (4 waves per SIMD of 256 CUs)
    int parts = (int)std::max<int64_t>(1, std::min<int64_t>((4096 + count - 1) / count, iters / 1024));
    const int64_t per = (iters + parts - 1) / parts;
    parts = (int)((iters + per - 1) / per);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t b_items = al(sizeof(GurvitsItem) * items.size());
    const size_t b_part = al(sizeof(double) * 3 * (size_t)count * parts);
    char *blk = nullptr;
    PCVL_HIP_CHECK(ws_alloc(reinterpret_cast<void **>(&blk), b_items + b_part, stream));
    GurvitsItem *d_items = reinterpret_cast<GurvitsItem *>(blk);
    double *d_part = reinterpret_cast<double *>(blk + b_items);
    hipError_t e = hipMemcpyAsync(d_items, items.data(), sizeof(GurvitsItem) * items.size(), hipMemcpyHostToDevice,
                                  stream);
    int rc = OK;
    if (e != hipSuccess) {
        set_error(std::string("hipMemcpyAsync(gurvits items): ") + hipGetErrorString(e));
        rc = ERR_HIP;
    } else {
        const cd *mats = reinterpret_cast<const cd *>(d_mats);
        if (nmax <= 8) rc = launch_gurvits<8>(mats, d_items, count, parts, iters, per, seed, stream0, d_part, stream);
        else if (nmax <= 16) rc = launch_gurvits<16>(mats, d_items, count, parts, iters, per, seed, stream0, d_part, stream);
        else if (nmax <= 32) rc = launch_gurvits<32>(mats, d_items, count, parts, iters, per, seed, stream0, d_part, stream);
        else rc = launch_gurvits<48>(mats, d_items, count, parts, iters, per, seed, stream0, d_part, stream);
        if (rc == OK) {
            hipLaunchKernelGGL(gurvits_finalize_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                               d_part, count, parts, iters, reinterpret_cast<cd *>(d_est), d_err);
            e = hipGetLastError();
            if (e != hipSuccess) { set_error(std::string("gurvits finalize: ") + hipGetErrorString(e)); rc = ERR_HIP; }
        }
    }
    // the table must outlive the kernel: wait for it, then release it
    hipError_t se = hipStreamSynchronize(stream);
    hipError_t fe = hipFreeAsync(blk, stream);
    if (rc == OK && se != hipSuccess) { set_error(std::string("gurvits: ") + hipGetErrorString(se)); rc = ERR_HIP; }
    if (rc == OK && fe != hipSuccess) { set_error(std::string("hipFree(gurvits items): ") + hipGetErrorString(fe)); rc = ERR_HIP; }
    return rc;
}
