// Autoregressive proposal network for the Metropolis independence sampler:
// forward, backward and optimiser kernels (bf16 weights, fp32 arithmetic).
//
// No counterpart in the reference.  The model proposes an output state of n
// photons over m modes as a non-decreasing sequence of modes s_0 <= ... <=
// s_{n-1}; step p sees the counts of the modes drawn so far and a one-hot of
// p (D = m + n inputs), and emits logits over the m modes, masked below
// s_{p-1}.  The sequence is unique per Fock state, so q(state) is exact.
//   h = relu(x W1^T + b1)  (hidden H),  logits = h W2^T + b2.
// Matrices are row-major; W1 is H x D and W2 is m x H, stored as bf16 for the
// products with fp32 master copies for the optimiser.
#include "common.h"
#include "../../include/pcvl_hip.h"

namespace pcvl {

__device__ __forceinline__ float bf16_to_f32(uint16_t b)
{
    return __uint_as_float(((uint32_t)b) << 16);
}

__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}

template <typename T> __device__ __forceinline__ float ld(const T *p, size_t i);
template <> __device__ __forceinline__ float ld<float>(const float *p, size_t i) { return p[i]; }
template <> __device__ __forceinline__ float ld<uint16_t>(const uint16_t *p, size_t i) { return bf16_to_f32(p[i]); }

// Y[i][j] = act(alpha * sum_k A(i,k) B(k,j) + bias[j]) (+ Y[i][j] if accumulate)
// with A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj].  64 x 64 output
// tile per 256-thread block, 16-deep K slices staged through LDS, 4 x 4
// outputs per thread.
constexpr int GT = 64, GK = 16;

template <typename TB>
__global__ void __launch_bounds__(256) gemm_kernel(const float *__restrict__ A, int64_t sai, int64_t sak,
                                                   const TB *__restrict__ B, int64_t sbk, int64_t sbj,
                                                   float *__restrict__ Y, int64_t ldy, int M, int N, int K,
                                                   const float *__restrict__ bias, int relu, int accumulate)
{
    __shared__ float As[GK][GT + 1];
    __shared__ float Bs[GK][GT + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += GK) {
        for (int e = threadIdx.x; e < GK * GT; e += 256) {
            const int kk = e / GT, ii = e % GT;
            const int i = i0 + ii, k = k0 + kk;
            As[kk][ii] = (i < M && k < K) ? A[i * sai + k * sak] : 0.0f;
            const int j = j0 + ii;
            Bs[kk][ii] = (j < N && k < K) ? ld<TB>(B, (size_t)(k * sbk + j * sbj)) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; r++) a[r] = As[kk][ty + 16 * r];
#pragma unroll
            for (int c = 0; c < 4; c++) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[r][c] = fmaf(a[r], b[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + ty + 16 * r;
        if (i >= M) continue;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int j = j0 + tx + 16 * c;
            if (j >= N) continue;
            float v = acc[r][c];
            if (bias) v += bias[j];
            if (accumulate) v += Y[(size_t)i * ldy + j];
            if (relu) v = v > 0.0f ? v : 0.0f;
            Y[(size_t)i * ldy + j] = v;
        }
    }
}

// Weight gradient over a long row dimension, split deterministically:
//   part[s][i][j] = sum_{r in split s} dY[r][i] * Xe[r][j],  j < No + 1,
// with Xe = [X | 1], so column No of the result is db = colsum(dY).  Split s
// covers rows [s*rows_per, min(R, (s+1)*rows_per)); the second kernel sums
// the splits in order, so the result does not depend on the launch.
constexpr int WG_ROWS = 16;

__global__ void __launch_bounds__(256) wgrad_partial_kernel(const float *__restrict__ dY, int64_t ldd,
                                                            const float *__restrict__ X, int64_t ldx, int R,
                                                            int Mo, int No, int rows_per, float *__restrict__ part)
{
    __shared__ float As[WG_ROWS][GT + 1];
    __shared__ float Bs[WG_ROWS][GT + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
    const int split = blockIdx.z;
    const int r0 = split * rows_per;
    const int r1 = min(R, r0 + rows_per);
    const int Ne = No + 1;
    float acc[4][4] = {};
    for (int rb = r0; rb < r1; rb += WG_ROWS) {
        for (int e = threadIdx.x; e < WG_ROWS * GT; e += 256) {
            const int kk = e / GT, ii = e % GT;
            const int r = rb + kk;
            const bool rv = r < r1;
            const int i = i0 + ii, j = j0 + ii;
            As[kk][ii] = (rv && i < Mo) ? dY[(size_t)r * ldd + i] : 0.0f;
            Bs[kk][ii] = (rv && j < No) ? X[(size_t)r * ldx + j] : ((rv && j == No) ? 1.0f : 0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < WG_ROWS; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; q++) a[q] = As[kk][ty + 16 * q];
#pragma unroll
            for (int c = 0; c < 4; c++) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[q][c] = fmaf(a[q], b[c], acc[q][c]);
        }
        __syncthreads();
    }
    float *P = part + (size_t)split * Mo * Ne;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = i0 + ty + 16 * q;
        if (i >= Mo) continue;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int j = j0 + tx + 16 * c;
            if (j < Ne) P[(size_t)i * Ne + j] = acc[q][c];
        }
    }
}

__global__ void wgrad_reduce_kernel(const float *__restrict__ part, int splits, int Mo, int No,
                                    float *__restrict__ dW, float *__restrict__ db)
{
    const int Ne = No + 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)Mo * Ne) return;
    float s = 0.0f;
    for (int k = 0; k < splits; k++) s += part[(size_t)k * Mo * Ne + e];
    const int i = (int)(e / Ne), j = (int)(e % Ne);
    if (j < No) {
        if (dW) dW[(size_t)i * No + j] = s;
    } else if (db) {
        db[i] = s;
    }
}

// dH *= (H > 0)
__global__ void relu_bwd_kernel(float *__restrict__ dH, const float *__restrict__ H, int64_t count)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count && !(H[e] > 0.0f)) dH[e] = 0.0f;
}

// db[j] = sum_i dY[i][j]: 64 columns per block, four row groups combined in
// a fixed order (coalesced rows; for long row counts use the weight-gradient
// kernels, which split the rows over many blocks).
__global__ void __launch_bounds__(256) colsum_kernel(const float *__restrict__ dY, int M, int N, float *__restrict__ db)
{
    __shared__ float part[4][64];
    const int c = threadIdx.x % 64, g = threadIdx.x / 64;
    const int j = blockIdx.x * 64 + c;
    float s = 0.0f;
    if (j < N)
        for (int i = g; i < M; i += 4) s += dY[(size_t)i * N + j];
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && j < N) db[j] = (part[0][c] + part[1][c]) + (part[2][c] + part[3][c]);
}

// Masked softmax cross-entropy of row i with target t[i], modes < lo[i]
// masked: loss[i] = -log q(t | row); dlogits = (softmax - onehot) * scale.
__global__ void ce_kernel(const float *__restrict__ logits, int M, int m, const int32_t *__restrict__ target,
                          const int32_t *__restrict__ lo, float scale, float *__restrict__ loss,
                          float *__restrict__ dlogits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float *z = logits + (size_t)i * m;
    const int l0 = lo[i];
    const int t = target[i];
    if (t < l0 || t >= m || l0 < 0) {   // target outside the allowed modes: q = 0
        loss[i] = __int_as_float(0x7f800000);
        if (dlogits)
            for (int j = 0; j < m; j++) dlogits[(size_t)i * m + j] = 0.0f;
        return;
    }
    float mx = -3.0e38f;
    for (int j = l0; j < m; j++) mx = fmaxf(mx, z[j]);
    float se = 0.0f;
    for (int j = l0; j < m; j++) se += __expf(z[j] - mx);
    const float lse = mx + __logf(se);
    loss[i] = lse - z[t];
    if (dlogits) {
        float *d = dlogits + (size_t)i * m;
        for (int j = 0; j < m; j++) {
            const float p = (j >= l0) ? __expf(z[j] - lse) : 0.0f;
            d[j] = (p - (j == t ? 1.0f : 0.0f)) * scale;
        }
    }
}

// Adam on fp32 master weights; writes the bf16 copy used by the products.
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, uint16_t *__restrict__ wb, int64_t count, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float gg = g[e];
    const float a = b1 * m1[e] + (1.0f - b1) * gg;
    const float b = b2 * m2[e] + (1.0f - b2) * gg * gg;
    m1[e] = a;
    m2[e] = b;
    const float v = w[e] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[e] = v;
    if (wb) wb[e] = f32_to_bf16(v);
}

__global__ void to_bf16_kernel(const float *__restrict__ w, uint16_t *__restrict__ wb, int64_t count)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) wb[e] = f32_to_bf16(w[e]);
}

// Features of step p for each sequence row b: x = [counts of seq[b][0..p-1] | onehot(p)],
// and the mask bound lo[b] = seq[b][p-1] (0 at p = 0).
__global__ void features_kernel(const int32_t *__restrict__ seq, int B, int n, int m, int p,
                                float *__restrict__ x, int32_t *__restrict__ lo)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int D = m + n;
    float *xb = x + (size_t)b * D;
    for (int j = 0; j < D; j++) xb[j] = 0.0f;
    for (int q = 0; q < p; q++) {
        const int j = seq[(size_t)b * n + q];
        if (j >= 0 && j < m) xb[j] += 1.0f;
    }
    xb[m + p] = 1.0f;
    const int l = p > 0 ? seq[(size_t)b * n + p - 1] : 0;
    lo[b] = (l >= 0 && l < m) ? l : 0;
}

// Teacher-forcing features for all (row, step) pairs of a batch of sorted
// sequences: row b*n + p.  Also writes the targets seq[b][p].
__global__ void features_all_kernel(const int32_t *__restrict__ seq, int B, int n, int m, float *__restrict__ x,
                                    int32_t *__restrict__ lo, int32_t *__restrict__ target)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (int64_t)B * n) return;
    const int b = (int)(r / n), p = (int)(r % n);
    const int D = m + n;
    float *xr = x + (size_t)r * D;
    for (int j = 0; j < D; j++) xr[j] = 0.0f;
    for (int q = 0; q < p; q++) {
        const int j = seq[(size_t)b * n + q];
        if (j >= 0 && j < m) xr[j] += 1.0f;
    }
    xr[m + p] = 1.0f;
    const int l = p > 0 ? seq[(size_t)b * n + p - 1] : 0;
    lo[r] = (l >= 0 && l < m) ? l : 0;
    target[r] = seq[(size_t)b * n + p];   // validated by ce_kernel
}

// Draw step p of every row from the masked softmax of its logits; the uniform
// is the counter-based stream (seed, first + b, p).  Accumulates log q.
__global__ void sample_step_kernel(const float *__restrict__ logits, int B, int m, int n, int p,
                                   const int32_t *__restrict__ lo, uint64_t seed, uint64_t first,
                                   int32_t *__restrict__ seq, double *__restrict__ logq)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float *z = logits + (size_t)b * m;
    const int l0 = lo[b];
    float mx = -3.0e38f;
    for (int j = l0; j < m; j++) mx = fmaxf(mx, z[j]);
    double se = 0.0;
    for (int j = l0; j < m; j++) se += exp((double)(z[j] - mx));
    const double target = uniform01(seed, first + (uint64_t)b, (uint64_t)p) * se;
    double cum = 0.0;
    int pick = m - 1;
    for (int j = l0; j < m; j++) {
        cum += exp((double)(z[j] - mx));
        if (target < cum) { pick = j; break; }
    }
    seq[(size_t)b * n + p] = pick;
    const double lp = (double)(z[pick] - mx) - log(se);
    logq[b] = (p == 0 ? 0.0 : logq[b]) + lp;
}

// All n autoregressive steps of a tile of proposals in one launch.  The first
// layer's input is [counts of the modes drawn so far | one-hot of the step],
// so its pre-activation is updated in place: + W1[:, drawn mode], and the
// one-hot column moves from m+p to m+p+1 (no GEMM, no HBM round trip).  The
// second layer runs from LDS (relu(h)^T in fp32, W2^T in bf16), each thread
// computing a 4-sample x 2-logit tile.  The draw is the masked softmax of the
// logits with the uniform (seed, first + b, p), done by the sample's 8
// threads: group max, group sum of exp (double), ordered partial sums for the
// inverse CDF; log q accumulates (z_pick - max) - log(sum) from the same
// logits, so q is exactly the law of the draw.
constexpr int FS_TB = 16;            // proposals per workgroup
constexpr int FS_THREADS = 128;      // 8 per proposal
constexpr int FS_MW = 64;            // logits padded to 64 modes
constexpr int FS_MAX_H = 256;
constexpr int FS_HPT = FS_MAX_H / 8; // h_pre entries per thread

__global__ void __launch_bounds__(FS_THREADS) sample_fused_kernel(
    const uint16_t *__restrict__ W1, const float *__restrict__ b1, const uint16_t *__restrict__ W2,
    const float *__restrict__ b2, int B, int m, int n, int H, uint64_t seed, uint64_t first,
    int32_t *__restrict__ seq, double *__restrict__ logq)
{
    extern __shared__ float s_fs[];
    float *hT = s_fs;                                           // H x TB
    float *zs = hT + (size_t)H * FS_TB;                         // TB x 64
    uint16_t *W2T = reinterpret_cast<uint16_t *>(zs + FS_TB * FS_MW);   // H x 64
    const int tid = threadIdx.x;
    const int D = m + n;
    for (int e = tid; e < H * FS_MW; e += FS_THREADS) {
        const int h = e / FS_MW, j = e % FS_MW;
        W2T[e] = j < m ? W2[(size_t)j * H + h] : (uint16_t)0;
    }
    // ownership for h_pre and the draw: proposal s, sub-thread sub
    const int s = tid >> 3, sub = tid & 7;
    const int64_t b = (int64_t)blockIdx.x * FS_TB + s;
    const bool valid = b < B;
    const int hpt = H / 8;
    float hp[FS_HPT];
#pragma unroll
    for (int t = 0; t < FS_HPT; t++) {
        const int h = sub + 8 * t;
        hp[t] = (t < hpt) ? b1[h] + bf16_to_f32(W1[(size_t)h * D + m]) : 0.0f;
    }
    // z tile: proposals 4*sq .. 4*sq+3, logits 2*jp, 2*jp+1
    const int sq = tid & 3, jp = tid >> 2;
    int lo = 0;
    double lq = 0.0;
    __syncthreads();
    for (int p = 0; p < n; p++) {
#pragma unroll
        for (int t = 0; t < FS_HPT; t++)
            if (t < hpt) hT[(sub + 8 * t) * FS_TB + s] = hp[t] > 0.0f ? hp[t] : 0.0f;
        __syncthreads();
        float acc[4][2] = {};
        for (int h = 0; h < H; h++) {
            const float4 hv = *reinterpret_cast<const float4 *>(hT + h * FS_TB + 4 * sq);
            const uint32_t w2 = *reinterpret_cast<const uint32_t *>(W2T + h * FS_MW + 2 * jp);
            const float w0 = __uint_as_float(w2 << 16), w1 = __uint_as_float(w2 & 0xffff0000u);
            acc[0][0] = fmaf(hv.x, w0, acc[0][0]); acc[0][1] = fmaf(hv.x, w1, acc[0][1]);
            acc[1][0] = fmaf(hv.y, w0, acc[1][0]); acc[1][1] = fmaf(hv.y, w1, acc[1][1]);
            acc[2][0] = fmaf(hv.z, w0, acc[2][0]); acc[2][1] = fmaf(hv.z, w1, acc[2][1]);
            acc[3][0] = fmaf(hv.w, w0, acc[3][0]); acc[3][1] = fmaf(hv.w, w1, acc[3][1]);
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int j = 2 * jp + c;
                zs[(4 * sq + r) * FS_MW + j] = j < m ? acc[r][c] + b2[j] : 0.0f;
            }
        __syncthreads();
        // draw of proposal s by its 8 threads; thread sub owns modes 8*sub .. 8*sub+7
        const float *z = zs + s * FS_MW;
        float mx = -3.0e38f;
        for (int u = 0; u < 8; u++) {
            const int j = 8 * sub + u;
            if (j >= lo && j < m) mx = fmaxf(mx, z[j]);
        }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        double ev[8];   // this thread's exp(z_j - mx), reused by the draw below
        double part = 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = 8 * sub + u;
            ev[u] = (j >= lo && j < m) ? exp((double)(z[j] - mx)) : 0.0;
            if (j >= lo && j < m) part += ev[u];
        }
        double se = part;
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) se += __shfl_xor(se, o, 64);
        // exclusive prefix of the partial sums in sub order
        double base = 0.0;
        for (int v = 0; v < 7; v++) {
            const double pv = __shfl(part, (tid & ~7) + v, 64);
            if (v < sub) base += pv;
        }
        const double target = uniform01(seed, first + (uint64_t)(valid ? b : 0), (uint64_t)p) * se;
        int cand = -1;
        if (target >= base || sub == 0) {
            double cum = base;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int j = 8 * sub + u;
                if (cand < 0 && j >= lo && j < m) {
                    cum += ev[u];
                    if (target < cum) cand = j;
                }
            }
            if (target < base) cand = -1;   // belongs to an earlier thread
        }
        // the first thread (in sub order) that holds the target; none -> last mode
        int pick = m - 1;
        for (int v = 7; v >= 0; v--) {
            const int cv = __shfl(cand, (tid & ~7) + v, 64);
            if (cv >= 0) pick = cv;
        }
        lq += (double)(z[pick] - mx) - log(se);
        if (valid && sub == 0) seq[b * n + p] = pick;
        lo = pick;
        // next pre-activation: + W1[:, pick], one-hot m+p -> m+p+1
#pragma unroll
        for (int t = 0; t < FS_HPT; t++) {
            if (t < hpt) {
                const size_t row = (size_t)(sub + 8 * t) * D;
                float v = hp[t] + bf16_to_f32(W1[row + pick]) - bf16_to_f32(W1[row + m + p]);
                if (p + 1 < n) v += bf16_to_f32(W1[row + m + p + 1]);
                hp[t] = v;
            }
        }
        __syncthreads();   // zs and hT are rewritten by the next step
    }
    if (valid && sub == 0) logq[b] = lq;
}

} // namespace pcvl

using namespace pcvl;

static int grid1(int64_t count) { return (int)((count + 255) / 256); }

extern "C" int pcvl_mlp_gemm(const float *d_A, int64_t sai, int64_t sak, const void *d_B, int b_is_bf16,
                             int64_t sbk, int64_t sbj, float *d_Y, int64_t ldy, int M, int N, int K,
                             const float *d_bias, int relu, int accumulate, hipStream_t stream)
{
    PCVL_ARG_CHECK(M >= 0 && N >= 0 && K >= 0, "negative GEMM size");
    if (M == 0 || N == 0) return OK;
    PCVL_ARG_CHECK(d_A && d_B && d_Y, "null device buffer");
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    if (b_is_bf16) {
        hipLaunchKernelGGL(gemm_kernel<uint16_t>, grid, dim3(256), 0, stream, d_A, sai, sak,
                           static_cast<const uint16_t *>(d_B), sbk, sbj, d_Y, ldy, M, N, K, d_bias, relu, accumulate);
    } else {
        hipLaunchKernelGGL(gemm_kernel<float>, grid, dim3(256), 0, stream, d_A, sai, sak,
                           static_cast<const float *>(d_B), sbk, sbj, d_Y, ldy, M, N, K, d_bias, relu, accumulate);
    }
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

static int wgrad_splits(int R)
{
    int s = R / 256;
    return s < 1 ? 1 : (s > 128 ? 128 : s);
}

extern "C" int pcvl_mlp_wgrad_workspace(int R, int Mo, int No, int64_t *floats)
{
    PCVL_ARG_CHECK(floats != nullptr, "null output pointer");
    PCVL_ARG_CHECK(R >= 0 && Mo >= 0 && No >= 0, "negative size");
    *floats = (int64_t)wgrad_splits(R) * Mo * (No + 1);
    return OK;
}

extern "C" int pcvl_mlp_wgrad(const float *d_dY, int64_t ldd, const float *d_X, int64_t ldx, int R, int Mo, int No,
                              float *d_dW, float *d_db, float *d_work, int64_t work_floats, hipStream_t stream)
{
    PCVL_ARG_CHECK(R >= 0 && Mo >= 0 && No >= 0, "negative size");
    if (Mo == 0) return OK;
    PCVL_ARG_CHECK(d_dY && d_work && (No == 0 || d_X), "null device buffer");
    PCVL_ARG_CHECK(ldd >= Mo && (No == 0 || ldx >= No), "leading dimension below the row length");
    const int splits = wgrad_splits(R);
    PCVL_ARG_CHECK(work_floats >= (int64_t)splits * Mo * (No + 1), "workspace too small");
    const int rows_per = R > 0 ? (R + splits - 1) / splits : 0;
    dim3 grid((No + 1 + GT - 1) / GT, (Mo + GT - 1) / GT, splits);
    hipLaunchKernelGGL(wgrad_partial_kernel, grid, dim3(256), 0, stream, d_dY, ldd, d_X, ldx, R, Mo, No, rows_per,
                       d_work);
    PCVL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(grid1((int64_t)Mo * (No + 1))), dim3(256), 0, stream, d_work, splits,
                       Mo, No, d_dW, d_db);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mlp_relu_bwd(float *d_dH, const float *d_H, int64_t count, hipStream_t stream)
{
    if (count <= 0) return OK;
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid1(count)), dim3(256), 0, stream, d_dH, d_H, count);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mlp_colsum(const float *d_dY, int M, int N, float *d_db, hipStream_t stream)
{
    if (N <= 0) return OK;
    hipLaunchKernelGGL(colsum_kernel, dim3((N + 63) / 64), dim3(256), 0, stream, d_dY, M, N, d_db);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mlp_ce(const float *d_logits, int M, int m, const int32_t *d_target, const int32_t *d_lo,
                           float scale, float *d_loss, float *d_dlogits, hipStream_t stream)
{
    if (M <= 0) return OK;
    hipLaunchKernelGGL(ce_kernel, dim3(grid1(M)), dim3(256), 0, stream, d_logits, M, m, d_target, d_lo, scale, d_loss,
                       d_dlogits);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mlp_adam(float *d_w, const float *d_g, float *d_m1, float *d_m2, uint16_t *d_wbf16, int64_t count,
                             float lr, float beta1, float beta2, float eps, int step, hipStream_t stream)
{
    PCVL_ARG_CHECK(step >= 1, "Adam step counts from 1");
    if (count <= 0) return OK;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(count)), dim3(256), 0, stream, d_w, d_g, d_m1, d_m2, d_wbf16, count, lr,
                       beta1, beta2, eps, c1, c2);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mlp_to_bf16(const float *d_w, uint16_t *d_wbf16, int64_t count, hipStream_t stream)
{
    if (count <= 0) return OK;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1(count)), dim3(256), 0, stream, d_w, d_wbf16, count);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_features(const int32_t *d_seq, int B, int n, int m, int p, float *d_x, int32_t *d_lo,
                                 hipStream_t stream)
{
    PCVL_ARG_CHECK(p >= 0 && p < n, "step outside [0, n)");
    if (B <= 0) return OK;
    hipLaunchKernelGGL(features_kernel, dim3(grid1(B)), dim3(256), 0, stream, d_seq, B, n, m, p, d_x, d_lo);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_features_all(const int32_t *d_seq, int B, int n, int m, float *d_x, int32_t *d_lo,
                                     int32_t *d_target, hipStream_t stream)
{
    if (B <= 0) return OK;
    hipLaunchKernelGGL(features_all_kernel, dim3(grid1((int64_t)B * n)), dim3(256), 0, stream, d_seq, B, n, m, d_x,
                       d_lo, d_target);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_sample_fused(const uint16_t *d_W1, const float *d_b1, const uint16_t *d_W2,
                                     const float *d_b2, int B, int m, int n, int H, uint64_t seed, uint64_t first,
                                     int32_t *d_seq, double *d_logq, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= FS_MW && n >= 1 && H >= 8 && H <= FS_MAX_H && H % 8 == 0,
                   "fused sampler needs 1 <= m <= 64, n >= 1, 8 <= H <= 256, H % 8 == 0");
    if (B <= 0) return OK;
    PCVL_ARG_CHECK(d_W1 && d_b1 && d_W2 && d_b2 && d_seq && d_logq, "null device buffer");
    const size_t lds = sizeof(float) * ((size_t)H * FS_TB + FS_TB * FS_MW) + sizeof(uint16_t) * (size_t)H * FS_MW;
    const unsigned grid = (unsigned)(((int64_t)B + FS_TB - 1) / FS_TB);
    hipLaunchKernelGGL(sample_fused_kernel, dim3(grid), dim3(FS_THREADS), lds, stream, d_W1, d_b1, d_W2, d_b2, B, m, n,
                       H, seed, first, d_seq, d_logq);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mis_sample_step(const float *d_logits, int B, int m, int n, int p, const int32_t *d_lo,
                                    uint64_t seed, uint64_t first, int32_t *d_seq, double *d_logq, hipStream_t stream)
{
    PCVL_ARG_CHECK(p >= 0 && p < n, "step outside [0, n)");
    if (B <= 0) return OK;
    hipLaunchKernelGGL(sample_step_kernel, dim3(grid1(B)), dim3(256), 0, stream, d_logits, B, m, n, p, d_lo, seed,
                       first, d_seq, d_logq);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}
