// Measurement helper (no counterpart in the reference): the FP64 FMA rate the
// device sustains.  Each lane runs 16 independent FMA chains in registers, so
// the kernel issues nothing but v_fma_f64 in its loop and its rate is the
// attainable vector FP64 peak at the clock the chip holds under that load.
#include "common.h"
#include "../../include/pcvl_hip.h"

namespace pcvl {

constexpr int PROBE_CHAINS = 16;

__global__ void __launch_bounds__(256) probe_fma_kernel(int64_t iters, double a, double b, double *__restrict__ out)
{
    double c[PROBE_CHAINS];
#pragma unroll
    for (int k = 0; k < PROBE_CHAINS; k++) c[k] = (double)(threadIdx.x + k);
    for (int64_t it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < PROBE_CHAINS; k++) c[k] = fma(c[k], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < PROBE_CHAINS; k++) s += c[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64 matrix cores beside the vector ALU: each wave runs NM independent
// v_mfma_f64_16x16x4f64 accumulators (2,048 flops per instruction) and NF
// independent v_fma_f64 chains per lane in the same loop.  Timing NM-only,
// NF-only and both shows whether the two units overlap.
typedef double probe_d4 __attribute__((ext_vector_type(4)));

template <int NM, int NF>
__global__ void __launch_bounds__(256) probe_mix_kernel(int64_t iters, double a, double b, double *__restrict__ out)
{
    probe_d4 acc[NM > 0 ? NM : 1];
    double c[NF > 0 ? NF : 1];
#pragma unroll
    for (int k = 0; k < NM; k++) acc[k] = probe_d4{(double)k, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NF; k++) c[k] = (double)(threadIdx.x + k);
    // operands near 1 and 1e-9 keep every accumulator finite
    const double ma = 1e-9 * (double)(threadIdx.x & 7), mb = 1e-9;
    for (int64_t it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < NM; k++) acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(ma, mb, acc[k], 0, 0, 0);
#pragma unroll
        for (int k = 0; k < NF; k++) c[k] = fma(c[k], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NM; k++) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
#pragma unroll
    for (int k = 0; k < NF; k++) s += c[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Read bandwidth of a buffer that is read `reps` times over: a working set
// that fits a cache level (4 MB L2 per XCD, 256 MB last-level cache) shows
// that level's rate; a larger one the HBM rate.
__global__ void __launch_bounds__(256) probe_read_kernel(const double2 *__restrict__ p, uint64_t n2, int reps,
                                                         double *__restrict__ out)
{
    double s = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (int r = 0; r < reps; r++)
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
            const double2 v = p[i];
            s += v.x + v.y;
        }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Which XCD (accelerator complex die) runs each workgroup: its hardware
// XCC_ID register, read by lane 0 and stored by a vector store.
__global__ void __launch_bounds__(64) probe_xcc_kernel(int32_t *__restrict__ out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = (int32_t)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15);
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_probe_read(const double *d_buf, uint64_t bytes, int reps, int blocks, double *d_out,
                               hipStream_t stream)
{
    PCVL_ARG_CHECK(d_buf && d_out && bytes >= 16 && reps > 0 && blocks > 0, "bad probe arguments");
    hipLaunchKernelGGL(probe_read_kernel, dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const double2 *>(d_buf), bytes / 16, reps, d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_probe_xcc(int blocks, int32_t *d_out, hipStream_t stream)
{
    PCVL_ARG_CHECK(blocks > 0 && d_out, "bad probe arguments");
    hipLaunchKernelGGL(probe_xcc_kernel, dim3(blocks), dim3(64), 0, stream, d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_probe_fp64_mix(int blocks, int64_t iters, int mode, double *d_out, int64_t *flops_mfma,
                                   int64_t *flops_valu, hipStream_t stream)
{
    PCVL_ARG_CHECK(blocks > 0 && iters > 0 && d_out && flops_mfma && flops_valu, "bad probe arguments");
    PCVL_ARG_CHECK(mode >= 0 && mode <= 2, "mode must be 0 (matrix cores), 1 (vector FMA) or 2 (both)");
    const int nm = mode == 1 ? 0 : 4, nf = mode == 0 ? 0 : 16;
    auto kp = mode == 0 ? probe_mix_kernel<4, 0> : mode == 1 ? probe_mix_kernel<0, 16> : probe_mix_kernel<4, 16>;
    hipLaunchKernelGGL(kp, dim3(blocks), dim3(256), 0, stream, iters, 0.999999999, 1e-9, d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    *flops_mfma = (int64_t)blocks * 4 * nm * iters * 2048;   // 4 waves per block
    *flops_valu = (int64_t)blocks * 256 * nf * iters * 2;
    return OK;
}

extern "C" int pcvl_probe_fp64_fma(int blocks, int64_t iters, double *d_out, int64_t *flops, hipStream_t stream)
{
    PCVL_ARG_CHECK(blocks > 0 && iters > 0 && d_out && flops, "bad probe arguments");
    // a close to 1 and b small keep the chains finite for any iteration count
    hipLaunchKernelGGL(probe_fma_kernel, dim3(blocks), dim3(256), 0, stream, iters, 0.999999999, 1e-9, d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    *flops = (int64_t)blocks * 256 * PROBE_CHAINS * iters * 2;
    return OK;
}
