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

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_probe_fp64_fma(int blocks, int64_t iters, double *d_out, int64_t *flops, hipStream_t stream)
{
    PCVL_ARG_CHECK(blocks > 0 && iters > 0 && d_out && flops, "bad probe arguments");
    // a close to 1 and b small keep the chains finite for any iteration count
    hipLaunchKernelGGL(probe_fma_kernel, dim3(blocks), dim3(256), 0, stream, iters, 0.999999999, 1e-9, d_out);
    PCVL_HIP_CHECK(hipGetLastError());
    *flops = (int64_t)blocks * 256 * PROBE_CHAINS * iters * 2;
    return OK;
}
