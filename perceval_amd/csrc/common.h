// Shared helpers for the HIP simulation core: error reporting across the
// C ABI, complex FP64 arithmetic on interleaved doubles, and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace pcvl {

// Thread-local last-error text behind pcvl_last_error().
void set_error(const std::string &msg);

// Status codes returned by every entry point.
enum Status : int {
    OK = 0,
    ERR_ARG = -1,      // invalid argument (size, null pointer, range)
    ERR_HIP = -2,      // a HIP runtime call failed
    ERR_ALLOC = -3,    // host or device allocation failed
    ERR_RANGE = -4,    // result does not fit (e.g. Fock space above 2^62)
};

// Stream-ordered device workspace from the library's own pool on the current
// device (freed with hipFreeAsync).  Freed memory, up to 512 MiB, stays
// mapped in the pool for the next call.
hipError_t ws_alloc(void **p, size_t bytes, hipStream_t stream);

#define PCVL_HIP_CHECK(expr)                                                     \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            ::pcvl::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return ::pcvl::ERR_HIP;                                              \
        }                                                                        \
    } while (0)

#define PCVL_ARG_CHECK(cond, msg)                                                \
    do {                                                                         \
        if (!(cond)) {                                                           \
            ::pcvl::set_error(std::string("invalid argument: ") + (msg));        \
            return ::pcvl::ERR_ARG;                                              \
        }                                                                        \
    } while (0)

// Complex FP64 value on interleaved storage.
struct cd {
    double r, i;
};

__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return {fma(a.r, b.r, -a.i * b.i), fma(a.r, b.i, a.i * b.r)};
}

__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }

// acc += a * b with two fused steps per component.
// |a|^2 * pf * inv in one fixed rounding order: left to the compiler, the
// sum of squares may be contracted into either fused form, so two kernels
// turning the same coefficient into a probability could differ in the last
// bit.  Every SLOS epilogue goes through this.
__device__ __forceinline__ double prob_from_coef(cd a, double pf, double inv)
{
    return fma(a.r, a.r, a.i * a.i) * pf * inv;
}

__device__ __forceinline__ void cfma_acc(cd &acc, cd a, cd b)
{
    acc.r = fma(a.r, b.r, acc.r);
    acc.r = fma(-a.i, b.i, acc.r);
    acc.i = fma(a.r, b.i, acc.i);
    acc.i = fma(a.i, b.r, acc.i);
}

// acc -= a * b.
__device__ __forceinline__ void cfms_acc(cd &acc, cd a, cd b)
{
    acc.r = fma(-a.r, b.r, acc.r);
    acc.r = fma(a.i, b.i, acc.r);
    acc.i = fma(-a.r, b.i, acc.i);
    acc.i = fma(-a.i, b.r, acc.i);
}

// Counter-based uniform in [0,1): SplitMix64 mixing (Steele, Lea & Flood,
// OOPSLA 2014) of (seed, stream, counter).  The same definition is used by
// every sampler in this library so a run is fully determined by its seed.
__host__ __device__ __forceinline__ uint64_t smix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// 64 random bits of (seed, stream, counter).
__host__ __device__ __forceinline__ uint64_t rand64(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const uint64_t key = smix64(seed ^ smix64(stream));
    return smix64(key + ctr * 0x9E3779B97F4A7C15ull);
}

__host__ __device__ __forceinline__ double uniform01(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    return (double)(rand64(seed, stream, ctr) >> 11) * 0x1.0p-53;
}

} // namespace pcvl
