// Batched complex permanents by Glynn's formula in Gray-code order.
//
// Replaces xq.permanent_cx as called from NaiveBackend._compute_permanent
// (perceval/backends/_naive.py:70-71).  Glynn (2010):
//     perm(A) = 2^{-(n-1)} sum_{delta in {+-1}^n, delta_0 = +1} (prod_i delta_i) prod_j sum_i delta_i a_ij
// Step t of the binary-reflected Gray code flips row 1 + ctz(t); the sign of
// term t is (-1)^t.
//
// Work decomposition (orders 9..28): the 2^{n-1} steps of a matrix are cut
// into lane chunks of 2^c steps; one wavefront takes 64 consecutive chunks.
// Because a chunk starts at a multiple of 2^c, every lane flips the same row
// at the same step, so the flipped row is wave-uniform and is read from LDS
// with a broadcast address; only the flip direction of rows >= 2 differs
// between lanes and is applied as a per-lane +-1 FMA.  Row 1 (flipped every
// other step) stays in VGPRs with compile-time flip directions.  Column sums
// are rebuilt from scratch at each chunk start, which bounds the drift of the
// running sums to 2^c updates.  One persistent launch drains a cost-sorted
// queue of wave items of mixed order (largest first), then a second kernel
// sums each matrix's wave partials in a fixed order (bit-reproducible).
// Orders 0..8 take one matrix per thread.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <mutex>
#include <vector>

namespace pcvl {

constexpr int PERM_WAVE_MIN = 9;
constexpr int PERM_WAVE_MAX = 28;   // the mixed-order launch
// Orders 29..32 run in a launch of their own (their registers would lower the
// occupancy of the mixed launch), from their own part of the queue.
constexpr int PERM_BIG_MIN = PERM_WAVE_MAX + 1;
constexpr int PERM_BIG_MAX = 32;
constexpr int PERM_SMALL_MAX = 8;

struct PermItem {
    int64_t mat_off;   // complex offset of the matrix in the packed array
    uint32_t chunk0;   // first lane chunk of this wave
    int32_t slot;      // partial-sum slot written by this wave
    int32_t n;         // order
    int32_t c;         // log2 of Gray-code steps per lane
};

// ---------------------------------------------------------------- products

// Orders one Gray-code step after the previous term: the column sums are
// passed through an empty asm that also reads the accumulator, so no update
// of cs can be hoisted above the product that reads the old cs.  Without it
// the optimiser computes all column-sum generations of an unrolled block
// first and keeps them live together (4x the VGPRs, then spills).
template <int N>
__device__ __forceinline__ void pin_after(cd *cs, const cd &acc)
{
#pragma unroll
    for (int j = 0; j < N; j++)
        asm volatile("" : "+v"(cs[j].r), "+v"(cs[j].i) : "v"(acc.r), "v"(acc.i));
}

// prod_{j < N-1} cs[j] as four independent chains (ILP for FP64 latency).
template <int N>
__device__ __forceinline__ cd prod_head(const cd *cs)
{
    static_assert(N >= 5, "wave path needs N >= 5");
    cd p0 = cs[0], p1 = cs[1], p2 = cs[2], p3 = cs[3];
#pragma unroll
    for (int j = 4; j < N - 1; j++) {
        if ((j & 3) == 0) p0 = cmul(p0, cs[j]);
        else if ((j & 3) == 1) p1 = cmul(p1, cs[j]);
        else if ((j & 3) == 2) p2 = cmul(p2, cs[j]);
        else p3 = cmul(p3, cs[j]);
    }
    return cmul(cmul(p0, p1), cmul(p2, p3));
}

template <int N>
__device__ __forceinline__ void term_add(cd &acc, const cd *cs)
{
    cfma_acc(acc, prod_head<N>(cs), cs[N - 1]);
}

template <int N>
__device__ __forceinline__ void term_sub(cd &acc, const cd *cs)
{
    cfms_acc(acc, prod_head<N>(cs), cs[N - 1]);
}

// cs += s * row, row read from LDS at a wave-uniform address.
template <int N>
__device__ __forceinline__ void flip_lds(cd *cs, const cd *__restrict__ row, double s)
{
#pragma unroll
    for (int j = 0; j < N; j++) {
        cd a = row[j];
        cs[j].r = fma(s, a.r, cs[j].r);
        cs[j].i = fma(s, a.i, cs[j].i);
    }
}

// c += s * a, placed after dep: the empty asm makes the update wait for the
// product that consumed the old c, so both generations are never live.
__device__ __forceinline__ void flip_one(cd &c, const cd &a, double s, const cd &dep)
{
    // Not volatile: a volatile asm has unmodelled side effects, which would
    // also stop the scheduler from issuing LDS reads of the row above it.
    asm("" : "+v"(c.r), "+v"(c.i) : "v"(dep.r), "v"(dep.i));
    c.r = fma(s, a.r, c.r);
    c.i = fma(s, a.i, c.i);
}

// Returns v, made to look dependent on dep: keeps loads at an offset of v
// inside the loop that dep counts (they would otherwise be hoisted as loop
// invariant, holding whole rows in VGPRs).
__device__ __forceinline__ int loop_variant(int v, uint32_t dep)
{
    asm("" : "+s"(v) : "s"(dep));
    return v;
}

// One Gray-code term (added, or subtracted when SUB) fused with the flip to the
// next step: column j is flipped as soon as its product chain has consumed it,
// so row reads from LDS spread over the product's arithmetic instead of
// forming a separate latency-bound phase.
template <int N, bool SUB>
__device__ __forceinline__ void term_flip(cd &acc, cd *cs, const cd *row, double s)
{
    static_assert(N >= 9, "every chain needs a second factor");
    cd p0 = cs[0], p1 = cs[1], p2 = cs[2], p3 = cs[3];
#pragma unroll
    for (int j = 4; j < N - 1; j++) {
        cd &pk = ((j & 3) == 0) ? p0 : ((j & 3) == 1) ? p1 : ((j & 3) == 2) ? p2 : p3;
        pk = cmul(pk, cs[j]);
        if (j < 8) flip_one(cs[j - 4], row[j - 4], s, pk);
        flip_one(cs[j], row[j], s, pk);
    }
    const cd h = cmul(cmul(p0, p1), cmul(p2, p3));
    if (SUB) cfms_acc(acc, h, cs[N - 1]);
    else cfma_acc(acc, h, cs[N - 1]);
    flip_one(cs[N - 1], row[N - 1], s, acc);
}

// --------------------------------------------------------------- wave path

// One wave item: 64 lanes x 2^c steps of one order-N matrix.  sA is this
// wave's LDS slice holding 2*A (row-major).
template <int N, int R1 = 0>
__device__ __forceinline__ cd glynn_wave_item(const double *__restrict__ mats, const PermItem &it, cd *sA)
{
    const int lane = threadIdx.x & 63;
    const cd *A = reinterpret_cast<const cd *>(mats) + it.mat_off;
    for (int e = lane; e < N * N; e += 64) {
        cd v = A[e];
        sA[e] = {2.0 * v.r, 2.0 * v.i};
    }
    __syncthreads();

    const int c = it.c;
    const uint64_t steps = 1ull << c;
    const uint64_t base = (uint64_t)(it.chunk0 + (uint32_t)lane) << c;
    const uint64_t g0 = base ^ (base >> 1);

    // Column sums for delta = Gray(base): cs_j = (1/2) sum_i delta_i (2 a_ij).
    cd cs[N];
#pragma unroll
    for (int j = 0; j < N; j++) cs[j] = sA[j];
#pragma unroll 1
    for (int i = 1; i < N; i++) {
        const double d = ((g0 >> (i - 1)) & 1) ? -1.0 : 1.0;
        const cd *row = sA + i * N;
#pragma unroll
        for (int j = 0; j < N; j++) {
            cd a = row[j];
            cs[j].r = fma(d, a.r, cs[j].r);
            cs[j].i = fma(d, a.i, cs[j].i);
        }
    }
#pragma unroll
    for (int j = 0; j < N; j++) cs[j] = {0.5 * cs[j].r, 0.5 * cs[j].i};

    // Row 1 flips every other step: keep it in VGPRs (R1 = 1, 3), read it
    // from LDS like the other rows (R1 = 2, 4: fewer VGPRs and more waves), or
    // pick by order (R1 = 0: in VGPRs up to order 26, the largest that fits
    // without spilling; at orders 23-24 that is 5% faster than three waves
    // per SIMD reading it from LDS).  R1 = 3, 4 fuse each flip into the
    // preceding term.
    constexpr bool kRow1InRegs = (R1 == 1) || (R1 == 3) || (R1 == 0 && N <= 26);
    constexpr bool kFused = (R1 >= 3);
    cd r1[kRow1InRegs ? N : 1];
    if constexpr (kRow1InRegs) {
#pragma unroll
        for (int j = 0; j < N; j++) r1[j] = sA[N + j];
    }

    cd acc = {0.0, 0.0};
    if constexpr (kFused) {
        for (uint64_t u = 0; u < steps; u += 4) {
            const cd *row1 = kRow1InRegs ? r1 : sA + loop_variant(N, (uint32_t)u);
            term_flip<N, false>(acc, cs, row1, -1.0);          // t = base+u, then row 1 -> delta -1
            uint64_t t2 = base + u + 2;
            const double s2 = (((t2 >> 1) ^ (t2 >> 2)) & 1) ? -1.0 : 1.0;
            term_flip<N, true>(acc, cs, sA + loop_variant(2 * N, (uint32_t)u), s2);   // t+1, then row 2
            term_flip<N, false>(acc, cs, row1, 1.0);           // t+2, then row 1 -> delta +1
            // t+3, then row 1 + ctz(u+4) >= 3 (on the last block: first step of
            // the next chunk, row c+1 <= N-6, result discarded)
            const int b = __builtin_ctzll(u + 4);
            uint64_t t4 = base + u + 4;
            const double s4 = (((t4 >> b) ^ (t4 >> (b + 1))) & 1) ? -1.0 : 1.0;
            term_flip<N, true>(acc, cs, sA + (b + 1) * N, s4);
        }
        __syncthreads();
        return acc;
    }
    // Steps are taken four at a time; c >= 2 so base has two zero low bits.
    for (uint64_t u = 0; u < steps; u += 4) {
        term_add<N>(acc, cs);                 // t = base+u, even
        pin_after<N>(cs, acc);
        if constexpr (kRow1InRegs) {          // t = base+u+1: row 1 -> delta -1
#pragma unroll
            for (int j = 0; j < N; j++) { cs[j].r -= r1[j].r; cs[j].i -= r1[j].i; }
        } else {
            flip_lds<N>(cs, sA + N, -1.0);
        }
        term_sub<N>(acc, cs);
        pin_after<N>(cs, acc);
        {                                     // t = base+u+2: row 2
            uint64_t t = base + u + 2;
            double s = (((t >> 1) ^ (t >> 2)) & 1) ? -1.0 : 1.0;
            flip_lds<N>(cs, sA + 2 * N, s);
        }
        term_add<N>(acc, cs);
        pin_after<N>(cs, acc);
        if constexpr (kRow1InRegs) {          // t = base+u+3: row 1 -> delta +1
#pragma unroll
            for (int j = 0; j < N; j++) { cs[j].r += r1[j].r; cs[j].i += r1[j].i; }
        } else {
            flip_lds<N>(cs, sA + N, 1.0);
        }
        term_sub<N>(acc, cs);
        pin_after<N>(cs, acc);
        {
            // t = base+u+4: row 1 + ctz(u+4) >= 3.  On the last block this is
            // the first step of the next chunk (row c+1 <= N-6, in bounds); its
            // result is discarded, and taking it unconditionally keeps the
            // loop body a single basic block.
            const int b = __builtin_ctzll(u + 4);
            uint64_t t = base + u + 4;
            double s = (((t >> b) ^ (t >> (b + 1))) & 1) ? -1.0 : 1.0;
            flip_lds<N>(cs, sA + (b + 1) * N, s);
        }
    }
    __syncthreads();   // sA is reused by the next item of this wave
    return acc;
}

__device__ __forceinline__ cd wave_sum(cd v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v.r += __shfl_xor(v.r, off, 64);
        v.i += __shfl_xor(v.i, off, 64);
    }
    return v;
}

template <int NLO, int NHI, int R1>
__device__ __forceinline__ cd dispatch_item(const double *__restrict__ mats, const PermItem &it, cd *sA)
{
    if constexpr (NLO > NHI) {
        return {0.0, 0.0};
    } else {
        if (it.n == NLO) return glynn_wave_item<NLO, R1>(mats, it, sA);
        return dispatch_item<NLO + 1, NHI, R1>(mats, it, sA);
    }
}

// Persistent wave kernel: one wavefront per workgroup, items pulled from a
// cost-sorted queue with an atomic cursor.  The cursor is returned to zero by
// perm_finalize so the same plan can be replayed (also inside a HIP graph).
template <int NLO, int NHI, int R1 = 0>
__global__ void __launch_bounds__(64, 2) perm_wave_kernel(const double *__restrict__ mats,
                                                       const PermItem *__restrict__ items, int64_t nitems,
                                                       unsigned long long *cursor, cd *__restrict__ partials)
{
    __shared__ cd sA[NHI * NHI];
    // A wave can take at most nitems items, so the loop is bounded whatever
    // the cursor holds; items whose order or chunk size lies outside what
    // this instantiation handles are skipped (their slot is left untouched).
    for (int64_t iter = 0; iter < nitems; iter++) {
        unsigned long long k = 0;
        if (threadIdx.x == 0) k = atomicAdd(cursor, 1ull);
        k = __shfl(k, 0, 64);
        // The item index and every field of the item are wave-uniform; say so,
        // so that loop control and LDS row addresses live in SGPRs.
        const uint32_t k_lo = __builtin_amdgcn_readfirstlane((uint32_t)k);
        const uint32_t k_hi = __builtin_amdgcn_readfirstlane((uint32_t)(k >> 32));
        const uint64_t ku = ((uint64_t)k_hi << 32) | k_lo;
        if (ku >= (uint64_t)nitems) break;
        PermItem it;
        {
            const PermItem raw = items[ku];
            it.mat_off = (int64_t)(((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)raw.mat_off >> 32)) << 32) |
                                   __builtin_amdgcn_readfirstlane((uint32_t)raw.mat_off));
            it.chunk0 = __builtin_amdgcn_readfirstlane(raw.chunk0);
            it.slot = __builtin_amdgcn_readfirstlane(raw.slot);
            it.n = __builtin_amdgcn_readfirstlane(raw.n);
            it.c = __builtin_amdgcn_readfirstlane(raw.c);
        }
        if (it.n < NLO || it.n > NHI || it.c < 2 || it.c > it.n - 7) continue;
        cd v = dispatch_item<NLO, NHI, R1>(mats, it, sA);
        v = wave_sum(v);
        if (threadIdx.x == 0) partials[it.slot] = v;
    }
}

// --------------------------------------------------------------- small path

template <int N>
__device__ cd glynn_small(const cd *__restrict__ A)
{
    if constexpr (N == 0) {
        return {1.0, 0.0};
    } else {
        cd cs[N];
#pragma unroll
        for (int j = 0; j < N; j++) {
            cd s = {0.0, 0.0};
            for (int i = 0; i < N; i++) s = cadd(s, A[i * N + j]);
            cs[j] = s;
        }
        cd acc = {0.0, 0.0};
        uint32_t g = 0;
        for (uint32_t t = 0; t < (1u << (N - 1)); t++) {
            if (t) {
                int b = __builtin_ctz(t);
                g ^= 1u << b;
                double s = ((g >> b) & 1) ? -2.0 : 2.0;
                const cd *row = A + (b + 1) * N;
#pragma unroll
                for (int j = 0; j < N; j++) {
                    cs[j].r = fma(s, row[j].r, cs[j].r);
                    cs[j].i = fma(s, row[j].i, cs[j].i);
                }
            }
            cd p = cs[0];
#pragma unroll
            for (int j = 1; j < N; j++) p = cmul(p, cs[j]);
            if (t & 1) { acc.r -= p.r; acc.i -= p.i; } else { acc.r += p.r; acc.i += p.i; }
        }
        double sc = ldexp(1.0, -(N - 1));
        return {acc.r * sc, acc.i * sc};
    }
}

template <int N>
__device__ __forceinline__ cd dispatch_small(int n, const cd *A)
{
    if constexpr (N > PERM_SMALL_MAX) {
        return {0.0, 0.0};
    } else {
        if (n == N) return glynn_small<N>(A);
        return dispatch_small<N + 1>(n, A);
    }
}

__global__ void perm_small_kernel(const double *__restrict__ mats, const int64_t *__restrict__ idx,
                                  const int64_t *__restrict__ offs, const int32_t *__restrict__ ns,
                                  int64_t count, cd *__restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    int64_t b = idx[k];
    const cd *A = reinterpret_cast<const cd *>(mats) + offs[k];
    out[b] = dispatch_small<0>(ns[k], A);
}

// Sum each wave-path matrix's partials in slot order, scale by 2^{-(n-1)},
// and return the queue cursor to zero for the next replay.
__global__ void perm_finalize_kernel(const int64_t *__restrict__ idx, const int32_t *__restrict__ slot0,
                                     const int32_t *__restrict__ nslot, const int32_t *__restrict__ ns,
                                     int64_t count, const cd *__restrict__ partials, cd *__restrict__ out,
                                     unsigned long long *cursor)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) { cursor[0] = 0ull; cursor[1] = 0ull; }
    if (k >= count) return;
    cd s = {0.0, 0.0};
    for (int q = 0; q < nslot[k]; q++) s = cadd(s, partials[slot0[k] + q]);
    double sc = ldexp(1.0, -(ns[k] - 1));
    out[idx[k]] = {s.r * sc, s.i * sc};
}

// ------------------------------------------------------------------- plans

struct PermPlanImpl {
    int64_t count = 0;
    int64_t nitems = 0, nwave_mats = 0, nsmall = 0;
    int64_t nbig = 0;          // items of orders 29..32, after the nitems others in the queue
    int grid_big = 0;
    int single_n = 0;          // order shared by every wave item, or 0 if mixed
    int variant = 0;           // 0 auto, 1..4 (single-order plans): row 1 in VGPRs / LDS, separate / fused flips
    int grid = 0;
    PermItem *d_items = nullptr;
    unsigned long long *d_cursor = nullptr;   // [0]: the queue of orders <= 28, [1]: orders 29..32
    cd *d_partials = nullptr;
    int64_t *d_wave_idx = nullptr;
    int32_t *d_wave_slot0 = nullptr, *d_wave_nslot = nullptr, *d_wave_n = nullptr;
    int64_t *d_small_idx = nullptr, *d_small_off = nullptr;
    int32_t *d_small_n = nullptr;
    void *d_block = nullptr;   // one allocation backing every array above
};

static int pick_chunk_log2(int n, int max_c)
{
    // 2^{n-1} steps = 2^{n-7} lane chunks of 2^{n-7-w} steps per wave when a
    // matrix is split over 2^w waves; take the longest chunk up to 2^max_c.
    int c = n - 7;
    if (c > max_c) c = max_c;
    if (c < 2) c = 2;
    return c;
}

template <int NLO, int NHI, int R1>
static int launch_wave(const PermPlanImpl *p, const double *mats, hipStream_t st, const PermItem *items,
                       int64_t nitems, unsigned long long *cursor, int grid)
{
    hipLaunchKernelGGL((perm_wave_kernel<NLO, NHI, R1>), dim3(grid), dim3(64), 0, st, mats, items, nitems, cursor,
                       p->d_partials);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

template <int NLO, int NHI, int R1>
static int launch_wave(const PermPlanImpl *p, const double *mats, hipStream_t st)
{
    return launch_wave<NLO, NHI, R1>(p, mats, st, p->d_items, p->nitems, p->d_cursor, p->grid);
}

template <int N>
static int launch_single(PermPlanImpl *p, const double *mats, hipStream_t st)
{
    if constexpr (N > PERM_WAVE_MAX) {
        set_error("internal: order out of range");
        return ERR_ARG;
    } else {
        if (p->single_n == N) {
            if (p->variant == 1) return launch_wave<N, N, 1>(p, mats, st);
            if (p->variant == 2) return launch_wave<N, N, 2>(p, mats, st);
            if (p->variant == 3) return launch_wave<N, N, 3>(p, mats, st);
            if (p->variant == 4) return launch_wave<N, N, 4>(p, mats, st);
            return launch_wave<N, N, 0>(p, mats, st);
        }
        return launch_single<N + 1>(p, mats, st);
    }
}

template <int N>
static const void *single_kernel_ptr(int n, int variant)
{
    if constexpr (N > PERM_WAVE_MAX) {
        return nullptr;
    } else {
        if (n == N) {
            if (variant == 1) return reinterpret_cast<const void *>(&perm_wave_kernel<N, N, 1>);
            if (variant == 2) return reinterpret_cast<const void *>(&perm_wave_kernel<N, N, 2>);
            if (variant == 3) return reinterpret_cast<const void *>(&perm_wave_kernel<N, N, 3>);
            if (variant == 4) return reinterpret_cast<const void *>(&perm_wave_kernel<N, N, 4>);
            return reinterpret_cast<const void *>(&perm_wave_kernel<N, N, 0>);
        }
        return single_kernel_ptr<N + 1>(n, variant);
    }
}

static int resident_waves(const void *kernel, int *out)
{
    int dev = 0, cus = 0, per_cu = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0));
    if (per_cu < 1) per_cu = 1;
    *out = cus * per_cu;
    return OK;
}

} // namespace pcvl

using namespace pcvl;

struct pcvl_perm_plan_s {
    PermPlanImpl impl;
};

// Work items above this count are refused (24 bytes each on the host and the
// device); partial-sum slots are int32 on the device.
static constexpr int64_t PERM_MAX_ITEMS = int64_t(1) << 27;

// Device image of a plan: one block holding items | cursor | partials | wave
// arrays | small arrays, at these byte offsets.
struct PermLayout {
    size_t o_items = 0, o_cursor = 0, o_part = 0, o_widx = 0, o_ws0 = 0, o_wns = 0, o_wn = 0, o_sidx = 0,
           o_soff = 0, o_sn = 0, total = 0;
};

// Builds a plan on the host: the scheduling fields of p, the layout of its
// device block and the block's contents (cursor zero).  No device call but
// occupancy queries.
static int perm_plan_build(const int32_t *h_n, const int64_t *h_offsets, int64_t count, int max_chunk_log2,
                           int variant, PermPlanImpl &p, PermLayout &L, std::vector<char> &host)
{
    PCVL_ARG_CHECK(variant >= 0 && variant <= 4, "variant outside [0, 4]");
    PCVL_ARG_CHECK(count >= 0, "count < 0");
    PCVL_ARG_CHECK(count == 0 || (h_n && h_offsets), "null order/offset array");
    if (max_chunk_log2 <= 0) max_chunk_log2 = 11;
    PCVL_ARG_CHECK(max_chunk_log2 >= 2 && max_chunk_log2 <= 24, "max_chunk_log2 outside [2, 24]");

    std::vector<PermItem> items;
    std::vector<int64_t> wave_idx, small_idx, small_off;
    std::vector<int32_t> wave_slot0, wave_nslot, wave_n, small_n;
    std::vector<int> wave_c;
    int nmin = 1 << 30, nmax = -1;
    bool any_big = false;
    for (int64_t b = 0; b < count; b++) {
        int n = h_n[b];
        PCVL_ARG_CHECK(n >= 0 && n <= PERM_BIG_MAX, "matrix order outside [0, 32]");
        PCVL_ARG_CHECK(h_offsets[b] >= 0, "negative matrix offset");
        if (n <= PERM_SMALL_MAX) {
            small_idx.push_back(b);
            small_off.push_back(h_offsets[b]);
            small_n.push_back(n);
            continue;
        }
        wave_idx.push_back(b);
        wave_n.push_back(n);
        wave_c.push_back(pick_chunk_log2(n, max_chunk_log2));
        if (n >= PERM_BIG_MIN) {
            any_big = true;
            continue;
        }
        nmin = std::min(nmin, n);
        nmax = std::max(nmax, n);
    }
    // orders <= 28: one launch, specialised when they share one order
    const int single_n = (nmax >= 0 && nmin == nmax) ? nmax : 0;
    const int kvariant = single_n ? variant : 0;
    int rw = 0, rwb = 0;
    if (nmax >= 0) {
        const void *kp = single_n ? single_kernel_ptr<PERM_WAVE_MIN>(single_n, kvariant)
                                  : reinterpret_cast<const void *>(&perm_wave_kernel<PERM_WAVE_MIN, PERM_WAVE_MAX, 0>);
        int rc = resident_waves(kp, &rw);
        if (rc != OK) return rc;
    }
    if (any_big) {
        int rc = resident_waves(reinterpret_cast<const void *>(&perm_wave_kernel<PERM_BIG_MIN, PERM_BIG_MAX, 0>), &rwb);
        if (rc != OK) return rc;
    }
    // Tail refinement, per launch: the matrices whose items form the last
    // round of its queue (the last `resident waves` items in cost order) use
    // chunks 4x shorter, so the waves that finish last wait for at most a
    // quarter item.
    for (int big = 0; big < 2; big++) {
        std::vector<size_t> order;
        for (size_t i = 0; i < wave_idx.size(); i++)
            if ((wave_n[i] >= PERM_BIG_MIN) == (big == 1)) order.push_back(i);
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return ((int64_t)wave_n[a] << wave_c[a]) > ((int64_t)wave_n[b] << wave_c[b]);
        });
        const int resident = big ? rwb : rw;
        int64_t tail = 0;
        for (size_t r = order.size(); r-- > 0 && tail < resident;) {
            const size_t i = order[r];
            tail += (int64_t)((1ull << (wave_n[i] - 1 - wave_c[i])) / 64);
            wave_c[i] = std::max(2, wave_c[i] - 2);
        }
    }
    int64_t total_slots = 0;
    for (size_t i = 0; i < wave_idx.size(); i++)
        total_slots += (int64_t)((1ull << (wave_n[i] - 1 - wave_c[i])) / 64);
    if (total_slots > PERM_MAX_ITEMS) {
        set_error("permanent batch needs " + std::to_string(total_slots) + " work items, above the limit of " +
                  std::to_string(PERM_MAX_ITEMS) + " (use a longer chunk or split the batch)");
        return ERR_RANGE;
    }
    items.reserve((size_t)total_slots);
    std::vector<PermItem> big_items;
    int32_t slots = 0;
    for (size_t i = 0; i < wave_idx.size(); i++) {
        const int n = wave_n[i], c = wave_c[i];
        const int64_t waves = (int64_t)((1ull << (n - 1 - c)) / 64);
        wave_slot0.push_back(slots);
        wave_nslot.push_back((int32_t)waves);
        std::vector<PermItem> &dst = n >= PERM_BIG_MIN ? big_items : items;
        for (int64_t w = 0; w < waves; w++)
            dst.push_back(PermItem{h_offsets[wave_idx[i]], (uint32_t)(w * 64), slots + (int32_t)w, n, c});
        slots += (int32_t)waves;
    }
    // Largest items first: the atomic queue then gives a greedy
    // longest-processing-time schedule over the resident waves.  The queue
    // holds the items of orders <= 28, then those of orders 29..32.
    auto by_cost = [](const PermItem &a, const PermItem &b) {
        return ((int64_t)a.n << a.c) > ((int64_t)b.n << b.c);
    };
    std::stable_sort(items.begin(), items.end(), by_cost);
    std::stable_sort(big_items.begin(), big_items.end(), by_cost);
    const int64_t nregular = (int64_t)items.size();
    items.insert(items.end(), big_items.begin(), big_items.end());

    p.count = count;
    p.nitems = nregular;
    p.nbig = (int64_t)big_items.size();
    p.grid_big = p.nbig > 0 ? (int)std::min<int64_t>(p.nbig, (int64_t)rwb) : 0;
    p.nwave_mats = (int64_t)wave_idx.size();
    p.nsmall = (int64_t)small_idx.size();
    p.single_n = single_n;
    p.variant = kvariant;
    p.grid = p.nitems > 0 ? (int)std::min<int64_t>(p.nitems, (int64_t)rw) : 0;

    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    L.o_items = 0;
    L.o_cursor = align(L.o_items + sizeof(PermItem) * items.size());
    L.o_part = align(L.o_cursor + 2 * sizeof(unsigned long long));
    L.o_widx = align(L.o_part + sizeof(cd) * (size_t)slots);
    L.o_ws0 = align(L.o_widx + sizeof(int64_t) * wave_idx.size());
    L.o_wns = align(L.o_ws0 + sizeof(int32_t) * wave_idx.size());
    L.o_wn = align(L.o_wns + sizeof(int32_t) * wave_idx.size());
    L.o_sidx = align(L.o_wn + sizeof(int32_t) * wave_idx.size());
    L.o_soff = align(L.o_sidx + sizeof(int64_t) * small_idx.size());
    L.o_sn = align(L.o_soff + sizeof(int64_t) * small_idx.size());
    L.total = align(L.o_sn + sizeof(int32_t) * small_idx.size());

    host.assign(L.total, 0);
    auto put = [&](size_t off, const void *src, size_t bytes) { if (bytes) std::memcpy(host.data() + off, src, bytes); };
    put(L.o_items, items.data(), sizeof(PermItem) * items.size());
    put(L.o_widx, wave_idx.data(), sizeof(int64_t) * wave_idx.size());
    put(L.o_ws0, wave_slot0.data(), sizeof(int32_t) * wave_idx.size());
    put(L.o_wns, wave_nslot.data(), sizeof(int32_t) * wave_idx.size());
    put(L.o_wn, wave_n.data(), sizeof(int32_t) * wave_idx.size());
    put(L.o_sidx, small_idx.data(), sizeof(int64_t) * small_idx.size());
    put(L.o_soff, small_off.data(), sizeof(int64_t) * small_idx.size());
    put(L.o_sn, small_n.data(), sizeof(int32_t) * small_idx.size());
    return OK;
}

// Points the plan's device arrays into a block laid out as L.
static void perm_plan_bind(PermPlanImpl &p, const PermLayout &L, void *block)
{
    char *base = static_cast<char *>(block);
    p.d_items = reinterpret_cast<PermItem *>(base + L.o_items);
    p.d_cursor = reinterpret_cast<unsigned long long *>(base + L.o_cursor);
    p.d_partials = reinterpret_cast<cd *>(base + L.o_part);
    p.d_wave_idx = reinterpret_cast<int64_t *>(base + L.o_widx);
    p.d_wave_slot0 = reinterpret_cast<int32_t *>(base + L.o_ws0);
    p.d_wave_nslot = reinterpret_cast<int32_t *>(base + L.o_wns);
    p.d_wave_n = reinterpret_cast<int32_t *>(base + L.o_wn);
    p.d_small_idx = reinterpret_cast<int64_t *>(base + L.o_sidx);
    p.d_small_off = reinterpret_cast<int64_t *>(base + L.o_soff);
    p.d_small_n = reinterpret_cast<int32_t *>(base + L.o_sn);
}

static int perm_plan_create_impl(const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                                 int max_chunk_log2, int variant, pcvl_perm_plan *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out plan pointer is null");
    *out = nullptr;
    auto *plan = new (std::nothrow) pcvl_perm_plan_s();
    if (!plan) { set_error("host allocation failed"); return ERR_ALLOC; }
    PermPlanImpl &p = plan->impl;
    PermLayout L;
    std::vector<char> host;
    int rc = perm_plan_build(h_n, h_offsets, count, max_chunk_log2, variant, p, L, host);
    if (rc != OK) { delete plan; return rc; }
    hipError_t e = hipMalloc(&p.d_block, L.total);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(perm plan): ") + hipGetErrorString(e));
        delete plan;
        return ERR_ALLOC;
    }
    perm_plan_bind(p, L, p.d_block);
    e = hipMemcpy(p.d_block, host.data(), L.total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("hipMemcpy(perm plan): ") + hipGetErrorString(e));
        (void)hipFree(p.d_block);
        delete plan;
        return ERR_HIP;
    }
    *out = plan;
    return OK;
}

extern "C" int pcvl_perm_plan_create_ex(const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                                        int max_chunk_log2, int variant, pcvl_perm_plan *out)
{
    try {
        return perm_plan_create_impl(h_n, h_offsets, count, max_chunk_log2, variant, out);
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed while building the permanent plan");
        return ERR_ALLOC;
    } catch (...) {
        set_error("unexpected exception while building the permanent plan");
        return ERR_ALLOC;
    }
}

extern "C" int pcvl_perm_plan_create(const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                                     int max_chunk_log2, pcvl_perm_plan *out)
{
    return pcvl_perm_plan_create_ex(h_n, h_offsets, count, max_chunk_log2, 0, out);
}

extern "C" int pcvl_perm_plan_execute(pcvl_perm_plan plan, const double *d_mats, double *d_out, hipStream_t stream)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    PermPlanImpl &p = plan->impl;
    if (p.count == 0) return OK;
    PCVL_ARG_CHECK(d_mats && d_out, "null device buffer");
    cd *out = reinterpret_cast<cd *>(d_out);
    if (p.nsmall > 0) {
        int blocks = (int)((p.nsmall + 255) / 256);
        hipLaunchKernelGGL(perm_small_kernel, dim3(blocks), dim3(256), 0, stream, d_mats, p.d_small_idx,
                           p.d_small_off, p.d_small_n, p.nsmall, out);
        PCVL_HIP_CHECK(hipGetLastError());
    }
    if (p.nwave_mats > 0) {
        int rc = OK;
        if (p.nbig > 0)   // orders 29..32 first: the longest items
            rc = launch_wave<PERM_BIG_MIN, PERM_BIG_MAX, 0>(&p, d_mats, stream, p.d_items + p.nitems, p.nbig,
                                                            p.d_cursor + 1, p.grid_big);
        if (rc == OK && p.nitems > 0)
            rc = p.single_n ? launch_single<PERM_WAVE_MIN>(&p, d_mats, stream)
                            : launch_wave<PERM_WAVE_MIN, PERM_WAVE_MAX, 0>(&p, d_mats, stream);
        if (rc != OK) return rc;
        int blocks = (int)((p.nwave_mats + 255) / 256);
        hipLaunchKernelGGL(perm_finalize_kernel, dim3(blocks), dim3(256), 0, stream, p.d_wave_idx,
                           p.d_wave_slot0, p.d_wave_nslot, p.d_wave_n, p.nwave_mats, p.d_partials, out,
                           p.d_cursor);
        PCVL_HIP_CHECK(hipGetLastError());
    }
    return OK;
}

extern "C" int pcvl_perm_plan_destroy(pcvl_perm_plan plan)
{
    if (!plan) return OK;
    hipError_t e = hipSuccess;
    if (plan->impl.d_block) e = hipFree(plan->impl.d_block);
    delete plan;
    if (e != hipSuccess) {
        set_error(std::string("hipFree(perm plan): ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    return OK;
}

extern "C" int pcvl_perm_plan_info(pcvl_perm_plan plan, int64_t *nitems, int *grid, int *single_n)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    if (nitems) *nitems = plan->impl.nitems + plan->impl.nbig;
    if (grid) *grid = plan->impl.grid;
    if (single_n) *single_n = plan->impl.single_n;
    return OK;
}

extern "C" int pcvl_perm_batch(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                               double *d_out, hipStream_t stream)
{
    pcvl_perm_plan plan = nullptr;
    int rc = pcvl_perm_plan_create(h_n, h_offsets, count, 0, &plan);
    if (rc != OK) return rc;
    rc = pcvl_perm_plan_execute(plan, d_mats, d_out, stream);
    hipError_t e = hipStreamSynchronize(stream);
    int rc2 = pcvl_perm_plan_destroy(plan);
    if (rc != OK) return rc;
    if (e != hipSuccess) { set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e)); return ERR_HIP; }
    return rc2;
}

// ------------------------------------------------------- no allocation
//
// pcvl_perm_batch_noalloc: the same batch with no device allocation.  The
// plan's block (work items, queue cursor, partial sums, per-matrix arrays)
// goes into `scratch`, a file-scope device array that each GPU holds from
// module load until the process ends.  Host code never passes the array's
// name to a kernel: its device address is looked up once per device with
// hipGetSymbolAddress, on the entry point's first call on that device, and
// that pointer is used for the copy and the kernel arguments.  Every call on
// a GPU shares its array: calls that overlap on one GPU (several host threads
// or streams) race.  Each call rewrites the whole block it uses, the queue
// cursor included (zero), before its kernels read it.
constexpr size_t PERM_SCRATCH_BYTES = size_t(16) << 20;
static __device__ __align__(256) char scratch[PERM_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};

static int scratch_ptr(void **out)
{
    int dev = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) { set_error("device index outside [0, 64)"); return ERR_ARG; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

static int perm_batch_noalloc_impl(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets, int64_t count,
                                   double *d_out, hipStream_t stream)
{
    pcvl_perm_plan_s plan;
    PermPlanImpl &p = plan.impl;
    PermLayout L;
    std::vector<char> host;
    int rc = perm_plan_build(h_n, h_offsets, count, 0, 0, p, L, host);
    if (rc != OK) return rc;
    if (L.total > PERM_SCRATCH_BYTES) {
        set_error("permanent batch needs " + std::to_string(L.total) + " bytes of plan scratch, above the " +
                  std::to_string(PERM_SCRATCH_BYTES) + " of pcvl_perm_batch_noalloc (split the batch or use "
                  "pcvl_perm_batch)");
        return ERR_RANGE;
    }
    if (count == 0) return OK;
    PCVL_ARG_CHECK(d_mats && d_out, "null device buffer");
    void *base = nullptr;
    rc = scratch_ptr(&base);
    if (rc != OK) return rc;
    perm_plan_bind(p, L, base);
    PCVL_HIP_CHECK(hipMemcpyAsync(base, host.data(), L.total, hipMemcpyHostToDevice, stream));
    rc = pcvl_perm_plan_execute(&plan, d_mats, d_out, stream);
    // the host image must outlive its copy, and results are ready on return
    hipError_t e = hipStreamSynchronize(stream);
    if (rc != OK) return rc;
    if (e != hipSuccess) { set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e)); return ERR_HIP; }
    return OK;
}

extern "C" int pcvl_perm_batch_noalloc(const double *d_mats, const int32_t *h_n, const int64_t *h_offsets,
                                       int64_t count, double *d_out, hipStream_t stream)
{
    try {
        return perm_batch_noalloc_impl(d_mats, h_n, h_offsets, count, d_out, stream);
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed while building the permanent plan");
        return ERR_ALLOC;
    } catch (...) {
        set_error("unexpected exception while building the permanent plan");
        return ERR_ALLOC;
    }
}

extern "C" size_t pcvl_perm_noalloc_max_bytes(void) { return PERM_SCRATCH_BYTES; }
