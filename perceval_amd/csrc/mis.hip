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

#include <algorithm>

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

// v of lane `lane` (wave-uniform) as a scalar: two v_readlane into SGPRs,
// no LDS round trip as a shuffle by an arbitrary lane would take.
__device__ __forceinline__ uint64_t mis_readlane_u64(uint64_t v, int lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double mis_readlane(double v, int lane)
{
    return __longlong_as_double((long long)mis_readlane_u64((uint64_t)__double_as_longlong(v), lane));
}

// ---- accept/reject scan ----------------------------------------------------
//
// The chain takes step t iff lw_t - cur >= 0 or log u_t < lw_t - cur, where
// cur is the log weight of the state it holds, and after taking step t it
// holds cur = lw_t.  Cut the steps into groups of 64.  Which steps of a group
// the chain takes after holding step s of that group depends on s alone:
// next(s) = the first later step of the group taken from cur = lw_s, and the
// chain then takes next(s), next(next(s)), ... to the group's end.  So the
// whole set, path(s), is computed for every step in parallel, and the
// sequential scan does one ballot per group: the first step taken from the
// state held on entry, a, and then path(a).  The decisions are the same
// floating-point comparisons as a step-by-step scan, so the states are too.

// One wave per group: lu[t] = log u_t (uniform of step t: stream
// (seed, first + t), counter 0) and path[t] (bit u: step 64 g + u).
__global__ void __launch_bounds__(64) mis_chain_path_kernel(const double *__restrict__ logw, int64_t B, uint64_t seed,
                                                            uint64_t first, double *__restrict__ lu_out,
                                                            uint64_t *__restrict__ path_out)
{
    const int64_t g = blockIdx.x;
    const int s = threadIdx.x;
    const int64_t t = g * 64 + s;
    const bool valid = t < B;
    const double lw = valid ? logw[t] : 0.0;
    const double lu = valid ? log(uniform01(seed, first + (uint64_t)t, 0)) : 0.0;
    const int nv = (int)(B - g * 64 < 64 ? B - g * 64 : 64);   // steps of this group
    int nxt = 64;                                              // none
    for (int u = 1; u < nv; u++) {
        const double lwu = mis_readlane(lw, u), luu = mis_readlane(lu, u);
        const double d = lwu - lw;
        if (nxt == 64 && u > s && (d >= 0.0 || luu < d)) nxt = u;
    }
    // pointer doubling: after r rounds pm = {s, next(s), ..., next^(2^r - 1)(s)}
    // and j = next^(2^r)(s); 6 rounds cover the 64 steps of a group.
    uint64_t pm = 1ull << s;
    int j = nxt;
    for (int r = 0; r < 6; r++) {
        const int src = j < 64 ? j : s;
        const uint32_t plo = (uint32_t)__shfl((int)(uint32_t)pm, src, 64);
        const uint32_t phi = (uint32_t)__shfl((int)(uint32_t)(pm >> 32), src, 64);
        const int jj = __shfl(j, src, 64);
        if (j < 64) {
            pm |= ((uint64_t)phi << 32) | plo;
            j = jj;
        }
    }
    if (valid) {
        lu_out[t] = lu;
        path_out[t] = pm;
    }
}

// The sequential scan, one workgroup.  Wave 0 walks the groups; the other
// waves stage the next block of log weights, log uniforms and paths in LDS
// (double buffered) meanwhile.  state[t] = index of the proposal the chain
// holds after step t, or -1 for the initial state; *accepted = moves taken.
constexpr int MIS_CHAIN_WAVES = 8;
constexpr int MIS_CHAIN_BLOCK = 1024;   // steps per LDS buffer (a multiple of 64)

__device__ __forceinline__ void mis_chain_stage(const double *__restrict__ logw, const double *__restrict__ lu_g,
                                                const uint64_t *__restrict__ path_g, int64_t B, int64_t blk,
                                                double *lw, double *lu, uint64_t *pm, int id, int nth)
{
    for (int i = id; i < MIS_CHAIN_BLOCK; i += nth) {
        const int64_t t = blk * MIS_CHAIN_BLOCK + i;
        const bool valid = t < B;
        lw[i] = valid ? logw[t] : 0.0;
        lu[i] = valid ? lu_g[t] : 0.0;
        pm[i] = valid ? path_g[t] : 0ull;
    }
}

__global__ void __launch_bounds__(64 * MIS_CHAIN_WAVES) mis_chain_kernel(const double *__restrict__ logw,
                                                                         const double *__restrict__ lu_g,
                                                                         const uint64_t *__restrict__ path_g, int64_t B,
                                                                         double logw0, int64_t *__restrict__ state,
                                                                         int64_t *__restrict__ accepted)
{
    __shared__ double s_lw[2][MIS_CHAIN_BLOCK], s_lu[2][MIS_CHAIN_BLOCK];
    __shared__ uint64_t s_pm[2][MIS_CHAIN_BLOCK];
    if (blockIdx.x != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t nblk = (B + MIS_CHAIN_BLOCK - 1) / MIS_CHAIN_BLOCK;
    mis_chain_stage(logw, lu_g, path_g, B, 0, s_lw[0], s_lu[0], s_pm[0], tid, 64 * MIS_CHAIN_WAVES);
    __syncthreads();
    double cur = logw0;
    int64_t idx = -1, acc = 0;
    for (int64_t k = 0; k < nblk; k++) {
        const int buf = (int)(k & 1);
        if (wave != 0) {
            if (k + 1 < nblk)
                mis_chain_stage(logw, lu_g, path_g, B, k + 1, s_lw[buf ^ 1], s_lu[buf ^ 1], s_pm[buf ^ 1], tid - 64,
                                64 * (MIS_CHAIN_WAVES - 1));
        } else {
            const int64_t b0 = k * MIS_CHAIN_BLOCK;
            const int bn = (int)(B - b0 < MIS_CHAIN_BLOCK ? B - b0 : MIS_CHAIN_BLOCK);
            for (int g0 = 0; g0 < bn; g0 += 64) {
                const int i = g0 + lane;
                const bool valid = i < bn;
                const double lw = s_lw[buf][i], lu = s_lu[buf][i];
                const uint64_t pm = s_pm[buf][i];
                const double d = lw - cur;
                const uint64_t first_taken = __ballot(valid && (d >= 0.0 || lu < d));
                uint64_t taken = 0;   // steps of this group that the chain takes
                if (first_taken) {
                    taken = mis_readlane_u64(pm, __builtin_ctzll(first_taken));
                    cur = mis_readlane(lw, 63 - __builtin_clzll(taken));
                }
                const int64_t t0 = b0 + g0;
                const uint64_t upto = taken & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
                const int64_t mine = upto ? t0 + (63 - __builtin_clzll(upto)) : idx;
                if (valid) state[t0 + lane] = mine;
                if (taken) idx = t0 + (63 - __builtin_clzll(taken));
                acc += __builtin_popcountll(taken);
            }
        }
        __syncthreads();   // block k's buffer is free, block k+1's is staged
    }
    if (tid == 0) *accepted = acc;
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


// ------------------------------------------------------------------------
// Permanents of proposals with repeated rows (multiset Glynn).
//
// A proposal s (sorted modes) selects rows U[s_r, cols]; output photons in the
// same mode give identical rows.  In Glynn's sum the k copies of a row enter
// the column sums only through their sign total k - 2t (t = copies with sign
// -1), with weight C(k, t) (-1)^t, so a group of k copies needs k+1 terms
// instead of 2^k.  Rows are ordered [fixed row | u unique rows | G groups];
// the unique rows are walked in Gray-code order, the group digits form an
// outer mixed-radix index o with weight W_o = prod_g C(k_g, t_g) (-1)^{t_g}:
//     perm = 2 sum_o W_o sum_{t < 2^u} (-1)^t prod_c h_c(o, t),
//     h_c = (1/2) (a_fc + sum_unique delta_i a_ic + sum_g (k_g - 2 t_g) a_gc).
// (Half column sums let a flip add +-a directly; prod c = 2^n prod h.)
// One wavefront per proposal; lanes take (o, inner chunk) units.
constexpr int MISP_MAX_N = 28;

struct MispShared {
    cd A[MISP_MAX_N * MISP_MAX_N];   // rows of the fixed / unique / group modes
    int kg[MISP_MAX_N];              // multiplicity of each group
    int rmode[MISP_MAX_N];           // mode of each stored row
    int gmode[MISP_MAX_N], gcnt[MISP_MAX_N];   // distinct modes of the sequence and their counts
    int nu, ng;                      // unique rows (besides the fixed one), groups
};

// prod_c h[c]; four independent chains from N >= 8 (ILP for FP64 latency).
template <int N>
__device__ __forceinline__ cd misp_prod(const cd *h)
{
    if constexpr (N < 8) {
        cd p = h[0];
#pragma unroll
        for (int c = 1; c < N; c++) p = cmul(p, h[c]);
        return p;
    } else {
        cd p0 = h[0], p1 = h[1], p2 = h[2], p3 = h[3];
#pragma unroll
        for (int c = 4; c < N; c++) {
            if ((c & 3) == 0) p0 = cmul(p0, h[c]);
            else if ((c & 3) == 1) p1 = cmul(p1, h[c]);
            else if ((c & 3) == 2) p2 = cmul(p2, h[c]);
            else p3 = cmul(p3, h[c]);
        }
        return cmul(cmul(p0, p1), cmul(p2, p3));
    }
}

template <int N>
__device__ __forceinline__ void misp_prod_acc(cd &acc, const cd *h, bool neg)
{
    const cd p = misp_prod<N>(h);
    if (neg) { acc.r -= p.r; acc.i -= p.i; } else { acc.r += p.r; acc.i += p.i; }
}

// Returns v, made to look dependent on dep, so LDS reads at offset v stay
// inside the loop that dep counts instead of being hoisted into VGPRs.
__device__ __forceinline__ int misp_loop_variant(int v, uint32_t dep)
{
    asm("" : "+s"(v) : "s"(dep));
    return v;
}

// Keeps the next update of h behind the term that read it (no hoisting of
// column-sum generations, which would multiply the live VGPRs).
template <int N>
__device__ __forceinline__ void misp_pin(cd *h, const cd &part)
{
#pragma unroll
    for (int cc = 0; cc < N; cc++) asm volatile("" : "+v"(h[cc].r), "+v"(h[cc].i) : "v"(part.r), "v"(part.i));
}

// ---- multiset permanents of proposals, as a balanced work list -----------
//
// Proposal p (a sorted mode sequence) is planned by one thread: its rows are
// a fixed row (delta_0 = +1), nu unique rows (Gray-coded) and ng groups of
// k_g equal rows (outer mixed-radix digits t_g in 0..k_g, weight
// C(k_g, t_g)(-1)^t_g, column-sum offset (k_g - 2 t_g)/2 times the row).  Its
// O * 2^nu terms are cut into units of 2^c Gray steps (units = O 2^(nu-c),
// c <= 10, chosen per proposal so that the wave passes fill their lanes),
// and the units into at most 64 slices.  Waves pull slices from one atomic cursor, so proposals of very
// different cost (a collision-free proposal has up to 4000 times the terms
// of a bunched one) balance over the device; each slice writes one partial,
// and a finalize pass sums a proposal's partials in slice order
// (deterministic).
constexpr int MISP_REC = 8 + 2 * MISP_MAX_N;   // ints per proposal record
constexpr int MISP_MAX_SLICES = 64;
constexpr int MISP_MAX_C = 10;

// record: [0] nu [1] ng [2] c [3] slices [4..5] units (int64) [6] nrow
//         [8 ..] row modes (fixed, unique, groups)  [8 + MAX_N ..] group sizes
__global__ void mis_plan_kernel(const int32_t *__restrict__ seq, int64_t B, int n, int m, int32_t *__restrict__ rec,
                                int32_t *__restrict__ nslices, unsigned long long *__restrict__ cursor)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) *cursor = 0ull;
    if (b >= B) return;
    int32_t *R = rec + b * MISP_REC;
    const int32_t *sq = seq + b * n;
    // The sorted sequence's runs (of modes clamped to [0, m)) are its distinct
    // modes with their counts; they are walked in place, with no per-thread
    // arrays.  The fixed row is the first run of count 1 if there is one, else
    // one copy of the first run.  Products of (count + 1) stay below 2^28
    // (at most 2^(n-1) terms), so 32-bit counts suffice.
    auto mode_at = [&](int r) {
        const int j = sq[r];
        return (j >= 0 && j < m) ? j : 0;
    };
    auto run_end = [&](int r, int j) {
        int e = r + 1;
        while (e < n && mode_at(e) == j) e++;
        return e;
    };
    int fixed_mode = -1;
    for (int r = 0; r < n;) {
        const int j = mode_at(r), e = run_end(r, j);
        if (e - r == 1) { fixed_mode = j; break; }
        r = e;
    }
    int nrow = 0, nu = 0, ng = 0;
    uint32_t O = 1;
    if (fixed_mode >= 0) {
        R[8 + nrow++] = fixed_mode;
        bool skipped = false;   // the fixed row's own run
        for (int r = 0; r < n;) {
            const int j = mode_at(r), e = run_end(r, j);
            if (e - r == 1) {
                if (skipped) { R[8 + nrow++] = j; nu++; }
                skipped = true;
            }
            r = e;
        }
        for (int r = 0; r < n;) {
            const int j = mode_at(r), e = run_end(r, j);
            if (e - r >= 2) {
                R[8 + nrow++] = j;
                R[8 + MISP_MAX_N + ng++] = e - r;
                O *= (uint32_t)(e - r + 1);
            }
            r = e;
        }
    } else {
        for (int r = 0; r < n;) {
            const int j = mode_at(r), e = run_end(r, j);
            const int k = r == 0 ? e - r - 1 : e - r;   // one copy of the first run is the fixed row
            if (r == 0) R[8 + nrow++] = j;
            if (k >= 1) {
                R[8 + nrow++] = j;
                R[8 + MISP_MAX_N + ng++] = k;
                O *= (uint32_t)(k + 1);
            }
            r = e;
        }
    }
    // Unit length 2^c: a wave runs its slice's units 64 at a time, so a unit
    // count that is not a multiple of 64 idles lanes in the last pass, while
    // short units repeat the column-sum set-up more often.  Take the c with
    // the least estimated wave time, in per-lane FP64 operations: passes x
    // (2^c Gray steps of 6n+2 operations + the set-up) + about 1000 per slice
    // (record and rows into LDS, search, reduction).
    const int cmax = nu < MISP_MAX_C ? nu : MISP_MAX_C;
    const double step_ops = 6.0 * n + 2.0;
    const double setup_ops = 2.0 * n * (2 + nu + ng) + 30.0 * ng;
    int c = cmax;
    double best = -1.0;
    for (int cc = cmax; cc >= 0 && cc >= cmax - 5; cc--) {
        const uint32_t un = O << (nu - cc);
        const uint32_t sl_c = (un + 63) / 64 < (uint32_t)MISP_MAX_SLICES ? (un + 63) / 64 : (uint32_t)MISP_MAX_SLICES;
        const uint32_t per = (un + sl_c - 1) / sl_c;
        const uint32_t full = un / per, rest = un - full * per;
        const uint32_t passes = full * ((per + 63) / 64) + (rest + 63) / 64;
        const double est = (double)passes * ((double)(1u << cc) * step_ops + setup_ops) + 1000.0 * (double)sl_c;
        if (best < 0.0 || est < best) { best = est; c = cc; }
    }
    const uint32_t units = O << (nu - c);
    const uint32_t sl = (units + 63) / 64;
    const int slices = (int)(sl < (uint32_t)MISP_MAX_SLICES ? sl : (uint32_t)MISP_MAX_SLICES);
    R[0] = nu; R[1] = ng; R[2] = c; R[3] = slices;
    R[4] = (int32_t)units; R[5] = 0; R[6] = nrow; R[7] = 0;
    nslices[b] = slices;
}

// Exclusive prefix of the slice counts, one block: slice0[b], total at
// slice0[B].  4096 counts per pass, four consecutive ones per thread (their
// loads in flight together); a shuffle scan of the thread sums per wave and
// the 16 wave totals through LDS.
__global__ void __launch_bounds__(1024) mis_scan_kernel(const int32_t *__restrict__ nslices, int64_t B,
                                                        int64_t *__restrict__ slice0)
{
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int64_t carry = 0;
    for (int64_t base = 0; base < B; base += 4096) {
        const int64_t i0 = base + 4 * (int64_t)t;
        int v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = i0 + q < B ? nslices[i0 + q] : 0;   // at most 64 each
        const int own = v[0] + v[1] + v[2] + v[3];
        int x = own;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int w = wsum[k];
            if (k < wv) before += w;
            total += w;
        }
        int64_t run = carry + before + x - own;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (i0 + q < B) slice0[i0 + q] = run;
            run += v[q];
        }
        carry += total;
        __syncthreads();   // wsum is rewritten by the next pass
    }
    if (t == 0) slice0[B] = carry;
}

// Owner of every slice, packed as proposal * 64 + slice (slices per
// proposal <= 64), so a wave finds its slice's proposal with one load.
__global__ void mis_owner_kernel(const int64_t *__restrict__ slice0, int64_t B, uint64_t *__restrict__ owner)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t g0 = slice0[b], g1 = slice0[b + 1];
    for (int64_t g = g0; g < g1; g++) owner[g] = ((uint64_t)b << 6) | (uint64_t)(g - g0);
}

template <int N>
__global__ void __launch_bounds__(64) mis_slice_kernel(const cd *__restrict__ U, int m, const int32_t *__restrict__ cols,
                                                       const int32_t *__restrict__ rec, const int64_t *__restrict__ slice0,
                                                       const uint64_t *__restrict__ owner, int64_t B,
                                                       unsigned long long *__restrict__ cursor, cd *__restrict__ partial)
{
    __shared__ MispShared sh;
    __shared__ int32_t srec[MISP_REC];
    const int lane = threadIdx.x;
    const int64_t total = slice0[B];
    // a wave takes at most `total` slices, so the loop is bounded whatever the cursor holds
    for (int64_t iter = 0; iter < total; iter++) {
        unsigned long long gk = 0;
        if (lane == 0) gk = atomicAdd(cursor, 1ull);
        gk = __shfl(gk, 0, 64);
        const int64_t g = (int64_t)gk;
        if (g >= total) break;
        // proposal p and slice s of global slice g, packed as p * 64 + s
        const uint64_t own = owner[g];
        const int64_t pidx = (int64_t)(own >> 6);
        const int s = (int)(own & 63u);
        for (int e = lane; e < MISP_REC; e += 64) srec[e] = rec[pidx * MISP_REC + e];
        __syncthreads();
        const int nu = srec[0], ng = srec[1], c = srec[2], slices = srec[3], nrow = srec[6];
        const uint64_t units = (uint64_t)(uint32_t)srec[4] | ((uint64_t)(uint32_t)srec[5] << 32);
        for (int e = lane; e < nrow * N; e += 64) {
            const int r = e / N, cc = e % N;
            sh.A[e] = U[(size_t)srec[8 + r] * m + cols[cc]];
        }
        if (lane < ng) sh.kg[lane] = srec[8 + MISP_MAX_N + lane];
        __syncthreads();
        const uint64_t nchunks = 1ull << (nu - c);
        const uint64_t per = (units + (uint64_t)slices - 1) / (uint64_t)slices;
        const uint64_t u0 = (uint64_t)s * per;
        const uint64_t u1 = (u0 + per < units) ? u0 + per : units;
        cd acc = {0.0, 0.0};
        for (uint64_t unit = u0 + (uint64_t)lane; unit < u1; unit += 64) {
            uint64_t o = unit / nchunks;
            const uint64_t chunk = unit % nchunks;
            const uint64_t base = chunk << c;
            const uint64_t g0 = base ^ (base >> 1);
            cd h[N];
#pragma unroll
            for (int cc = 0; cc < N; cc++) h[cc] = {0.5 * sh.A[cc].r, 0.5 * sh.A[cc].i};
            for (int i = 1; i <= nu; i++) {
                const double d = ((g0 >> (i - 1)) & 1) ? -0.5 : 0.5;
#pragma unroll
                for (int cc = 0; cc < N; cc++) {
                    h[cc].r = fma(d, sh.A[i * N + cc].r, h[cc].r);
                    h[cc].i = fma(d, sh.A[i * N + cc].i, h[cc].i);
                }
            }
            // group digits of o (mixed radix k_g + 1) and the weight W_o
            double W = 1.0;
            for (int gg = 0; gg < ng; gg++) {
                const int k = sh.kg[gg];
                const int t = (int)(o % (uint64_t)(k + 1));
                o /= (uint64_t)(k + 1);
                double binom = 1.0;
                for (int i = 1; i <= t; i++) binom = binom * (double)(k - t + i) / (double)i;
                W *= (t & 1) ? -binom : binom;
                const double f = 0.5 * (double)(k - 2 * t);
#pragma unroll
                for (int cc = 0; cc < N; cc++) {
                    h[cc].r = fma(f, sh.A[(1 + nu + gg) * N + cc].r, h[cc].r);
                    h[cc].i = fma(f, sh.A[(1 + nu + gg) * N + cc].i, h[cc].i);
                }
            }
            cd part = {0.0, 0.0};
            const uint64_t steps = 1ull << c;
            if (c >= 2) {
                // Four steps per iteration; base is a multiple of 4, so row 1
                // (first unique row) flips to delta -1 after the first term
                // and back after the third.  Row 1 sits in VGPRs up to order
                // 22 and is read from LDS above (VGPR budget of two waves).
                constexpr bool kR1Regs = N <= 22;
                cd r1[kR1Regs ? N : 1];
                if constexpr (kR1Regs) {
#pragma unroll
                    for (int cc = 0; cc < N; cc++) r1[cc] = sh.A[N + cc];
                }
                for (uint64_t u = 0; u < steps; u += 4) {
                    const cd *row1 = kR1Regs ? r1 : sh.A + misp_loop_variant(N, (uint32_t)u);
                    const cd *row2 = sh.A + misp_loop_variant(2 * N, (uint32_t)u);
                    misp_prod_acc<N>(part, h, false);             // t = base+u
                    misp_pin<N>(h, part);
#pragma unroll
                    for (int cc = 0; cc < N; cc++) { h[cc].r -= row1[cc].r; h[cc].i -= row1[cc].i; }
                    misp_prod_acc<N>(part, h, true);              // t+1
                    misp_pin<N>(h, part);
                    {
                        const uint64_t t = base + u + 2;          // row 2
                        const double sg = (((t >> 1) ^ (t >> 2)) & 1) ? -1.0 : 1.0;
#pragma unroll
                        for (int cc = 0; cc < N; cc++) {
                            h[cc].r = fma(sg, row2[cc].r, h[cc].r);
                            h[cc].i = fma(sg, row2[cc].i, h[cc].i);
                        }
                    }
                    misp_prod_acc<N>(part, h, false);             // t+2
                    misp_pin<N>(h, part);
#pragma unroll
                    for (int cc = 0; cc < N; cc++) { h[cc].r += row1[cc].r; h[cc].i += row1[cc].i; }
                    misp_prod_acc<N>(part, h, true);              // t+3
                    misp_pin<N>(h, part);
                    if (u + 4 < steps) {                          // row 1 + ctz(u+4) <= c <= nu
                        const uint64_t t = base + u + 4;
                        const int bb = __builtin_ctzll(u + 4);
                        const double sg = (((t >> bb) ^ (t >> (bb + 1))) & 1) ? -1.0 : 1.0;
                        const cd *row = sh.A + (bb + 1) * N;
#pragma unroll
                        for (int cc = 0; cc < N; cc++) {
                            h[cc].r = fma(sg, row[cc].r, h[cc].r);
                            h[cc].i = fma(sg, row[cc].i, h[cc].i);
                        }
                    }
                }
            } else {
                for (uint64_t u = 0; u < steps; u++) {
                    const uint64_t t = base + u;
                    if (u) {
                        const int bb = __builtin_ctzll(t);
                        const double sg = (((t >> bb) ^ (t >> (bb + 1))) & 1) ? -1.0 : 1.0;
                        const cd *row = sh.A + (bb + 1) * N;
#pragma unroll
                        for (int cc = 0; cc < N; cc++) {
                            h[cc].r = fma(sg, row[cc].r, h[cc].r);
                            h[cc].i = fma(sg, row[cc].i, h[cc].i);
                        }
                    }
                    misp_prod_acc<N>(part, h, (t & 1) != 0);
                    misp_pin<N>(h, part);
                }
            }
            acc.r = fma(W, part.r, acc.r);
            acc.i = fma(W, part.i, acc.i);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            acc.r += __shfl_xor(acc.r, off, 64);
            acc.i += __shfl_xor(acc.i, off, 64);
        }
        if (lane == 0) partial[g] = acc;
        __syncthreads();   // srec and sh are reused by the next slice
    }
}

// out[p] = 2 * sum of proposal p's partials, in slice order.
__global__ void mis_finalize_kernel(const cd *__restrict__ partial, const int64_t *__restrict__ slice0, int64_t B,
                                    cd *__restrict__ out)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    cd acc = {0.0, 0.0};
    for (int64_t g = slice0[b]; g < slice0[b + 1]; g++) { acc.r += partial[g].r; acc.i += partial[g].i; }
    out[b] = {2.0 * acc.r, 2.0 * acc.i};
}

template <int N>
static int launch_mis_slices(int n, const cd *U, int m, const int32_t *cols, const int32_t *rec, const int64_t *slice0,
                             const uint64_t *owner, int64_t B, unsigned long long *cursor, cd *partial, hipStream_t st)
{
    if constexpr (N > MISP_MAX_N) {
        set_error("photon number above 28");
        return ERR_ARG;
    } else {
        if (n != N) return launch_mis_slices<N + 1>(n, U, m, cols, rec, slice0, owner, B, cursor, partial, st);
        int dev = 0, cus = 0, per_cu = 0;
        PCVL_HIP_CHECK(hipGetDevice(&dev));
        PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&mis_slice_kernel<N>), 64, 0));
        if (per_cu < 1) per_cu = 1;
        const int grid = (int)std::min<int64_t>(B * MISP_MAX_SLICES, (int64_t)cus * per_cu);
        hipLaunchKernelGGL((mis_slice_kernel<N>), dim3(grid), dim3(64), 0, st, U, m, cols, rec, slice0, owner, B,
                           cursor, partial);
        PCVL_HIP_CHECK(hipGetLastError());
        return OK;
    }
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_mis_permanents(const double *d_U, int m, const int32_t *d_cols, int n, const int32_t *d_seq,
                                   int64_t B, double *d_out, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && n >= 1 && n <= MISP_MAX_N && B >= 0, "bad sizes (1 <= n <= 28)");
    if (B == 0) return OK;
    PCVL_ARG_CHECK(d_U && d_cols && d_seq && d_out, "null device buffer");
    // stream-ordered workspace: records | slice counts | slice offsets | cursor | partials | slice owners
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t b_rec = al(sizeof(int32_t) * (size_t)B * MISP_REC);
    const size_t b_ns = al(sizeof(int32_t) * (size_t)B);
    const size_t b_s0 = al(sizeof(int64_t) * ((size_t)B + 1));
    const size_t b_cur = al(sizeof(unsigned long long));
    const size_t b_part = al(sizeof(cd) * (size_t)B * MISP_MAX_SLICES);
    const size_t b_own = al(sizeof(uint64_t) * (size_t)B * MISP_MAX_SLICES);
    char *blk = nullptr;
    PCVL_HIP_CHECK(ws_alloc(reinterpret_cast<void **>(&blk), b_rec + b_ns + b_s0 + b_cur + b_part + b_own, stream));
    int32_t *rec = reinterpret_cast<int32_t *>(blk);
    int32_t *ns = reinterpret_cast<int32_t *>(blk + b_rec);
    int64_t *s0 = reinterpret_cast<int64_t *>(blk + b_rec + b_ns);
    unsigned long long *cur = reinterpret_cast<unsigned long long *>(blk + b_rec + b_ns + b_s0);
    cd *part = reinterpret_cast<cd *>(blk + b_rec + b_ns + b_s0 + b_cur);
    uint64_t *own = reinterpret_cast<uint64_t *>(blk + b_rec + b_ns + b_s0 + b_cur + b_part);
    int rc = OK;
    hipLaunchKernelGGL(mis_plan_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream, d_seq, B, n, m, rec, ns,
                       cur);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mis_scan_kernel, dim3(1), dim3(1024), 0, stream, ns, B, s0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mis_owner_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream, s0, B, own);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        set_error(std::string("mis plan: ") + hipGetErrorString(e));
        rc = ERR_HIP;
    } else {
        rc = launch_mis_slices<1>(n, reinterpret_cast<const cd *>(d_U), m, d_cols, rec, s0, own, B, cur, part, stream);
    }
    if (rc == OK) {
        hipLaunchKernelGGL(mis_finalize_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, stream, part, s0, B,
                           reinterpret_cast<cd *>(d_out));
        e = hipGetLastError();
        if (e != hipSuccess) { set_error(std::string("mis finalize: ") + hipGetErrorString(e)); rc = ERR_HIP; }
    }
    hipError_t fe = hipFreeAsync(blk, stream);
    if (rc == OK && fe != hipSuccess) { set_error(std::string("hipFreeAsync(mis): ") + hipGetErrorString(fe)); rc = ERR_HIP; }
    return rc;
}

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
    PCVL_ARG_CHECK(B >= 0 && B <= ((int64_t)1 << 36), "bad size (0 <= B <= 2^36)");
    if (B == 0) return OK;
    PCVL_ARG_CHECK(d_logw && d_state && d_accepted, "null device buffer");
    // stream-ordered workspace: log uniforms | paths
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t b_lu = al(sizeof(double) * (size_t)B);
    const size_t b_pm = al(sizeof(uint64_t) * (size_t)B);
    char *blk = nullptr;
    PCVL_HIP_CHECK(ws_alloc(reinterpret_cast<void **>(&blk), b_lu + b_pm, stream));
    double *lu = reinterpret_cast<double *>(blk);
    uint64_t *pm = reinterpret_cast<uint64_t *>(blk + b_lu);
    int rc = OK;
    hipLaunchKernelGGL(mis_chain_path_kernel, dim3((unsigned)((B + 63) / 64)), dim3(64), 0, stream, d_logw, B, seed,
                       first, lu, pm);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mis_chain_kernel, dim3(1), dim3(64 * MIS_CHAIN_WAVES), 0, stream, d_logw, lu, pm, B, logw0,
                           d_state, d_accepted);
        e = hipGetLastError();
    }
    if (e != hipSuccess) { set_error(std::string("mis chain: ") + hipGetErrorString(e)); rc = ERR_HIP; }
    hipError_t fe = hipFreeAsync(blk, stream);
    if (rc == OK && fe != hipSuccess) { set_error(std::string("hipFreeAsync(chain): ") + hipGetErrorString(fe)); rc = ERR_HIP; }
    return rc;
}
