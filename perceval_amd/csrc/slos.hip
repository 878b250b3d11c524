// SLOS layered Fock-state amplitude recursion and Fock-space ranking.
//
// Replaces _Path.compute -> FSMap.compute_slos_layer and
// xq.all_prob_normalize_output (perceval/backends/_slos.py:88-102, 195-214),
// and FSArray iteration / FSArray.find (perceval/utils/states.py:255-298,
// _slos.py:190).
//
// Fock order.  The n-photon states of m modes are ranked in descending
// lexicographic order (pinned by tests/utils/test_statevector.py:430-438).
// With R_q = number of photons in modes q..m-1 (R_0 = n),
//     rank(s) = sum_{q=0}^{m-2} T_q(R_{q+1}),   T_q(x) = C(x + m-q-2, m-q-1),
// which is the combinatorial number system with K = m-1.  Unranking walks the
// levels once: R_{q+1} is the largest x <= R_q with T_q(x) <= remaining, and
// since x only decreases the walk costs m + n table reads in total.
//
// Layer recursion (gather form).  Child c (k photons) sums its parents c-e_j
// over occupied modes j:  coefs_k[c] = sum_j coefs_{k-1}[c - e_j] * U[j, mk].
// The parent rank is rank(c) - sum_{q<j} E_q(R_{q+1}) with
// E_q(x) = C(x + m-q-3, m-q-2), so one pass over the modes unranks the child
// and gathers all its parents.  The last layer can emit probabilities
// |c|^2 * prod s! / prod t! directly (fused epilogue).
#include "common.h"
#include "slos_block.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

namespace pcvl {

constexpr int SLOS_MAX_MODES = 256;
constexpr int SLOS_MAX_PHOTONS = 40;
constexpr int SLOS_THREADS = 256;

// Exact binomial with saturation at UINT64_MAX.
static uint64_t binom_u64(int64_t a, int64_t b)
{
    if (b < 0 || a < 0 || b > a) return 0;
    if (b > a - b) b = a - b;
    unsigned __int128 r = 1;
    for (int64_t i = 1; i <= b; i++) {
        r = r * (unsigned __int128)(a - b + i) / (unsigned __int128)i;
        if (r > (unsigned __int128)UINT64_MAX) return UINT64_MAX;
    }
    return (uint64_t)r;
}

// Rank tables for (m, nmax): T[q*(nmax+1)+x] and E[q*(nmax+1)+x] for
// q in [0, m), x in [0, nmax]; T/E of q = m-1 are zero (no term).
// Walk tables of the successor kernels, one array of 64-bit words: the packed
// table P[q][x] = T[q][x] | E[q][x] << 32 for x = 0..nmax+1, then the 32-bit
// cumulative columns CT[R][q] = sum_{j<q} T[j][R] and CE[R][q] (the same
// for E) for R = 0..nmax, q = 0..m+3, with CT = 0xffffffff for q >= m.
// Values above 2^32 are truncated; the kernels run only where every layer
// has fewer than 2^32 states, which bounds all of them.
static void build_packed(int m, int nmax, std::vector<uint64_t> &P)
{
    const int wp = nmax + 2, cw = m + 4;
    const size_t nc = (size_t)(nmax + 1) * cw;
    P.assign((size_t)m * wp + nc, 0);   // nc u64 words hold the 2 * nc u32 of CT and CE
    for (int q = 0; q + 1 < m; q++)
        for (int x = 0; x <= nmax + 1; x++) {
            const uint64_t t = binom_u64((int64_t)x + m - q - 2, m - q - 1);
            const uint64_t e = binom_u64((int64_t)x + m - q - 3, m - q - 2);
            P[(size_t)q * wp + x] = (t & 0xffffffffull) | (e << 32);
        }
    uint32_t *ct = reinterpret_cast<uint32_t *>(P.data() + (size_t)m * wp);
    uint32_t *ce = ct + nc;
    for (int R = 0; R <= nmax; R++) {
        uint64_t at = 0, ae = 0;
        for (int q = 0; q < cw; q++) {
            if (q < m) {
                ct[(size_t)R * cw + q] = (uint32_t)at;
                ce[(size_t)R * cw + q] = (uint32_t)ae;
                if (q + 1 < m) {
                    at += binom_u64((int64_t)R + m - q - 2, m - q - 1);
                    ae += binom_u64((int64_t)R + m - q - 3, m - q - 2);
                }
            } else {
                ct[(size_t)R * cw + q] = 0xffffffffu;
                ce[(size_t)R * cw + q] = 0;
            }
        }
    }
}

static void build_tables(int m, int nmax, std::vector<uint64_t> &T, std::vector<uint64_t> &E)
{
    const int w = nmax + 1;
    T.assign((size_t)m * w, 0);
    E.assign((size_t)m * w, 0);
    for (int q = 0; q + 1 < m; q++)
        for (int x = 0; x <= nmax; x++) {
            T[(size_t)q * w + x] = binom_u64((int64_t)x + m - q - 2, m - q - 1);
            E[(size_t)q * w + x] = binom_u64((int64_t)x + m - q - 3, m - q - 2);
        }
}

__device__ __forceinline__ double fact_d(int v)
{
    double f = 1.0;
    for (int i = 2; i <= v; i++) f *= (double)i;
    return f;
}

// Layer kernel: children [0, nchild) of a k-photon layer.  MODE 0 writes
// complex coefficients, MODE 1 writes probabilities.
template <int MODE>
__global__ void __launch_bounds__(SLOS_THREADS) slos_layer_kernel(
    const cd *__restrict__ parent, uint64_t nparent, void *__restrict__ out, const cd *__restrict__ U, int m, int mk,
    int k, uint64_t nchild, const uint64_t *__restrict__ gT, const uint64_t *__restrict__ gE, int w, double inv_prod_in)
{
    extern __shared__ uint64_t s_tab[];   // T (m*w) then E (m*w)
    uint64_t *sT = s_tab;
    uint64_t *sE = s_tab + (size_t)m * w;
    for (int e = threadIdx.x; e < m * w; e += blockDim.x) {
        sT[e] = gT[e];
        sE[e] = gE[e];
    }
    __syncthreads();

    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nchild; idx += stride) {
        uint64_t rem = idx;      // rank still to place
        uint64_t pe = 0;         // sum_{q' < q} E_q'
        int R = k;               // photons in modes q..m-1
        cd acc = {0.0, 0.0};
        double pfact = 1.0;
        for (int q = 0; q < m; q++) {
            int Rn;
            if (q + 1 < m) {
                const uint64_t *tq = sT + q * w;
                Rn = R;
                uint64_t tv = tq[Rn];
                while (Rn > 0 && tv > rem) { --Rn; tv = tq[Rn]; }   // T_q(0) = 0 <= rem ends the walk
                rem -= tv;
            } else {
                Rn = 0;
            }
            const int sq = R - Rn;
            const uint64_t pidx = idx - pe;
            if (sq > 0 && pidx < nparent) {   // always true for a consistent rank; guards the read
                const cd p = parent[pidx];
                const cd u = U[(size_t)q * m + mk];
                acc.r = fma(p.r, u.r, acc.r);
                acc.r = fma(-p.i, u.i, acc.r);
                acc.i = fma(p.r, u.i, acc.i);
                acc.i = fma(p.i, u.r, acc.i);
                if (MODE == 1) pfact *= fact_d(sq);
            }
            if (q + 1 < m) pe += sE[q * w + Rn];
            R = Rn;
        }
        if (MODE == 0) {
            reinterpret_cast<cd *>(out)[idx] = acc;
        } else {
            reinterpret_cast<double *>(out)[idx] = prob_from_coef(acc, pfact, inv_prod_in);
        }
    }
}


// Per-state kernel in two phases (parent layers below 2^26 states): the walk first
// records, per occupied mode, the parent rank and the mode in per-thread LDS
// slots, and a second loop gathers; its loads do not depend on one another,
// so the unrolled loop keeps several in flight instead of waiting on each
// inside the walk.  Summation order is the per-state kernel's (increasing
// mode), so the results agree with it to the last bit for non-zero sums.
constexpr int SLOS4_THREADS = 256;

template <int MODE, int G>
__global__ void __launch_bounds__(SLOS4_THREADS) slos_layer2p_kernel(
    const cd *__restrict__ parent, uint32_t nparent, void *__restrict__ out, const cd *__restrict__ U, int m, int mk,
    int k, uint32_t nchild, const uint64_t *__restrict__ gT, const uint64_t *__restrict__ gE, int w,
    double inv_prod_in)
{
    extern __shared__ uint64_t s_dyn2p[];
    // layout: u (m cd) | fact (w double) | T (m*w u32) | E (m*w u32) | slots (k*THREADS u32)
    cd *su = reinterpret_cast<cd *>(s_dyn2p);
    double *sf = reinterpret_cast<double *>(su + m);
    uint32_t *sT = reinterpret_cast<uint32_t *>(sf + w);
    uint32_t *sE = sT + (size_t)m * w;
    uint32_t *slot = sE + (size_t)m * w;
    for (int e = threadIdx.x; e < m * w; e += blockDim.x) {
        sT[e] = (uint32_t)gT[e];
        sE[e] = (uint32_t)gE[e];
    }
    for (int e = threadIdx.x; e < m; e += blockDim.x) su[e] = U[(size_t)e * m + mk];
    if (threadIdx.x == 0) {
        double f = 1.0;
        for (int v = 0; v < w; v++) {
            if (v > 1) f *= (double)v;
            sf[v] = f;
        }
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i64 = (uint64_t)blockIdx.x * blockDim.x + tid; i64 < nchild; i64 += stride) {
        const uint32_t idx = (uint32_t)i64;
        uint32_t rem = idx, pe = 0;
        int R = k, depth = 0;
        double pfact = 1.0;
        for (int q = 0; q < m && R > 0; q++) {
            int Rn = 0;
            if (q + 1 < m) {
                const uint32_t *tq = sT + q * w;
                Rn = R;
                uint32_t tv = tq[Rn];
                while (Rn > 0 && tv > rem) { --Rn; tv = tq[Rn]; }
                rem -= tv;
            }
            const int sq = R - Rn;
            if (sq > 0) {
                // parent rank (< 2^26) and mode (< 64) packed in one word
                slot[depth * SLOS4_THREADS + tid] = ((idx - pe) << 6) | (uint32_t)q;
                depth++;
                if (MODE == 1) pfact *= sf[sq];
            }
            if (q + 1 < m) pe += sE[q * w + Rn];
            R = Rn;
        }
        // Gather G parents at a time with unconditional loads (indices
        // clamped to the layer), so the G are in flight together; an
        // entry past the depth, or an index outside the parent layer (never
        // for a consistent rank), contributes an exact zero.
        cd acc = {0.0, 0.0};
        for (int d0 = 0; d0 < depth; d0 += G) {
            cd pv[G];
            int qv[G];
            bool ok[G];
#pragma unroll
            for (int t = 0; t < G; t++) {
                const int d = d0 + t;
                const uint32_t w32 = d < depth ? slot[d * SLOS4_THREADS + tid] : 0u;
                const uint32_t pidx = w32 >> 6;
                ok[t] = d < depth && pidx < nparent;
                pv[t] = parent[ok[t] ? pidx : 0u];
                qv[t] = (int)(w32 & 63u);
            }
#pragma unroll
            for (int t = 0; t < G; t++) {
                const cd u0 = su[qv[t]];
                const cd u = {ok[t] ? u0.r : 0.0, ok[t] ? u0.i : 0.0};
                acc.r = fma(pv[t].r, u.r, acc.r);
                acc.r = fma(-pv[t].i, u.i, acc.r);
                acc.i = fma(pv[t].r, u.i, acc.i);
                acc.i = fma(pv[t].i, u.r, acc.i);
            }
        }
        if (MODE == 0) {
            reinterpret_cast<cd *>(out)[idx] = acc;
        } else {
            reinterpret_cast<double *>(out)[idx] = prob_from_coef(acc, pfact, inv_prod_in);
        }
    }
}

// Successor kernel with a transposed gather.  Thread t of a block unranks
// child blk0 + t*RUN once and steps through the RUN - 1 children after it in
// O(1) each (in descending lexicographic order only the end of the
// occupied-mode list changes: the last occupied mode i below m-1 gives one
// photon to mode i+1, and the photons of mode m-1 join it there), writing
// every child's parent slots to LDS.  The gather then gives consecutive
// lanes consecutive children, so each load and store instruction covers a
// contiguous run of parents and outputs.  A slot holds
// (parent rank << 6) | mode, so the slots of the next child are the current
// ones plus 64 except at the end of the list.  Slots, gather and summation
// order are the two-phase kernel's, so the results are bit-identical to it.
// Slot columns are padded by one word per 32 so that both the per-thread
// writes (stride RUN) and the per-lane reads avoid LDS bank conflicts.
// The walk reads the packed table P[q][x] = T[q][x] | E[q][x] << 32.
// Parent layers carry one zero coefficient after their last state (index
// nparent): gathers past a child's depth read it, so they need no masking.
// Intermediate layers (sentinel != 0) write that zero for the next layer.
// PFI: the product of factorials of a child's occupations is kept as an
// exact 32-bit integer (k <= 12), else as a double.
constexpr int SLOS7_THREADS = 256;

__device__ __forceinline__ uint32_t succ_pad(uint32_t c) { return c + (c >> 5); }

// Tiles of a layer in XCD-contiguous order.  Workgroups are dispatched to
// the 8 XCDs round robin (block b on XCD b % 8), each XCD with its own L2, so
// XCD x is given the contiguous tile range [x*T8, (x+1)*T8): neighbouring
// tiles gather many of the same parent lines, and this keeps them in one L2
// (at m=24, n=10 it cuts the last layer's L2 misses from 6.1 to 4.7 GB; the
// time is unchanged, the kernel being bound by instruction issue).
// gridDim.x is a multiple of 8; any tile order gives the same results.
struct XcdTiles {
    uint32_t first, end, step;
};
__device__ __forceinline__ XcdTiles xcd_tiles(uint32_t ntiles)
{
    const uint32_t x = blockIdx.x & 7u, j = blockIdx.x >> 3, t8 = (ntiles + 7u) / 8u;
    const uint32_t lo = min(x * t8, ntiles), hi = min(lo + t8, ntiles);
    return {lo + j, hi, gridDim.x >> 3};
}

// Unranks `first` in the k-photon layer and calls emit(q, s_q, pe) for each
// occupied mode q in increasing order, pe being the rank offset of modes
// below q (parent(c - e_q) = first - pe).  Empty modes are skipped four at
// a time: with CT[R][q] = sum_{j<q} T[j][R] (CE the same for E; entries
// q >= m hold 0xffffffff), the walk at mode q with R photons left and rank
// rem in its block next stops at the first q' >= q with
// CT[R][q'+1] > rem + CT[R][q], and CT is increasing in q, so one compare
// per table entry counts the empty modes.  Equivalent to the mode-by-mode
// walk T[q][Rn] <= rem < T[q][Rn+1].
template <typename Emit>
__device__ __forceinline__ void skip_walk(uint32_t first, int m, int k, const uint64_t *__restrict__ sP, int wp,
                                          const uint32_t *__restrict__ sCT, const uint32_t *__restrict__ sCE, Emit emit)
{
    const int cw = m + 4;
    uint32_t rem = first, pe = 0;
    int R = k, q = 0;
    while (R > 0) {
        const uint32_t *ct = sCT + R * cw;
        const uint32_t X = rem + ct[q];
        int qn = q;
        for (;;) {
            const int c = (int)(ct[qn + 1] <= X) + (int)(ct[qn + 2] <= X) + (int)(ct[qn + 3] <= X) +
                          (int)(ct[qn + 4] <= X);
            qn += c;
            if (c < 4) break;
        }
        rem = X - ct[qn];
        pe += sCE[R * cw + qn] - sCE[R * cw + q];
        q = qn;
        int Rn = 0;
        uint64_t e = 0;
        if (q + 1 < m) {   // mode q holds at least one photon
            const uint64_t *tq = sP + q * wp;
            Rn = R - 1;
            e = tq[Rn];
            while (Rn > 0 && (uint32_t)e > rem) { --Rn; e = tq[Rn]; }
            rem -= (uint32_t)e;
        }
        emit(q, R - Rn, pe);
        pe += (uint32_t)(e >> 32);
        R = Rn;
        q++;
    }
}

// Successor kernel, compact form.  As slos_layer_succ_kernel, but a thread
// stores the full slot list only for its first child (slot0); child s of
// the run is the first child's first keep_s entries (parent ranks + s) plus
// an own tail of at most s entries, since each successor step rewrites only
// the end of the list.  That cuts the LDS per child to a half or less, so
// runs of several children fit with 5-6 blocks per CU.  Results are
// bit-identical to the two-phase kernel.
template <int MODE, int RUN, bool PFI, int G>
__global__ void __launch_bounds__(SLOS7_THREADS) slos_layer_succ2_kernel(
    const cd *__restrict__ parent, uint32_t nparent, void *__restrict__ out, const cd *__restrict__ U, int m, int mk,
    int k, uint32_t nchild, const uint64_t *__restrict__ gP, int wp, double inv_prod_in, int sentinel)
{
    constexpr uint32_t TILE = (uint32_t)RUN * SLOS7_THREADS;
    constexpr int OWN_ROWS = RUN * (RUN - 1) / 2;
    using pf_t = typename std::conditional<PFI, uint32_t, double>::type;
    extern __shared__ uint64_t s_dyn8[];
    // layout: P (m*wp u64) | u (m cd) | fact (wp double) | pf (RUN*THREADS pf_t, probabilities only) |
    //         slot0 (k*THREADS u32) | own (OWN_ROWS*THREADS u32) | cnt ((k+1)*THREADS u8) |
    //         keep (RUN*THREADS u8) | dep (RUN*THREADS u8)
    uint64_t *sP = s_dyn8;
    uint32_t *sCT = reinterpret_cast<uint32_t *>(sP + (size_t)m * wp);
    uint32_t *sCE = sCT + (size_t)(wp - 1) * (m + 4);
    cd *su = reinterpret_cast<cd *>(sCE + (size_t)(wp - 1) * (m + 4));
    double *sf = reinterpret_cast<double *>(su + m);
    pf_t *spf = reinterpret_cast<pf_t *>(sf + wp);
    uint32_t *slot0 = reinterpret_cast<uint32_t *>(spf + (MODE == 1 ? TILE : 0));
    uint32_t *own = slot0 + (size_t)k * SLOS7_THREADS;
    uint8_t *cnt = reinterpret_cast<uint8_t *>(own + (size_t)OWN_ROWS * SLOS7_THREADS);
    uint8_t *skeep = cnt + (size_t)(k + 1) * SLOS7_THREADS;
    uint8_t *sdep = skeep + TILE;
    for (int e = threadIdx.x; e < m * wp; e += blockDim.x) sP[e] = gP[e];
    {
        const uint32_t *gC = reinterpret_cast<const uint32_t *>(gP + (size_t)m * wp);
        for (int e = threadIdx.x; e < 2 * (wp - 1) * (m + 4); e += blockDim.x) sCT[e] = gC[e];
    }
    for (int e = threadIdx.x; e < m; e += blockDim.x) su[e] = U[(size_t)e * m + mk];
    if (threadIdx.x == 0) {
        double f = 1.0;
        for (int v = 0; v < wp; v++) {
            if (v > 1) f *= (double)v;
            sf[v] = f;
        }
        if (sentinel && blockIdx.x == 0) reinterpret_cast<cd *>(out)[nchild] = cd{0.0, 0.0};
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const uint32_t sent_word = nparent << 6;
#define CNT(d) cnt[(size_t)(d) * SLOS7_THREADS + tid]
#define SLOT0(d, t) slot0[(size_t)(d) * SLOS7_THREADS + (t)]
#define OWN(s, j, t) own[(size_t)((s) * ((s) - 1) / 2 + (j)) * SLOS7_THREADS + (t)]
    const XcdTiles xt = xcd_tiles((uint32_t)(((uint64_t)nchild + TILE - 1) / TILE));
    for (uint32_t tile = xt.first; tile < xt.end; tile += xt.step) {
        const uint32_t blk0 = tile * TILE;
        // phase 1: this thread's RUN consecutive children
        const uint32_t first = blk0 + (uint32_t)tid * RUN;
        int depth = 0;
        pf_t pfact = (pf_t)1;
        if (first < nchild) {
            skip_walk(first, m, k, sP, wp, sCT, sCE, [&](int q, int sq, uint32_t pe) {
                SLOT0(depth, tid) = ((first - pe) << 6) | (uint32_t)q;
                CNT(depth) = (uint8_t)sq;
                depth++;
                if (MODE == 1) pfact *= (pf_t)sf[sq];
            });
        }
        int kp = depth;
        skeep[tid] = (uint8_t)depth;
        sdep[tid] = first < nchild ? (uint8_t)depth : (uint8_t)0;
        if (MODE == 1) spf[tid] = pfact;
#pragma unroll
        for (int s = 1; s < RUN; s++) {
            const uint32_t idx = first + (uint32_t)s;
            const int cs = s * SLOS7_THREADS + tid;
            if (idx >= nchild) { sdep[cs] = 0; skeep[cs] = 0; continue; }
            // slot word of entry d of child s-1 (parent ranks of child s-1)
            auto word = [&](int d) -> uint32_t {
                return d < kp ? SLOT0(d, tid) + 64u * (uint32_t)(s - 1) : OWN(s - 1, d - kp, tid);
            };
            int top = depth - 1;
            uint32_t ts = word(top) + 64u;
            int tail = 0;
            if ((int)(ts & 63u) == m - 1) {   // idx - 1 is not the last state: an entry below m-1 remains
                tail = CNT(top);
                top--;
                ts = word(top) + 64u;
            }
            const int i = (int)(ts & 63u);
            const int tc = CNT(top);
            const uint32_t pe_top = idx - (ts >> 6);
            int newlen = top;
            if (tc > 1) {
                CNT(top) = (uint8_t)(tc - 1);
                newlen = top + 1;
            }
            const int nkp = min(kp, newlen);
            for (int d = nkp; d < newlen; d++) OWN(s, d - nkp, tid) = word(d) + 64u;
            const uint32_t pe_new = pe_top + (uint32_t)(sP[i * wp + tail + 1] >> 32);
            OWN(s, newlen - nkp, tid) = ((idx - pe_new) << 6) | (uint32_t)(i + 1);
            CNT(newlen) = (uint8_t)(tail + 1);
            depth = newlen + 1;
            kp = nkp;
            skeep[cs] = (uint8_t)kp;
            sdep[cs] = (uint8_t)depth;
            if (MODE == 1) {
                pf_t pf = (pf_t)1;
                for (int d = 0; d < depth; d++) pf *= (pf_t)sf[CNT(d)];
                spf[cs] = pf;
            }
        }
        __syncthreads();
        // phase 2: lane-consecutive children; child c is step c % RUN of thread c / RUN
        for (int r = 0; r < RUN; r++) {
            const uint32_t c = (uint32_t)r * SLOS7_THREADS + (uint32_t)tid;
            const uint32_t idx = blk0 + c;
            if (idx >= nchild) break;
            const uint32_t o = c / RUN, s = c % RUN;
            const int cs = (int)(s * SLOS7_THREADS + o);
            const int dd = sdep[cs], kk = skeep[cs];
            cd acc = {0.0, 0.0};
            for (int d0 = 0; d0 < dd; d0 += G) {
                cd pv[G];
                int qv[G];
#pragma unroll
                for (int t = 0; t < G; t++) {
                    const int d = d0 + t;
                    uint32_t w32 = sent_word;
                    if (d < kk) w32 = SLOT0(d, o) + 64u * s;
                    else if (d < dd) w32 = own[(size_t)(s * (s - 1) / 2 + (uint32_t)(d - kk)) * SLOS7_THREADS + o];
                    pv[t] = parent[min(w32 >> 6, nparent)];
                    qv[t] = (int)(w32 & 63u);
                }
#pragma unroll
                for (int t = 0; t < G; t++) {
                    const cd u = su[qv[t]];
                    acc.r = fma(pv[t].r, u.r, acc.r);
                    acc.r = fma(-pv[t].i, u.i, acc.r);
                    acc.i = fma(pv[t].r, u.i, acc.i);
                    acc.i = fma(pv[t].i, u.r, acc.i);
                }
            }
            if (MODE == 0) {
                reinterpret_cast<cd *>(out)[idx] = acc;
            } else {
                reinterpret_cast<double *>(out)[idx] = prob_from_coef(acc, (double)spf[cs], inv_prod_in);
            }
        }
        __syncthreads();
    }
#undef OWN
#undef SLOT0
#undef CNT
}

// Writes the vacuum layer: coefficient (1, 0), or probability 1.
__global__ void slos_vacuum_kernel(double *dst, int as_prob)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dst[0] = 1.0;
        if (!as_prob) dst[1] = 0.0;
        // as_prob == 2: coefficient layer in the workspace, followed by the
        // zero coefficient the successor kernel's gathers read past a depth
        if (as_prob == 2) { dst[1] = 0.0; dst[2] = 0.0; dst[3] = 0.0; }
    }
}

// Rank of each occupation vector (int32 rows of m), -1 for an invalid row.
__global__ void fock_rank_kernel(const int32_t *__restrict__ occ, int64_t count, int m, int nmax,
                                 const uint64_t *__restrict__ gT, int w, int64_t *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int32_t *s = occ + i * m;
    int n = 0;
    bool ok = true;
    for (int q = 0; q < m; q++) { ok = ok && s[q] >= 0; n += s[q]; }
    if (!ok || n > nmax) { out[i] = -1; return; }
    int R = n;
    uint64_t r = 0;
    for (int q = 0; q + 1 < m; q++) {
        R -= s[q];
        r += gT[q * w + R];
    }
    out[i] = (int64_t)r;
}

// Occupation vector of each rank in the n-photon space; rows of m int32.
__global__ void fock_unrank_kernel(const int64_t *__restrict__ ranks, int64_t count, int m, int n,
                                   const uint64_t *__restrict__ gT, int w, uint64_t total, int32_t *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int32_t *s = out + i * m;
    const int64_t rk = ranks[i];
    if (rk < 0 || (uint64_t)rk >= total) {
        for (int q = 0; q < m; q++) s[q] = -1;
        return;
    }
    uint64_t rem = (uint64_t)rk;
    int R = n;
    for (int q = 0; q < m; q++) {
        int Rn = 0;
        if (q + 1 < m) {
            const uint64_t *tq = gT + q * w;
            Rn = R;
            uint64_t tv = tq[Rn];
            while (Rn > 0 && tv > rem) { --Rn; tv = tq[Rn]; }
            rem -= tv;
        }
        s[q] = R - Rn;
        R = Rn;
    }
}

// ---------------------------------------------------------------- masks
//
// FSMask semantics (AStrongSimulationBackend.set_mask,
// perceval/backends/_abstract_backends.py:103-146; tests/utils/test_mask.py:
// 33-45): a mask holds one condition per mode, -1 for any count, else a count
// that must match exactly, or be reached at least on the modes flagged in
// at_least; masks are OR-ed.  Selection runs over the ranks of the n-photon
// space: a count pass unranks each rank, tests it and counts matches per
// block; a one-block scan turns counts into offsets; a scatter pass writes the
// matching ranks in rank order (FSArray order).
constexpr int MASK_THREADS = 256;
constexpr int MASK_MAX = 16;   // masks per selection

// Walks rank rk of the n-photon space and tests every mask on the way; bit k
// of the result is set when mask k can be completed by `spare` more photons:
// no exact count exceeded, the photons missing from the exact counts and the
// floors at most `spare`, and exactly `spare` when mask k has no mode to take
// the surplus (bit k of `absorbs`: a free or at-least mode; spare 0: an exact
// match).
__device__ __forceinline__ uint32_t mask_walk(uint64_t rk, int m, int n, const uint64_t *sT, int w,
                                              const int32_t *cond, int nmask, const int32_t *at_least, int spare,
                                              uint32_t absorbs)
{
    uint32_t ok = (nmask >= 32) ? 0xffffffffu : ((1u << nmask) - 1u);
    int miss[MASK_MAX];
#pragma unroll
    for (int k = 0; k < MASK_MAX; k++) miss[k] = 0;
    uint64_t rem = rk;
    int R = n;
    for (int q = 0; q < m; q++) {
        int Rn = 0;
        if (q + 1 < m) {
            const uint64_t *tq = sT + q * w;
            Rn = R;
            uint64_t tv = tq[Rn];
            while (Rn > 0 && tv > rem) { --Rn; tv = tq[Rn]; }
            rem -= tv;
        }
        const int sq = R - Rn;
#pragma unroll
        for (int k = 0; k < MASK_MAX; k++) {
            if (k >= nmask) break;
            const int c = cond[k * m + q];
            if (c < 0) continue;
            if (at_least[q]) miss[k] += c > sq ? c - sq : 0;
            else if (sq > c) ok &= ~(1u << k);
            else miss[k] += c - sq;
        }
        R = Rn;
    }
#pragma unroll
    for (int k = 0; k < MASK_MAX; k++)
        if (k < nmask && (miss[k] > spare || (miss[k] < spare && !((absorbs >> k) & 1u)))) ok &= ~(1u << k);
    return ok;
}

// flags[i] = match of rank i; counts[b] = matches in block b.
__global__ void __launch_bounds__(MASK_THREADS) mask_count_kernel(uint64_t total, int m, int n,
                                                                  const uint64_t *__restrict__ gT, int w,
                                                                  const int32_t *__restrict__ gcond, int nmask,
                                                                  const int32_t *__restrict__ gat, int spare,
                                                                  uint32_t absorbs,
                                                                  uint8_t *__restrict__ flags,
                                                                  int32_t *__restrict__ counts)
{
    // rank table (m x w), conditions (nmask x m) and floors (m), sized per launch
    extern __shared__ uint64_t s_mask[];
    uint64_t *sT = s_mask;
    int32_t *scond = reinterpret_cast<int32_t *>(sT + (size_t)m * w);
    int32_t *sat = scond + (size_t)nmask * m;
    for (int e = threadIdx.x; e < m * w; e += MASK_THREADS) sT[e] = gT[e];
    for (int e = threadIdx.x; e < nmask * m; e += MASK_THREADS) scond[e] = gcond[e];
    for (int e = threadIdx.x; e < m; e += MASK_THREADS) sat[e] = gat[e];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * MASK_THREADS + threadIdx.x;
    bool hit = false;
    if (i < total) {
        hit = mask_walk(i, m, n, sT, w, scond, nmask, sat, spare, absorbs) != 0u;
        flags[i] = hit ? 1 : 0;
    }
    const int c = __syncthreads_count(hit);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// Exclusive prefix of the block counts (one block); offsets[nb] = total matches.
__global__ void __launch_bounds__(1024) mask_scan_kernel(const int32_t *__restrict__ counts, int64_t nb,
                                                         int64_t *__restrict__ offsets, int64_t *__restrict__ d_count)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t lo = t * per, hi = (lo + per < nb) ? lo + per : nb;
    int64_t sum = 0;
    for (int64_t i = lo; i < hi; i++) sum += counts[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - sum;
    for (int64_t i = lo; i < hi; i++) { offsets[i] = run; run += counts[i]; }
    if (t == 1023) { offsets[nb] = part[1023]; *d_count = part[1023]; }
}

// Matching ranks, in rank order.
__global__ void __launch_bounds__(MASK_THREADS) mask_scatter_kernel(uint64_t total, const uint8_t *__restrict__ flags,
                                                                    const int64_t *__restrict__ offsets,
                                                                    int64_t *__restrict__ out)
{
    __shared__ int32_t wave_base[MASK_THREADS / 64];
    const uint64_t i = (uint64_t)blockIdx.x * MASK_THREADS + threadIdx.x;
    const bool hit = i < total && flags[i];
    const uint64_t bal = __ballot(hit);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int before = __builtin_popcountll(bal & ((1ull << lane) - 1));
    if (lane == 0) wave_base[wv] = __builtin_popcountll(bal);
    __syncthreads();
    int base = 0;
    for (int k = 0; k < wv; k++) base += wave_base[k];
    if (hit) out[offsets[blockIdx.x] + base + before] = (int64_t)i;
}

__global__ void gather_f64_kernel(const double *__restrict__ src, const int64_t *__restrict__ idx, int64_t count,
                                  double *__restrict__ dst)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) dst[k] = src[idx[k]];
}

struct SlosPlanImpl {
    int m = 0, n = 0, w = 0;
    std::vector<int32_t> mk;          // input mode of each added photon
    std::vector<uint64_t> counts;     // layer sizes N_0..N_n
    double inv_prod_in = 1.0;         // 1 / prod t!
    uint64_t max_inter = 1;           // largest intermediate layer (k < n)
    uint64_t *d_T = nullptr, *d_E = nullptr;
    uint64_t *d_P = nullptr;          // packed table of the successor kernel (inside the d_T allocation)
    int grid_cap = 0;
    int kernel = 0;                   // 0 auto, 1 per state, 4 per state in two phases,
                                      // 6 successor runs with a transposed gather, 8 prefix/suffix blocks
    BlkLayout blk;                    // block layout (kernel 8), built when the layers allow it
    bool masked = false;              // layers restricted to a mask (kernel 8 only)
    BlkMasked mask;                   // their sub-layers
    uint32_t *d_perm = nullptr;       // final layer in FSArray order (masks with several final sub-layers)
};

static size_t twophase_lds(int m, int w, int k)
{
    return sizeof(cd) * m + sizeof(double) * w + 2 * sizeof(uint32_t) * (size_t)m * w +
           sizeof(uint32_t) * (size_t)k * SLOS4_THREADS;
}

// The two-phase kernel packs a parent rank and a mode in 32 bits.
static bool twophase_ok(const SlosPlanImpl &p)
{
    for (int k = 0; k < p.n; k++)
        if (p.counts[k] >= (1ull << 26)) return false;
    return p.counts[p.n] < (1ull << 32) && twophase_lds(p.m, p.w, p.n) <= 64 * 1024;
}

// Successor kernel: compact slot lists, runs of 3 children per thread.  At
// m=24, n=10 runs of 2 / 3 / 4 / 5 take 1.80 / 1.71 / 1.86 / 1.98 ms: longer
// runs save unranking work but need more LDS per block, so fewer waves.
static size_t succ2_lds(int m, int n, int k, int run, bool probs)
{
    const size_t rows = (size_t)k + (size_t)run * (run - 1) / 2;
    const size_t pf = probs ? (k <= 12 ? sizeof(uint32_t) : sizeof(double)) * run * SLOS7_THREADS : 0;
    return sizeof(uint64_t) * (size_t)m * (n + 2) + 2 * sizeof(uint32_t) * (size_t)(n + 1) * (m + 4) +
           sizeof(cd) * m + sizeof(double) * (n + 2) + pf + sizeof(uint32_t) * rows * SLOS7_THREADS +
           (size_t)(k + 1) * SLOS7_THREADS + 2 * (size_t)run * SLOS7_THREADS;
}

static bool succ_ok(const SlosPlanImpl &p)
{
    return twophase_ok(p) && succ2_lds(p.m, p.n, p.n, 3, true) <= 64 * 1024;
}

static bool blk_ok(const SlosPlanImpl &p) { return p.blk.d_mem != nullptr; }

// (Re)builds a masked plan's block layout with suffix width w (0: default).
static int masked_layout(SlosPlanImpl &p, int w)
{
    const BlkMasked &mk = p.mask;
    if (w == 0) w = blk_default_width(mk.m_free, mk.nmax_free);
    if (w < 1 || w >= mk.m_free || w > 16) { set_error("suffix width outside [1, min(free modes - 1, 16)]"); return ERR_ARG; }
    BlkLayout lay;
    std::vector<char> host;
    int rc = blk_layout_build_host_ex(mk.m_free, mk.nmax_free, w, mk.subs, mk.counts, mk.rowmap, &lay, &host);
    if (rc == OK) rc = blk_layout_upload(&lay, host);
    if (rc != OK) return rc;
    blk_layout_destroy(&p.blk);
    p.blk = std::move(lay);
    return OK;
}

static int device_grid_cap(int *cap)
{
    int dev = 0, cus = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *cap = cus * 16;
    return OK;
}

} // namespace pcvl

using namespace pcvl;

struct pcvl_slos_plan_s {
    SlosPlanImpl impl;
};

extern "C" int pcvl_slos_plan_create(int m, const int32_t *h_in_occ, pcvl_slos_plan *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out plan pointer is null");
    PCVL_ARG_CHECK(h_in_occ != nullptr, "input state is null");
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 256]");
    *out = nullptr;
    int n = 0;
    for (int i = 0; i < m; i++) {
        PCVL_ARG_CHECK(h_in_occ[i] >= 0, "negative occupation");
        n += h_in_occ[i];
    }
    PCVL_ARG_CHECK(n <= SLOS_MAX_PHOTONS, "more than 40 photons");
    auto *plan = new (std::nothrow) pcvl_slos_plan_s();
    if (!plan) { set_error("host allocation failed"); return ERR_ALLOC; }
    SlosPlanImpl &p = plan->impl;
    p.m = m; p.n = n; p.w = n + 1;
    // _Path._decompose for a single input (perceval/backends/_slos.py:61-86):
    // add photons to the mode holding the most remaining photons, lowest mode first.
    std::vector<int32_t> rem(h_in_occ, h_in_occ + m);
    for (int k = 0; k < n; k++) {
        int best = 0;
        for (int i = 1; i < m; i++) if (rem[i] > rem[best]) best = i;
        p.mk.push_back(best);
        rem[best]--;
    }
    double pin = 1.0;
    for (int i = 0; i < m; i++) for (int v = 2; v <= h_in_occ[i]; v++) pin *= v;
    p.inv_prod_in = 1.0 / pin;
    for (int k = 0; k <= n; k++) {
        uint64_t c = binom_u64((int64_t)m + k - 1, k);
        if (c == UINT64_MAX || c > (1ull << 40)) {
            delete plan;
            set_error("Fock space too large (more than 2^40 states in a layer)");
            return ERR_RANGE;
        }
        p.counts.push_back(c);
        if (k < n && c > p.max_inter) p.max_inter = c;
    }
    std::vector<uint64_t> T, E, P;
    build_tables(m, n, T, E);
    build_packed(m, n, P);
    size_t bytes = sizeof(uint64_t) * T.size();
    hipError_t e = hipMalloc(&p.d_T, 2 * bytes + sizeof(uint64_t) * P.size());
    if (e != hipSuccess) { delete plan; set_error(std::string("hipMalloc(slos tables): ") + hipGetErrorString(e)); return ERR_ALLOC; }
    p.d_E = p.d_T + T.size();
    p.d_P = p.d_E + E.size();
    e = hipMemcpy(p.d_T, T.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.d_E, E.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.d_P, P.data(), sizeof(uint64_t) * P.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p.d_T);
        delete plan;
        set_error(std::string("hipMemcpy(slos tables): ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    int rc = device_grid_cap(&p.grid_cap);
    // the block kernel where its layout fits (a layout out of range leaves the others)
    if (rc == OK && blk_layout_ok(m, n)) {
        rc = blk_layout_create(m, n, blk_default_width(m, n), &p.blk);
        if (rc == ERR_RANGE) rc = OK;
    }
    // the per-state kernel, which every other falls back to, stages its rank tables in LDS
    if (rc == OK && !blk_ok(p) && 2 * sizeof(uint64_t) * (size_t)m * (size_t)p.w > 160 * 1024) {
        set_error("SLOS: modes and photons too many for any layer kernel's tables");
        rc = ERR_RANGE;
    }
    if (rc != OK) { blk_layout_destroy(&p.blk); (void)hipFree(p.d_T); delete plan; return rc; }
    *out = plan;
    return OK;
}

extern "C" int pcvl_slos_plan_create_masked(int m, const int32_t *h_in_occ, const int32_t *h_cond,
                                            pcvl_slos_plan *out)
{
    return pcvl_slos_plan_create_masks(m, h_in_occ, 1, h_cond, nullptr, 0, out);
}

extern "C" int pcvl_slos_plan_create_masks(int m, const int32_t *h_in_occ, int nrows, const int32_t *h_conds,
                                           const int32_t *h_at_least, int mask_n, pcvl_slos_plan *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out plan pointer is null");
    PCVL_ARG_CHECK(h_in_occ != nullptr && h_conds != nullptr, "input state or mask is null");
    PCVL_ARG_CHECK(m >= 2 && m <= SLOS_MAX_MODES, "mode count outside [2, 256]");
    PCVL_ARG_CHECK(nrows >= 1 && nrows <= 64, "mask row count outside [1, 64]");
    *out = nullptr;
    int n = 0, nfree = 0;
    for (int i = 0; i < m; i++) {
        PCVL_ARG_CHECK(h_in_occ[i] >= 0, "negative occupation");
        n += h_in_occ[i];
        bool free_mode = true;
        for (int r = 0; r < nrows; r++) {
            PCVL_ARG_CHECK(h_conds[(size_t)r * m + i] >= -1, "mask condition below -1");
            free_mode = free_mode && h_conds[(size_t)r * m + i] < 0;
        }
        nfree += free_mode ? 1 : 0;
    }
    PCVL_ARG_CHECK(n >= 1 && n <= 30, "masked plans take 1 to 30 photons");
    PCVL_ARG_CHECK(mask_n == 0 || (mask_n >= n && mask_n <= 1024),
                   "mask photon count must be 0 (the input's) or between the input's and 1024");
    PCVL_ARG_CHECK(nfree >= 2, "a masked plan needs at least two modes that no mask row constrains");
    auto *plan = new (std::nothrow) pcvl_slos_plan_s();
    if (!plan) { set_error("host allocation failed"); return ERR_ALLOC; }
    SlosPlanImpl &p = plan->impl;
    try {
        p.m = m; p.n = n; p.w = n + 1;
        p.masked = true;
        p.kernel = 8;
        std::vector<int32_t> rem(h_in_occ, h_in_occ + m);
        for (int k = 0; k < n; k++) {   // _Path._decompose, as in pcvl_slos_plan_create
            int best = 0;
            for (int i = 1; i < m; i++) if (rem[i] > rem[best]) best = i;
            p.mk.push_back(best);
            rem[best]--;
        }
        double pin = 1.0;
        for (int i = 0; i < m; i++) for (int v = 2; v <= h_in_occ[i]; v++) pin *= v;
        int rc = device_grid_cap(&p.grid_cap);
        if (rc == OK) rc = blk_masked_layers_ex(m, n, mask_n > n ? mask_n : n, nrows, h_conds, h_at_least, &p.mask);
        if (rc != OK) { delete plan; return rc; }
        // the herald modes' factorials go with each final sub-layer (its groups)
        p.inv_prod_in = 1.0 / pin;
        p.counts = p.mask.counts;
        for (int k = 0; k < n; k++) p.max_inter = std::max<uint64_t>(p.max_inter, p.counts[(size_t)k]);
        if (!p.mask.perm.empty()) {
            const size_t b = sizeof(uint32_t) * p.mask.perm.size();
            hipError_t e = hipMalloc((void **)&p.d_perm, b);
            if (e == hipSuccess) e = hipMemcpy(p.d_perm, p.mask.perm.data(), b, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                if (p.d_perm) (void)hipFree(p.d_perm);
                delete plan;
                set_error(std::string("masked SLOS output order upload: ") + hipGetErrorString(e));
                return ERR_HIP;
            }
            std::vector<uint32_t>().swap(p.mask.perm);   // the host copy is no longer needed
        }
        if (!p.mask.empty) {
            rc = masked_layout(p, 0);
            if (rc != OK) { if (p.d_perm) (void)hipFree(p.d_perm); delete plan; return rc; }
        }
    } catch (const std::bad_alloc &) {
        if (p.d_perm) (void)hipFree(p.d_perm);
        delete plan;
        set_error("host allocation failed while building the masked plan");
        return ERR_ALLOC;
    }
    *out = plan;
    return OK;
}

// Complex entries of the two layer buffers at the start of the workspace.
static uint64_t slos_work_main(const SlosPlanImpl &p) { return p.n >= 2 ? 2 * (p.max_inter + 1) : 2; }

// Masked final layer in FSArray order: dst[i] = src[perm[i]].
template <typename T>
__global__ void slos_perm_gather_kernel(const T *__restrict__ src, const uint32_t *__restrict__ perm, uint64_t count,
                                        T *__restrict__ dst)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[perm[i]];
}

extern "C" int pcvl_slos_plan_info(pcvl_slos_plan plan, int *n, int64_t *out_count, size_t *work_bytes,
                                   int32_t *h_path)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    const SlosPlanImpl &p = plan->impl;
    if (n) *n = p.n;
    if (out_count) *out_count = (int64_t)p.counts[p.n];
    // two layer buffers of the largest intermediate layer (complex), each
    // with one zero coefficient after its states; masks with several final
    // sub-layers add the final layer in sub-layer order (reordered at the end)
    if (work_bytes) *work_bytes = sizeof(cd) * (slos_work_main(p) + (p.d_perm ? p.counts[(size_t)p.n] : 0));
    if (h_path) for (int k = 0; k < p.n; k++) h_path[k] = p.mk[k];
    return OK;
}

extern "C" int pcvl_slos_plan_execute(pcvl_slos_plan plan, const double *d_U, int mode, double *d_out,
                                      void *d_work, size_t work_bytes, hipStream_t stream)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    PCVL_ARG_CHECK(mode == PCVL_SLOS_COEFS || mode == PCVL_SLOS_PROBS, "mode must be PCVL_SLOS_COEFS or PCVL_SLOS_PROBS");
    PCVL_ARG_CHECK(d_U && d_out, "null device buffer");
    const SlosPlanImpl &p = plan->impl;
    if (p.masked && p.counts[(size_t)p.n] == 0) return OK;   // the heralds need more photons than there are
    size_t need = 0;
    pcvl_slos_plan_info(plan, nullptr, nullptr, &need, nullptr);
    PCVL_ARG_CHECK(d_work != nullptr && work_bytes >= need, "workspace missing or smaller than pcvl_slos_plan_info reports");
    const cd *U = reinterpret_cast<const cd *>(d_U);
    cd *bufA = reinterpret_cast<cd *>(d_work);
    cd *bufB = bufA + p.max_inter + 1;
    const size_t lds = 2 * sizeof(uint64_t) * (size_t)p.m * p.w;
    if (p.n == 0) {
        // Vacuum: one output state with coefficient 1 (probability 1).
        hipLaunchKernelGGL(slos_vacuum_kernel, dim3(1), dim3(64), 0, stream, d_out, mode == PCVL_SLOS_PROBS ? 1 : 0);
        PCVL_HIP_CHECK(hipGetLastError());
        return OK;
    }
    // Auto takes the block kernel where its layout was built (every layer
    // below 2^32 states), else the compact successor kernel where its packed
    // parent ranks and LDS fit (parent layers below 2^26 states), else the
    // two-phase kernel where its ranks fit, else the per-state kernel.
    const int kernel = p.masked ? 8 : p.kernel != 0 ? p.kernel : blk_ok(p) ? 8 : succ_ok(p) ? 6 : twophase_ok(p) ? 4 : 1;
    if ((kernel == 4 && !twophase_ok(p)) || (kernel == 6 && !succ_ok(p)) || (kernel == 8 && !blk_ok(p))) {
        set_error("SLOS kernel not applicable to this plan's layer sizes");
        return ERR_ARG;
    }
    // Layer 0 is the vacuum coefficient 1; layer 1 reads it from bufA[0].
    hipLaunchKernelGGL(slos_vacuum_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<double *>(bufA), 2);
    PCVL_HIP_CHECK(hipGetLastError());
    const cd *prev = bufA;
    // masks with several final sub-layers: the last layer goes to the end of
    // the workspace in sub-layer order, then to d_out in FSArray order
    void *last_dst = p.d_perm ? static_cast<void *>(bufA + slos_work_main(p)) : static_cast<void *>(d_out);
    for (int k = 1; k <= p.n; k++) {
        const uint64_t nc = p.counts[k];
        const bool last = (k == p.n);
        void *dst = last ? last_dst : static_cast<void *>(prev == bufA ? bufB : bufA);
        const bool probs = last && mode == PCVL_SLOS_PROBS;
        if (kernel == 8) {
            const int rc = blk_layer_launch(p.blk, prev, dst, U, p.m, p.mk[k - 1], k, probs, p.inv_prod_in, !last,
                                            stream);
            if (rc != OK) return rc;
        } else if (kernel == 6) {
            const uint64_t tile = 3ull * SLOS7_THREADS;
            // a multiple of 8 workgroups, for the XCD-contiguous tile order
            const int grid = (int)((std::min<uint64_t>((nc + tile - 1) / tile, (uint64_t)p.grid_cap) + 7) / 8 * 8);
            const bool pfi = k <= 12;
            auto kp = probs ? (pfi ? slos_layer_succ2_kernel<1, 3, true, 4> : slos_layer_succ2_kernel<1, 3, false, 4>)
                            : slos_layer_succ2_kernel<0, 3, true, 4>;
            hipLaunchKernelGGL(kp, dim3(grid), dim3(SLOS7_THREADS), succ2_lds(p.m, p.n, k, 3, probs), stream, prev,
                               (uint32_t)p.counts[k - 1], dst, U, p.m, p.mk[k - 1], k, (uint32_t)nc, p.d_P, p.n + 2,
                               p.inv_prod_in, last ? 0 : 1);
        } else if (kernel == 4) {
            const int grid = (int)std::min<uint64_t>((nc + SLOS4_THREADS - 1) / SLOS4_THREADS, (uint64_t)p.grid_cap);
            auto kp = probs ? slos_layer2p_kernel<1, 4> : slos_layer2p_kernel<0, 4>;
            hipLaunchKernelGGL(kp, dim3(grid), dim3(SLOS4_THREADS), twophase_lds(p.m, p.w, k), stream, prev,
                               (uint32_t)p.counts[k - 1], dst, U, p.m, p.mk[k - 1], k, (uint32_t)nc, p.d_T, p.d_E, p.w,
                               p.inv_prod_in);
        } else {
            const int grid = (int)std::min<uint64_t>((nc + SLOS_THREADS - 1) / SLOS_THREADS, (uint64_t)p.grid_cap);
            if (probs)
                hipLaunchKernelGGL(slos_layer_kernel<1>, dim3(grid), dim3(SLOS_THREADS), lds, stream, prev,
                                   p.counts[k - 1], dst, U, p.m, p.mk[k - 1], k, nc, p.d_T, p.d_E, p.w, p.inv_prod_in);
            else
                hipLaunchKernelGGL(slos_layer_kernel<0>, dim3(grid), dim3(SLOS_THREADS), lds, stream, prev,
                                   p.counts[k - 1], dst, U, p.m, p.mk[k - 1], k, nc, p.d_T, p.d_E, p.w, p.inv_prod_in);
        }
        PCVL_HIP_CHECK(hipGetLastError());
        prev = reinterpret_cast<const cd *>(dst);
    }
    if (p.d_perm) {
        const uint64_t cnt = p.counts[(size_t)p.n];
        const int grid = (int)std::min<uint64_t>((cnt + 255) / 256, (uint64_t)p.grid_cap);
        if (mode == PCVL_SLOS_PROBS)
            hipLaunchKernelGGL(slos_perm_gather_kernel<double>, dim3(grid), dim3(256), 0, stream,
                               reinterpret_cast<const double *>(last_dst), p.d_perm, cnt, d_out);
        else
            hipLaunchKernelGGL(slos_perm_gather_kernel<cd>, dim3(grid), dim3(256), 0, stream,
                               reinterpret_cast<const cd *>(last_dst), p.d_perm, cnt, reinterpret_cast<cd *>(d_out));
        PCVL_HIP_CHECK(hipGetLastError());
    }
    return OK;
}

extern "C" int pcvl_slos_plan_set_kernel(pcvl_slos_plan plan, int kernel)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    PCVL_ARG_CHECK(kernel == 0 || kernel == 1 || kernel == 4 || kernel == 6 || kernel == 8,
                   "kernel must be 0 (auto), 1 (per state), 4 (per state, two phases), 6 (successor runs) "
                   "or 8 (prefix/suffix blocks)");
    if (kernel == 4 || kernel == 6)
        PCVL_ARG_CHECK(twophase_ok(plan->impl), "two-phase and successor kernels need parent layers below 2^26 states");
    if (kernel == 6) PCVL_ARG_CHECK(succ_ok(plan->impl), "successor kernel needs its LDS below 64 KiB");
    if (kernel == 8) PCVL_ARG_CHECK(blk_ok(plan->impl), "block kernel needs every layer below 2^32 states and m >= 2");
    if (plan->impl.masked) {
        PCVL_ARG_CHECK(kernel == 0 || kernel == 8, "masked plans run the block kernel only (0 or 8)");
        return OK;
    }
    plan->impl.kernel = kernel;
    return OK;
}

extern "C" int pcvl_slos_plan_set_block_width(pcvl_slos_plan plan, int w)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    SlosPlanImpl &p = plan->impl;
    if (p.masked) {
        if (p.mask.empty) return OK;
        if (w == 0) w = blk_default_width(p.mask.m_free, p.mask.nmax_free);
        if (p.blk.d_mem && p.blk.w == w) return OK;
        PCVL_HIP_CHECK(hipDeviceSynchronize());   // the previous layout may still be read by enqueued work
        return masked_layout(p, w);
    }
    PCVL_ARG_CHECK(blk_layout_ok(p.m, p.n), "block kernel needs every layer below 2^32 states, m >= 2 and n in [1, 30]");
    if (w == 0) w = blk_default_width(p.m, p.n);
    PCVL_ARG_CHECK(w >= 1 && w < p.m && w <= 16, "suffix width outside [1, min(m - 1, 16)]");
    if (p.blk.d_mem && p.blk.w == w) return OK;
    BlkLayout lay;
    const int rc = blk_layout_create(p.m, p.n, w, &lay);
    if (rc != OK) return rc;
    // the previous layout may still be read by enqueued work
    PCVL_HIP_CHECK(hipDeviceSynchronize());
    blk_layout_destroy(&p.blk);
    p.blk = std::move(lay);
    return OK;
}

extern "C" int pcvl_slos_plan_block_width(pcvl_slos_plan plan)
{
    if (plan == nullptr) return 0;
    return plan->impl.blk.d_mem ? plan->impl.blk.w : 0;
}

extern "C" int pcvl_slos_plan_destroy(pcvl_slos_plan plan)
{
    if (!plan) return OK;
    hipError_t e = hipSuccess;
    if (plan->impl.d_T) e = hipFree(plan->impl.d_T);
    if (plan->impl.d_perm) {
        const hipError_t e2 = hipFree(plan->impl.d_perm);
        if (e == hipSuccess) e = e2;
    }
    blk_layout_destroy(&plan->impl.blk);
    delete plan;
    if (e != hipSuccess) { set_error(std::string("hipFree(slos plan): ") + hipGetErrorString(e)); return ERR_HIP; }
    return OK;
}

// ------------------------------------------------------------- input trees
//
// SLOSBackend.preprocess([inputs]) builds one _Path tree over several inputs
// (perceval/backends/_slos.py:37-86,170-185): at each node the photon counts
// still to add are summed over the node's targets, the mode with the largest
// sum (lowest on ties) becomes a child taking every target that still has a
// photon there, and the rest is decomposed again; a target with no photon
// left is mapped to the node.  Shared prefixes are computed once.  Execution
// is depth first: node layers live in one buffer per depth, so a node's
// parent layer is still in place when the node is computed.

struct TreeOp {
    int k;              // layer of the node (photons)
    int mk;             // input mode of the photon added by the node
    int to_level;       // 1: write the node's coefficients to its depth buffer
    int input;          // >= 0: also write this input's output from the node
    double inv_prod_in; // 1 / prod t! of that input
};

struct SlosTreeImpl {
    int m = 0, nmax = 0, w = 0, ninputs = 0;
    std::vector<uint64_t> counts;       // N_0..N_nmax
    std::vector<int64_t> out_off;       // element offset of each input's output
    std::vector<int> in_n;              // photons of each input
    std::vector<size_t> level_off;      // byte offset of depth buffer k in the workspace
    size_t work_bytes = 0;
    std::vector<TreeOp> ops;            // depth-first launch list
    uint64_t *d_T = nullptr, *d_E = nullptr;
    int grid_cap = 0;
    BlkLayout blk;                      // block layout of the layers, when they allow it
};

static void tree_decompose(SlosTreeImpl &t, int k, int mk, std::vector<std::vector<int32_t>> targets,
                           std::vector<int> ids, const std::vector<double> &inv_pin)
{
    std::vector<int> here;
    std::vector<std::vector<int32_t>> tg;
    std::vector<int> tid;
    for (size_t i = 0; i < targets.size(); i++) {
        int left = 0;
        for (int v : targets[i]) left += v;
        if (left == 0) here.push_back(ids[i]);
        else { tg.push_back(targets[i]); tid.push_back(ids[i]); }
    }
    // the node itself: its layer feeds children (to_level) and/or outputs
    const bool has_children = !tg.empty();
    if (k > 0) {
        if (has_children) t.ops.push_back(TreeOp{k, mk, 1, -1, 1.0});
        for (int id : here) t.ops.push_back(TreeOp{k, mk, 0, id, inv_pin[(size_t)id]});
    } else {
        for (int id : here) t.ops.push_back(TreeOp{0, -1, 0, id, 1.0});
    }
    while (!tg.empty()) {
        std::vector<int> cnt((size_t)t.m, 0);
        for (auto &x : tg) for (int q = 0; q < t.m; q++) cnt[(size_t)q] += x[(size_t)q];
        int best = 0;
        for (int q = 1; q < t.m; q++) if (cnt[(size_t)q] > cnt[(size_t)best]) best = q;
        std::vector<std::vector<int32_t>> child, rest;
        std::vector<int> cid, rid;
        for (size_t i = 0; i < tg.size(); i++) {
            if (tg[i][(size_t)best] > 0) {
                tg[i][(size_t)best] -= 1;
                child.push_back(tg[i]);
                cid.push_back(tid[i]);
            } else {
                rest.push_back(tg[i]);
                rid.push_back(tid[i]);
            }
        }
        tree_decompose(t, k + 1, best, std::move(child), std::move(cid), inv_pin);
        tg = std::move(rest);
        tid = std::move(rid);
    }
}

struct pcvl_slos_tree_s {
    SlosTreeImpl impl;
};

extern "C" int pcvl_slos_tree_create(int m, const int32_t *h_inputs, int count, pcvl_slos_tree *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out plan pointer is null");
    PCVL_ARG_CHECK(count >= 1 && h_inputs != nullptr, "no input states");
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 256]");
    *out = nullptr;
    auto *plan = new (std::nothrow) pcvl_slos_tree_s();
    if (!plan) { set_error("host allocation failed"); return ERR_ALLOC; }
    SlosTreeImpl &t = plan->impl;
    t.m = m;
    t.ninputs = count;
    std::vector<std::vector<int32_t>> targets((size_t)count);
    std::vector<int> ids((size_t)count);
    std::vector<double> inv_pin((size_t)count);
    int64_t total = 0;
    for (int i = 0; i < count; i++) {
        const int32_t *row = h_inputs + (size_t)i * m;
        int n = 0;
        double pin = 1.0;
        for (int q = 0; q < m; q++) {
            if (row[q] < 0) { delete plan; set_error("negative occupation"); return ERR_ARG; }
            n += row[q];
            for (int v = 2; v <= row[q]; v++) pin *= v;
        }
        if (n > SLOS_MAX_PHOTONS) { delete plan; set_error("more than 40 photons"); return ERR_ARG; }
        targets[(size_t)i].assign(row, row + m);
        ids[(size_t)i] = i;
        inv_pin[(size_t)i] = 1.0 / pin;
        t.in_n.push_back(n);
        t.nmax = std::max(t.nmax, n);
    }
    t.w = t.nmax + 1;
    for (int k = 0; k <= t.nmax; k++) {
        const uint64_t c = binom_u64((int64_t)m + k - 1, k);
        if (c == UINT64_MAX || c > (1ull << 40)) {
            delete plan;
            set_error("Fock space too large (more than 2^40 states in a layer)");
            return ERR_RANGE;
        }
        t.counts.push_back(c);
    }
    for (int i = 0; i < count; i++) {
        t.out_off.push_back(total);
        total += (int64_t)t.counts[(size_t)t.in_n[(size_t)i]];
    }
    t.out_off.push_back(total);
    size_t off = 0;
    for (int k = 0; k < t.nmax; k++) {
        t.level_off.push_back(off);
        off += ((sizeof(cd) * (size_t)t.counts[(size_t)k]) + 255) & ~size_t(255);
    }
    t.work_bytes = off;
    tree_decompose(t, 0, -1, std::move(targets), std::move(ids), inv_pin);

    std::vector<uint64_t> T, E;
    build_tables(m, t.nmax, T, E);
    const size_t bytes = sizeof(uint64_t) * T.size();
    hipError_t e = hipMalloc(&t.d_T, 2 * bytes);
    if (e != hipSuccess) { delete plan; set_error(std::string("hipMalloc(slos tables): ") + hipGetErrorString(e)); return ERR_ALLOC; }
    t.d_E = t.d_T + T.size();
    e = hipMemcpy(t.d_T, T.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.d_E, E.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t.d_T);
        delete plan;
        set_error(std::string("hipMemcpy(slos tables): ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    int rc = device_grid_cap(&t.grid_cap);
    if (rc == OK && blk_layout_ok(m, t.nmax)) {
        rc = blk_layout_create(m, t.nmax, blk_default_width(m, t.nmax), &t.blk);
        if (rc == ERR_RANGE) rc = OK;   // the tree's per-state kernel then runs every layer
    }
    if (rc == OK && !t.blk.d_mem && 2 * sizeof(uint64_t) * (size_t)m * (size_t)t.w > 160 * 1024) {
        set_error("SLOS tree: modes and photons too many for the per-state kernel's tables");
        rc = ERR_RANGE;
    }
    if (rc != OK) { blk_layout_destroy(&t.blk); (void)hipFree(t.d_T); delete plan; return rc; }
    *out = plan;
    return OK;
}

extern "C" int pcvl_slos_tree_info(pcvl_slos_tree plan, int *nmax, int64_t *h_out_offsets, size_t *work_bytes,
                                   int64_t *layer_launches)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    const SlosTreeImpl &t = plan->impl;
    if (nmax) *nmax = t.nmax;
    if (h_out_offsets) for (size_t i = 0; i < t.out_off.size(); i++) h_out_offsets[i] = t.out_off[i];
    if (work_bytes) *work_bytes = t.work_bytes;
    if (layer_launches) {
        int64_t n = 0;
        for (const TreeOp &op : t.ops) n += op.k > 0 ? 1 : 0;
        *layer_launches = n;
    }
    return OK;
}

extern "C" int pcvl_slos_tree_execute(pcvl_slos_tree plan, const double *d_U, int mode, void *d_out, void *d_work,
                                      size_t work_bytes, hipStream_t stream)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    PCVL_ARG_CHECK(mode == PCVL_SLOS_COEFS || mode == PCVL_SLOS_PROBS, "mode must be COEFS or PROBS");
    PCVL_ARG_CHECK(d_U && d_out, "null device buffer");
    const SlosTreeImpl &t = plan->impl;
    PCVL_ARG_CHECK(work_bytes >= t.work_bytes && (t.work_bytes == 0 || d_work), "workspace too small");
    const cd *U = reinterpret_cast<const cd *>(d_U);
    char *work = static_cast<char *>(d_work);
    const size_t esz = mode == PCVL_SLOS_PROBS ? sizeof(double) : sizeof(cd);
    const size_t lds = 2 * sizeof(uint64_t) * (size_t)t.m * t.w;
    if (t.nmax > 0) {   // depth-0 buffer holds the vacuum coefficient
        hipLaunchKernelGGL(slos_vacuum_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<double *>(work + t.level_off[0]), 0);
        PCVL_HIP_CHECK(hipGetLastError());
    }
    for (const TreeOp &op : t.ops) {
        if (op.k == 0) {   // input with no photon: vacuum output
            hipLaunchKernelGGL(slos_vacuum_kernel, dim3(1), dim3(64), 0, stream,
                               reinterpret_cast<double *>(static_cast<char *>(d_out) + esz * (size_t)t.out_off[(size_t)op.input]),
                               mode == PCVL_SLOS_PROBS ? 1 : 0);
            PCVL_HIP_CHECK(hipGetLastError());
            continue;
        }
        const cd *parent = reinterpret_cast<const cd *>(work + t.level_off[(size_t)op.k - 1]);
        const uint64_t nc = t.counts[(size_t)op.k];
        const int grid = (int)std::min<uint64_t>((nc + SLOS_THREADS - 1) / SLOS_THREADS, (uint64_t)t.grid_cap);
        void *dst;
        bool probs;
        if (op.to_level) {
            dst = work + t.level_off[(size_t)op.k];
            probs = false;
        } else {
            dst = static_cast<char *>(d_out) + esz * (size_t)t.out_off[(size_t)op.input];
            probs = mode == PCVL_SLOS_PROBS;
        }
        if (t.blk.d_mem) {
            // no sentinel: depth buffers hold exactly their layer, and only
            // the block and per-state kernels run here
            const int rc = blk_layer_launch(t.blk, parent, dst, U, t.m, op.mk, op.k, probs, op.inv_prod_in, false,
                                            stream);
            if (rc != OK) return rc;
        } else if (probs)
            hipLaunchKernelGGL(slos_layer_kernel<1>, dim3(grid), dim3(SLOS_THREADS), lds, stream, parent,
                               t.counts[(size_t)op.k - 1], dst, U, t.m, op.mk, op.k, nc, t.d_T, t.d_E, t.w,
                               op.inv_prod_in);
        else
            hipLaunchKernelGGL(slos_layer_kernel<0>, dim3(grid), dim3(SLOS_THREADS), lds, stream, parent,
                               t.counts[(size_t)op.k - 1], dst, U, t.m, op.mk, op.k, nc, t.d_T, t.d_E, t.w,
                               op.inv_prod_in);
        PCVL_HIP_CHECK(hipGetLastError());
    }
    return OK;
}

extern "C" int pcvl_slos_tree_destroy(pcvl_slos_tree plan)
{
    if (!plan) return OK;
    hipError_t e = hipSuccess;
    if (plan->impl.d_T) e = hipFree(plan->impl.d_T);
    blk_layout_destroy(&plan->impl.blk);
    delete plan;
    if (e != hipSuccess) { set_error(std::string("hipFree(slos tree): ") + hipGetErrorString(e)); return ERR_HIP; }
    return OK;
}


// ------------------------------------------------------------------ ranking

struct RankTables {
    uint64_t *d_T = nullptr;
    int w = 0;
};

static int make_rank_tables(int m, int nmax, RankTables *rt)
{
    std::vector<uint64_t> T, E;
    build_tables(m, nmax, T, E);
    hipError_t e = hipMalloc(&rt->d_T, sizeof(uint64_t) * T.size());
    if (e != hipSuccess) { set_error(std::string("hipMalloc(rank table): ") + hipGetErrorString(e)); return ERR_ALLOC; }
    e = hipMemcpy(rt->d_T, T.data(), sizeof(uint64_t) * T.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(rt->d_T); set_error(std::string("hipMemcpy(rank table): ") + hipGetErrorString(e)); return ERR_HIP; }
    rt->w = nmax + 1;
    return OK;
}

extern "C" int pcvl_fock_rank(const int32_t *d_occ, int64_t count, int m, int nmax, int64_t *d_out, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 256]");
    PCVL_ARG_CHECK(nmax >= 0 && nmax <= 255, "nmax outside [0, 255]");
    PCVL_ARG_CHECK(count >= 0, "count < 0");
    if (count == 0) return OK;
    PCVL_ARG_CHECK(d_occ && d_out, "null device buffer");
    RankTables rt;
    int rc = make_rank_tables(m, nmax, &rt);
    if (rc != OK) return rc;
    int blocks = (int)((count + 255) / 256);
    hipLaunchKernelGGL(fock_rank_kernel, dim3(blocks), dim3(256), 0, stream, d_occ, count, m, nmax, rt.d_T, rt.w, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(rt.d_T);
    if (e != hipSuccess) { set_error(std::string("fock_rank: ") + hipGetErrorString(e)); return ERR_HIP; }
    return OK;
}

extern "C" int pcvl_fock_unrank(const int64_t *d_rank, int64_t count, int m, int n, int32_t *d_occ, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 256]");
    PCVL_ARG_CHECK(n >= 0 && n <= 255, "photon count outside [0, 255]");
    PCVL_ARG_CHECK(count >= 0, "count < 0");
    if (count == 0) return OK;
    PCVL_ARG_CHECK(d_rank && d_occ, "null device buffer");
    uint64_t total = binom_u64((int64_t)m + n - 1, n);
    RankTables rt;
    int rc = make_rank_tables(m, n, &rt);
    if (rc != OK) return rc;
    int blocks = (int)((count + 255) / 256);
    hipLaunchKernelGGL(fock_unrank_kernel, dim3(blocks), dim3(256), 0, stream, d_rank, count, m, n, rt.d_T, rt.w,
                       total, d_occ);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(rt.d_T);
    if (e != hipSuccess) { set_error(std::string("fock_unrank: ") + hipGetErrorString(e)); return ERR_HIP; }
    return OK;
}

extern "C" int64_t pcvl_fock_count(int m, int n)
{
    if (m < 1 || n < 0) return (m == 0 && n == 0) ? 1 : 0;
    uint64_t c = binom_u64((int64_t)m + n - 1, n);
    return c > (uint64_t)INT64_MAX ? -1 : (int64_t)c;
}

extern "C" int pcvl_fock_mask_select(int m, int n, int mask_n, const int32_t *h_cond, int nmask,
                                     const int32_t *h_at_least, int64_t *d_out, int64_t *d_count, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 256]");
    PCVL_ARG_CHECK(n >= 0 && n <= 63, "photon count outside [0, 63]");
    PCVL_ARG_CHECK(mask_n >= 0 && mask_n <= 1024, "mask photon count outside [0, 1024]");
    const int spare = mask_n > n ? mask_n - n : 0;
    PCVL_ARG_CHECK(nmask >= 1 && nmask <= MASK_MAX, "between 1 and 16 masks");
    const size_t mask_lds = sizeof(uint64_t) * (size_t)m * (size_t)(n + 1) + sizeof(int32_t) * (size_t)(nmask + 1) * m;
    PCVL_ARG_CHECK(mask_lds <= 64 * 1024, "mode and photon counts too large for the selection kernel's tables");
    PCVL_ARG_CHECK(h_cond && d_out && d_count, "null buffer");
    const uint64_t total = binom_u64((int64_t)m + n - 1, n);
    PCVL_ARG_CHECK(total != UINT64_MAX && total < (1ull << 34), "Fock space too large (2^34 states or more)");
    const int w = n + 1;
    std::vector<uint64_t> T, E;
    build_tables(m, n, T, E);
    std::vector<int32_t> at((size_t)m, 0);
    if (h_at_least) for (int q = 0; q < m; q++) at[(size_t)q] = h_at_least[q] ? 1 : 0;
    uint32_t absorbs = 0;   // masks with a free or at-least mode
    for (int k = 0; k < nmask; k++)
        for (int q = 0; q < m; q++)
            if (h_cond[(size_t)k * m + q] < 0 || at[(size_t)q]) absorbs |= 1u << k;
    const int64_t nb = (int64_t)((total + MASK_THREADS - 1) / MASK_THREADS);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t b_T = al(sizeof(uint64_t) * T.size());
    const size_t b_c = al(sizeof(int32_t) * (size_t)nmask * m);
    const size_t b_a = al(sizeof(int32_t) * (size_t)m);
    const size_t b_cnt = al(sizeof(int32_t) * (size_t)nb);
    const size_t b_off = al(sizeof(int64_t) * ((size_t)nb + 1));
    const size_t b_fl = al((size_t)total);
    char *blk = nullptr;
    PCVL_HIP_CHECK(ws_alloc(reinterpret_cast<void **>(&blk), b_T + b_c + b_a + b_cnt + b_off + b_fl, stream));
    uint64_t *dT = reinterpret_cast<uint64_t *>(blk);
    int32_t *dc = reinterpret_cast<int32_t *>(blk + b_T);
    int32_t *da = reinterpret_cast<int32_t *>(blk + b_T + b_c);
    int32_t *dcnt = reinterpret_cast<int32_t *>(blk + b_T + b_c + b_a);
    int64_t *doff = reinterpret_cast<int64_t *>(blk + b_T + b_c + b_a + b_cnt);
    uint8_t *dfl = reinterpret_cast<uint8_t *>(blk + b_T + b_c + b_a + b_cnt + b_off);
    int rc = OK;
    hipError_t e = hipMemcpyAsync(dT, T.data(), sizeof(uint64_t) * T.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dc, h_cond, sizeof(int32_t) * (size_t)nmask * m, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(da, at.data(), sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mask_count_kernel, dim3((unsigned)nb), dim3(MASK_THREADS), mask_lds, stream, total, m, n, dT,
                           w, dc, nmask, da, spare, absorbs, dfl, dcnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mask_scan_kernel, dim3(1), dim3(1024), 0, stream, dcnt, nb, doff, d_count);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mask_scatter_kernel, dim3((unsigned)nb), dim3(MASK_THREADS), 0, stream, total, dfl, doff,
                           d_out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) { set_error(std::string("fock mask select: ") + hipGetErrorString(e)); rc = ERR_HIP; }
    // the host tables must outlive their copies
    hipError_t se = hipStreamSynchronize(stream);
    hipError_t fe = hipFreeAsync(blk, stream);
    if (rc == OK && se != hipSuccess) { set_error(std::string("fock mask select: ") + hipGetErrorString(se)); rc = ERR_HIP; }
    if (rc == OK && fe != hipSuccess) { set_error(std::string("hipFreeAsync(mask): ") + hipGetErrorString(fe)); rc = ERR_HIP; }
    return rc;
}

extern "C" int pcvl_gather_f64(const double *d_src, const int64_t *d_idx, int64_t count, double *d_dst,
                               hipStream_t stream)
{
    PCVL_ARG_CHECK(count >= 0, "count < 0");
    if (count == 0) return OK;
    PCVL_ARG_CHECK(d_src && d_idx && d_dst, "null device buffer");
    hipLaunchKernelGGL(gather_f64_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, d_src, d_idx,
                       count, d_dst);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}
