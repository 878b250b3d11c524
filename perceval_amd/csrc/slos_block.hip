// SLOS layer kernel over prefix/suffix blocks (layout in slos_block.h).
//
// Replaces FSMap.compute_slos_layer (perceval/backends/_slos.py:88-102) and,
// on the last layer, xq.all_prob_normalize_output (_slos.py:195-214), as the
// per-state and successor kernels of slos.hip do, with the same parents in
// the same summation order: results are bit-identical to theirs.
//
// Work.  The children of layer k are split into groups by r = photons in
// the w suffix modes; group r holds N(h, k-r) blocks of S(r) = N(w, r)
// children.  A wave is one work item:
//   * S(r) >= 32: one 64-child chunk of one block.  The block's walk
//     (B_k(alpha), PE_h, its occupied prefix modes with their PE_j, and
//     prod alpha!) is precomputed as a 128-byte record that one scalar load
//     brings in; each prefix parent load is then a uniform base plus the
//     lane's beta: one contiguous 1 KiB run per load instruction.
//   * S(r) < 32: 64 consecutive children of the group, across blocks; each
//     lane walks its own alpha (walk tables in LDS) and keeps its list in LDS.
// The suffix parents come from a table per r (every block with the same r
// has the same suffix structure): slot d of beta holds
// (rank(beta - e_j) << 5) | (j - h) for its d-th occupied suffix mode j, or
// all ones past its last one.  Groups run in increasing r, and each group
// gives XCD x a contiguous share of its items (each XCD has its own L2).
#include "common.h"
#include "slos_block.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <new>
#include <type_traits>

namespace pcvl {

constexpr int BLK_THREADS = 256;
constexpr int BLK_WMAX = 16;        // widest suffix
constexpr uint32_t BLK_A_MIN = 32;  // smallest block given whole waves
constexpr int BLK_WG_PER_CU = 8;    // workgroups per CU of the persistent launch
constexpr int BLK_GROUP_LDS = 64;   // groups of a layer cached in LDS (more: read from global memory)

// The walk tables (Th, TE, CTE, f32, fd: consecutive in the layout, from
// off_th): byte offsets from Th and the total, a multiple of 256.
struct BlkTabOffsets {
    uint32_t te, cte, f32, fd, bytes;
};

// C(a, b) for a < 128 by Pascal's rule, saturated at UINT64_MAX, built once
// (the host builders look binomials up millions of times).
struct BlkBinomTable {
    static constexpr int A = 128;
    uint64_t c[A][A];
    BlkBinomTable()
    {
        for (int a = 0; a < A; a++) {
            c[a][0] = 1;
            for (int b = 1; b < A; b++) {
                if (b > a) { c[a][b] = 0; continue; }
                const uint64_t x = c[a - 1][b - 1], y = (b <= a - 1) ? c[a - 1][b] : 0;
                c[a][b] = (x > UINT64_MAX - y) ? UINT64_MAX : x + y;
            }
        }
    }
};

static uint64_t blk_binom(int64_t a, int64_t b)
{
    if (b < 0 || a < 0 || b > a) return 0;
    if (a < BlkBinomTable::A) {
        static const BlkBinomTable table;   // thread-safe one-time construction
        return table.c[a][b];
    }
    if (b > a - b) b = a - b;
    unsigned __int128 r = 1;
    for (int64_t i = 1; i <= b; i++) {
        r = r * (unsigned __int128)(a - b + i) / (unsigned __int128)i;
        if (r > (unsigned __int128)UINT64_MAX) return UINT64_MAX;
    }
    return (uint64_t)r;
}

// N(modes, x): states of x photons in `modes` modes.
static uint64_t nstates(int modes, int x)
{
    if (x < 0) return 0;
    if (modes == 0) return x == 0 ? 1 : 0;
    return blk_binom((int64_t)x + modes - 1, x);
}

// Herald parents of the child at position loc of its sub-layer: entry (i, base)
// gives parent[base + loc] times U[H_i, mk] (su[mfree + i]).  They come first
// in the summation order (herald modes lead the local mode order).
// The group's herald factorial prod a_j! (1 outside masked final layers).
__device__ __forceinline__ double blk_hfact(const BlkGroup &g)
{
    return __builtin_bit_cast(double, ((uint64_t)g.hf_hi << 32) | g.hf_lo);
}

// Children are written once and read again only by the next layer's launch
// (after the whole layer): nontemporal stores keep them from displacing the
// parent lines that the running layer still re-reads from L2.
__device__ __forceinline__ void blk_store(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void blk_store(cd *p, cd v)
{
    __builtin_nontemporal_store(v.r, &p->r);
    __builtin_nontemporal_store(v.i, &p->i);
}

__device__ __forceinline__ void blk_heralds(cd &acc, const cd *__restrict__ parent, const uint2 *__restrict__ her,
                                            uint32_t nher, const cd *su, int mfree, uint32_t loc)
{
    for (uint32_t d0 = 0; d0 < nher; d0 += 4) {
        cd pv[4], uv[4];
#pragma unroll
        for (uint32_t t = 0; t < 4; t++) {
            pv[t] = cd{0.0, 0.0};
            uv[t] = cd{0.0, 0.0};
            if (d0 + t < nher) {
                const uint2 e = her[d0 + t];
                pv[t] = parent[(size_t)e.y + loc];
                uv[t] = su[mfree + (int)e.x];
            }
        }
#pragma unroll
        for (int t = 0; t < 4; t++) cfma_acc(acc, pv[t], uv[t]);   // zero past the list: adds an exact 0
    }
}

// Suffix parents of one child, from the group's slot table: slot d of beta
// holds (rank(beta - e_j) << 5) | (j - h) for its d-th occupied suffix mode j,
// or all ones past its last one.  Slot words are loaded eight at a time, then
// their parents four at a time.
__device__ __forceinline__ void blk_suffix(cd &acc, const char *ps, const uint32_t *__restrict__ ts, uint32_t S,
                                           int smax, const cd *su, int h)
{
#pragma unroll
    for (int h0 = 0; h0 < BLK_WMAX; h0 += 8) {
        if (h0 < smax) {
            uint32_t sw[8];
#pragma unroll
            for (int d = 0; d < 8; d++) sw[d] = h0 + d < smax ? ts[(size_t)(h0 + d) * S] : 0xffffffffu;
#pragma unroll
            for (int d0 = 0; d0 < 8; d0 += 4) {
                if (h0 + d0 < smax) {
                    cd pv[4], uv[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        pv[t] = cd{0.0, 0.0};
                        uv[t] = cd{0.0, 0.0};
                        if (sw[d0 + t] != 0xffffffffu) {
                            pv[t] = *reinterpret_cast<const cd *>(ps + ((sw[d0 + t] >> 5) << 4));
                            uv[t] = su[h + (int)(sw[d0 + t] & 31u)];
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 4; t++) cfma_acc(acc, pv[t], uv[t]);
                }
            }
        }
    }
}

// One wave item of a whole-wave group: one 64-child chunk of one block.  The
// block's walk is its record, fetched with one scalar load; the loads are
// issued in two rounds only: the record, the first eight suffix slot words
// and the factorial product, then the parents of the first eight prefix modes
// and the first eight suffix slots together.  Summation order is unchanged:
// herald modes, prefix modes, suffix modes, each increasing.
template <int MODE>
__device__ __forceinline__ void blk_item_a(
    const BlkGroup &g, uint32_t item, uint32_t lane, const cd *__restrict__ parent, void *__restrict__ out,
    const cd *su, int h, int wsfx, int mfree, const uint2 *__restrict__ her, const double *__restrict__ spf,
    const uint32_t *__restrict__ sslot, const BlkRec *__restrict__ recs, double inv_prod_in)
{
    const uint32_t b = item / g.nch;
    const uint32_t beta = (item - b * g.nch) * 64u + lane;
    const uint32_t *R = recs[g.rec0 + b].w;
    uint32_t rw[26];
#pragma unroll
    for (int i = 0; i < 26; i++) rw[i] = R[i];
    const uint32_t S = g.S;
    if (beta >= S) return;
    const int smax = min((int)g.r, wsfx);
    const uint32_t *ts = sslot + g.tslot + beta;
    uint32_t sw[8];
#pragma unroll
    for (int d = 0; d < 8; d++) sw[d] = d < smax ? ts[(size_t)d * S] : 0xffffffffu;
    const double pfs = MODE == 1 ? spf[g.tpf + beta] : 0.0;
    const uint32_t Bk = rw[0], PEh = rw[1];
    const int cnt = (int)rw[2];
    cd acc = {0.0, 0.0};
    if (g.nher) blk_heralds(acc, parent, her + g.her_off, g.nher, su, mfree, Bk + beta);
    const cd *pp = parent + g.par_base + Bk;
    const char *ps = reinterpret_cast<const char *>(parent + g.par_base + (Bk - PEh));
    const uint32_t boff = beta << 4;
    // round two: first eight prefix parents and first eight suffix parents
    cd pv[8], qv[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        pv[t] = cd{0.0, 0.0};
        if (t < cnt) pv[t] = *reinterpret_cast<const cd *>(reinterpret_cast<const char *>(pp - rw[10 + t]) + boff);
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        qv[t] = cd{0.0, 0.0};
        if (sw[t] != 0xffffffffu) qv[t] = *reinterpret_cast<const cd *>(ps + ((sw[t] >> 5) << 4));
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        if (t < cnt) cfma_acc(acc, pv[t], su[(rw[4 + t / 4] >> (8 * (t % 4))) & 255u]);
    }
#pragma unroll
    for (int t = 8; t < BLK_REC_MAX; t++) {   // more than eight occupied prefix modes
        if (t < cnt) {
            const cd p = *reinterpret_cast<const cd *>(reinterpret_cast<const char *>(pp - rw[10 + t]) + boff);
            cfma_acc(acc, p, su[(rw[4 + t / 4] >> (8 * (t % 4))) & 255u]);
        }
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        if (sw[t] != 0xffffffffu) cfma_acc(acc, qv[t], su[h + (int)(sw[t] & 31u)]);
    }
    if (smax > 8) {   // suffix slots 8..15
#pragma unroll
        for (int t = 0; t < 8; t++) sw[t] = 8 + t < smax ? ts[(size_t)(8 + t) * S] : 0xffffffffu;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            qv[t] = cd{0.0, 0.0};
            if (sw[t] != 0xffffffffu) qv[t] = *reinterpret_cast<const cd *>(ps + ((sw[t] >> 5) << 4));
        }
#pragma unroll
        for (int t = 0; t < 8; t++)
            if (sw[t] != 0xffffffffu) cfma_acc(acc, qv[t], su[h + (int)(sw[t] & 31u)]);
    }
    const size_t o = (size_t)g.out_base + Bk + beta;
    if (MODE == 0) {
        blk_store(reinterpret_cast<cd *>(out) + o, acc);
    } else {
        double pfa;
        __builtin_memcpy(&pfa, &rw[8], sizeof(double));
        blk_store(reinterpret_cast<double *>(out) + o, prob_from_coef(acc, pfa * pfs * blk_hfact(g), inv_prod_in));
    }
}

// One wave item of a group whose blocks are smaller than a wave: 64
// consecutive children of the group, each lane walking its own block (list of
// occupied prefix modes in LDS, per thread).
template <int MODE, bool PFI>
__device__ __forceinline__ void blk_item_b(
    const BlkGroup &g, uint32_t item, uint32_t lane, int tid, const cd *__restrict__ parent, void *__restrict__ out,
    const cd *su, uint32_t *lpe, uint8_t *lmd, int h, int wsfx, int mfree, const uint2 *__restrict__ her,
    const uint32_t *__restrict__ Th, const uint64_t *__restrict__ TE, const uint64_t *__restrict__ CTE, int wt,
    const uint32_t *__restrict__ f32, const double *__restrict__ fd, const double *__restrict__ spf,
    const uint32_t *__restrict__ sslot, double inv_prod_in)
{
    using pf_t = typename std::conditional<PFI, uint32_t, double>::type;
    const int r = (int)g.r, K = (int)g.K;
    const int smax = min(r, wsfx);
    const uint32_t S = g.S;
    const uint32_t i = item * 64u + lane;
    if (i >= g.nb * S) return;
    const uint32_t b = i / S;
    const uint32_t beta = i - b * S;
    pf_t pfp = (pf_t)1;
    uint32_t rem = b, PE = 0, Bk = 0;
    int P = K, cnt = 0;
    for (int q = 0; q < h; q++) {
        int x = 0;
        if (q < h - 1) {
            x = P;
            uint32_t t = Th[q * wt + x];
            while (t > rem) { --x; t = Th[q * wt + x]; }
            rem -= t;
        }
        const uint64_t te = TE[q * wt + x + r];
        Bk += (uint32_t)te;
        if (P > x) {
            lpe[cnt * BLK_THREADS + tid] = PE;
            lmd[cnt * BLK_THREADS + tid] = (uint8_t)q;
            cnt++;
            if (PFI) pfp *= (pf_t)f32[P - x];
            else pfp *= (pf_t)fd[P - x];
        }
        PE += (uint32_t)(te >> 32);
        P = x;
        if (P == 0) {
            const uint64_t ce = CTE[r * (h + 1) + q + 1];
            Bk += (uint32_t)ce;
            PE += (uint32_t)(ce >> 32);
            break;
        }
    }
    const uint32_t PEh = PE;
    const uint32_t cidx = Bk + beta;
    cd acc = {0.0, 0.0};
    if (g.nher) blk_heralds(acc, parent, her + g.her_off, g.nher, su, mfree, cidx);
    const cd *pp = parent + g.par_base;
    for (int d0 = 0; d0 < cnt; d0 += 4) {
        cd pv[4], uv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            pv[t] = cd{0.0, 0.0};
            uv[t] = cd{0.0, 0.0};
            if (d0 + t < cnt) {
                pv[t] = pp[cidx - lpe[(d0 + t) * BLK_THREADS + tid]];
                uv[t] = su[lmd[(d0 + t) * BLK_THREADS + tid]];
            }
        }
#pragma unroll
        for (int t = 0; t < 4; t++) cfma_acc(acc, pv[t], uv[t]);   // zero past the list: adds an exact 0
    }
    if (smax > 0)
        blk_suffix(acc, reinterpret_cast<const char *>(parent + g.par_base + (Bk - PEh)), sslot + g.tslot + beta, S,
                   smax, su, h);
    const size_t o = (size_t)g.out_base + Bk + beta;
    if (MODE == 0) {
        blk_store(reinterpret_cast<cd *>(out) + o, acc);
    } else {
        const double pf = (double)pfp * spf[g.tpf + beta] * blk_hfact(g);
        blk_store(reinterpret_cast<double *>(out) + o, prob_from_coef(acc, pf, inv_prod_in));
    }
}

// Persistent layer kernel: the workgroups of XCD x (blockIdx % 8, dispatch
// being round robin over the 8 XCDs) take the x-th eighth of every group's
// items, groups in increasing suffix photon count, so all XCDs work on the
// same group at a time and each on a contiguous part of it (each XCD has its
// own L2).  Waves take items at a stride of the XCD's waves.
template <int MODE, bool PFI>
__global__ void __launch_bounds__(BLK_THREADS) slos_block_kernel(
    const cd *__restrict__ parent, void *__restrict__ out, const cd *__restrict__ U, int m, int mk, int k,
    uint32_t nchild, int h, int wsfx, int mfree, int mloc, const int32_t *__restrict__ rowmap,
    const BlkGroup *__restrict__ groups, int ngroups, const uint2 *__restrict__ her,
    const char *__restrict__ tabs, BlkTabOffsets to, int wt, const double *__restrict__ spf,
    const uint32_t *__restrict__ sslot, const BlkRec *__restrict__ recs, double inv_prod_in, int sentinel)
{
    extern __shared__ uint64_t s_blk[];
    BlkGroup *sg = reinterpret_cast<BlkGroup *>(s_blk);                   // the layer's groups (up to BLK_GROUP_LDS)
    cd *su = reinterpret_cast<cd *>(sg + BLK_GROUP_LDS);                  // local mode -> U[row, mk]
    char *stab = reinterpret_cast<char *>(su + mloc);                     // walk tables of the per-lane walk
    uint32_t *lpe = reinterpret_cast<uint32_t *>(stab + to.bytes);        // k x THREADS (lists of waves over several blocks)
    uint8_t *lmd = reinterpret_cast<uint8_t *>(lpe + (size_t)k * BLK_THREADS);
    const int tid = threadIdx.x;
    // the per-lane walk reads its tables once per prefix mode, one read
    // depending on the last: from LDS, not through the memory hierarchy
    for (int e = tid; e < (int)(to.bytes / sizeof(uint4)); e += BLK_THREADS)
        reinterpret_cast<uint4 *>(stab)[e] = reinterpret_cast<const uint4 *>(tabs)[e];
    const uint32_t *Th = reinterpret_cast<const uint32_t *>(stab);
    const uint64_t *TE = reinterpret_cast<const uint64_t *>(stab + to.te);
    const uint64_t *CTE = reinterpret_cast<const uint64_t *>(stab + to.cte);
    const uint32_t *f32 = reinterpret_cast<const uint32_t *>(stab + to.f32);
    const double *fd = reinterpret_cast<const double *>(stab + to.fd);
    // rowmap null: local mode e is row e (unmasked layouts), one dependent load less
    for (int e = tid; e < mloc; e += BLK_THREADS) su[e] = U[(size_t)(rowmap ? rowmap[e] : e) * m + mk];
    // the group table is read once per group by every wave: one parallel
    // fetch into LDS instead of a dependent global load per group
    const bool glds = ngroups <= BLK_GROUP_LDS;
    if (glds) {
        const uint4 *gsrc = reinterpret_cast<const uint4 *>(groups);
        uint4 *gdst = reinterpret_cast<uint4 *>(sg);
        for (int e = tid; e < ngroups * (int)(sizeof(BlkGroup) / sizeof(uint4)); e += BLK_THREADS) gdst[e] = gsrc[e];
    }
    if (sentinel && blockIdx.x == 0 && tid == 0) reinterpret_cast<cd *>(out)[nchild] = cd{0.0, 0.0};
    __syncthreads();

    const uint32_t lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t x = blockIdx.x & 7u, j = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
    const uint32_t stride = per_xcd * (BLK_THREADS / 64);
    const uint32_t wl = j * (BLK_THREADS / 64) + wave;   // the wave's index among its XCD's waves
    uint32_t cum = 0;                                      // XCD items of the previous groups, mod stride
    for (int gi = 0; gi < ngroups; gi++) {
        BlkGroup g;
        if (glds) {   // uniform: keep the fields in scalar registers
            const uint32_t *src = reinterpret_cast<const uint32_t *>(sg + gi);
            uint32_t *dst = reinterpret_cast<uint32_t *>(&g);
#pragma unroll
            for (int f = 0; f < (int)(sizeof(BlkGroup) / 4); f++) dst[f] = __builtin_amdgcn_readfirstlane(src[f]);
        } else {
            g = groups[gi];
        }
        const uint32_t share = (g.nitems + 7u) >> 3;
        const uint32_t lo = min(x * share, g.nitems), hi = min(lo + share, g.nitems);
        // the XCD's waves take its items round robin, starting where the
        // previous group stopped: a layer with few items per group gives
        // each wave about one item in all, not one in every group
        const uint32_t first = lo + (wl + stride - cum) % stride;
        cum = (cum + (hi - lo)) % stride;
        if (g.nch != 0) {
            for (uint32_t item = first; item < hi; item += stride)
                blk_item_a<MODE>(g, item, lane, parent, out, su, h, wsfx, mfree, her, spf, sslot, recs, inv_prod_in);
        } else {
            for (uint32_t item = first; item < hi; item += stride)
                blk_item_b<MODE, PFI>(g, item, lane, tid, parent, out, su, lpe, lmd, h, wsfx, mfree, her, Th, TE, CTE,
                                      wt, f32, fd, spf, sslot, inv_prod_in);
        }
    }
}

// ------------------------------------------------------------------ host

bool blk_layout_ok(int m, int nmax)
{
    if (m < 2 || nmax < 1 || nmax > 30) return false;
    for (int k = 0; k <= nmax; k++)
        if (nstates(m, k) >= (1ull << 32)) return false;
    return true;
}

static uint64_t suffix_table_words(int w, int nmax)
{
    uint64_t words = 0;
    for (int r = 0; r <= nmax; r++) words += nstates(w, r) * (2 + (uint64_t)std::min(r, w));
    return words;
}

int blk_default_width(int m, int nmax)
{
    // measured on MI355X at n = 10: m = 24 is fastest at w = 12 (1.46 ms
    // against 1.51-1.77 at 13-16), m = 32 at w = 16 (21.5 ms against 23.2 at 12)
    int w = std::min({16, m - 1, std::max(12, m / 2)});
    while (w > 1 && suffix_table_words(w, nmax) * 4 > (512ull << 20)) w--;
    return std::max(w, 1);
}

// Rank of beta in F(w, |beta|), descending lexicographic order.
static uint64_t local_rank(const std::vector<int> &beta)
{
    const int w = (int)beta.size();
    int R = 0;
    for (int v : beta) R += v;
    uint64_t rk = 0;
    for (int q = 0; q + 1 < w; q++) {
        R -= beta[(size_t)q];
        rk += blk_binom((int64_t)R + w - q - 2, w - q - 1);
    }
    return rk;
}

static size_t blk_lds_bytes(const BlkLayout &lay, int k);

int blk_layout_build_host_ex(int m, int nmax, int w, const std::vector<std::vector<BlkSub>> &subs,
                             const std::vector<uint64_t> &counts, const std::vector<int32_t> &rowmap,
                             BlkLayout *out, std::vector<char> *image, size_t max_records)
{
    static_assert(sizeof(BlkRec) == 128, "BLK_MAX_RECORDS counts 128-byte records");
    if (m < 2 || nmax < 0 || nmax > 30) { set_error("block SLOS layout: needs 2 or more free modes and at most 30 photons"); return ERR_ARG; }
    if (w < 1 || w >= m || w > BLK_WMAX) { set_error("block SLOS layout: suffix width outside [1, min(m-1, 16)]"); return ERR_ARG; }
    if (subs.empty() || counts.size() < subs.size()) { set_error("block SLOS layout: inconsistent layers"); return ERR_ARG; }
    for (uint64_t c : counts)
        if (c >= (1ull << 32)) { set_error("block SLOS layout: layers must hold fewer than 2^32 states"); return ERR_ARG; }
    if ((int)rowmap.size() < m) { set_error("block SLOS layout: row map shorter than the free modes"); return ERR_ARG; }
    BlkLayout lay;
    lay.m = m; lay.nmax = nmax; lay.w = w; lay.h = m - w; lay.wt = nmax + 2;
    lay.nlayers = (int)subs.size() - 1;
    lay.mloc = (int)rowmap.size();
    lay.row_identity = true;
    for (size_t i = 0; i < rowmap.size(); i++) lay.row_identity = lay.row_identity && rowmap[i] == (int32_t)i;
    lay.counts = counts;
    const int h = lay.h, wt = lay.wt;
    // tables
    std::vector<uint32_t> th((size_t)h * wt, 0);
    std::vector<uint64_t> te((size_t)h * wt, 0);
    std::vector<uint64_t> cte((size_t)(nmax + 1) * (h + 1), 0);
    for (int q = 0; q < h; q++)
        for (int x = 0; x < wt; x++) {
            if (q + 1 < h) th[(size_t)q * wt + x] = (uint32_t)blk_binom((int64_t)x + h - q - 2, h - q - 1);
            const uint64_t t = blk_binom((int64_t)x + m - q - 2, m - q - 1);
            const uint64_t e = blk_binom((int64_t)x + m - q - 3, m - q - 2);
            te[(size_t)q * wt + x] = (t & 0xffffffffull) | (e << 32);
        }
    for (int r = 0; r <= nmax; r++) {
        uint64_t st = 0, se = 0;
        for (int q = h - 1; q >= 0; q--) {
            st += te[(size_t)q * wt + r] & 0xffffffffull;
            se += te[(size_t)q * wt + r] >> 32;
            cte[(size_t)r * (h + 1) + q] = (st & 0xffffffffull) | (se << 32);
        }
    }
    std::vector<uint32_t> f32(13, 1);
    std::vector<double> fd((size_t)nmax + 1, 1.0);
    for (int a = 2; a <= 12; a++) f32[(size_t)a] = f32[(size_t)a - 1] * (uint32_t)a;
    for (int a = 2; a <= nmax; a++) fd[(size_t)a] = fd[(size_t)a - 1] * (double)a;
    // suffix tables: per r, pf[beta] and slot[d][beta]
    std::vector<double> spf;
    std::vector<uint32_t> sslot;
    std::vector<uint32_t> tpf((size_t)nmax + 1), tslot((size_t)nmax + 1);
    for (int r = 0; r <= nmax; r++) {
        const uint64_t S = nstates(w, r);
        const int ns = std::min(r, w);
        tpf[(size_t)r] = (uint32_t)spf.size();
        tslot[(size_t)r] = (uint32_t)sslot.size();
        const size_t p0 = spf.size(), s0 = sslot.size();
        spf.resize(p0 + S);
        sslot.resize(s0 + S * (size_t)ns, 0xffffffffu);
        std::vector<int> beta((size_t)w, 0), after((size_t)w + 1, 0);
        beta[0] = r;
        for (uint64_t i = 0; i < S; i++) {
            double pf = 1.0;
            int d = 0;
            // beta has rank i; taking a photon from mode j lowers the photons
            // after each mode q < j by one, which lowers the rank by
            // sum_{q<j} C(after_q + w - q - 3, w - q - 2) (Pascal's rule on
            // each term of the rank sum)
            for (int q = w - 1; q >= 0; q--) after[(size_t)q] = after[(size_t)q + 1] + beta[(size_t)q];
            uint64_t drop = 0;
            for (int j = 0; j < w; j++) {
                for (int v = 2; v <= beta[(size_t)j]; v++) pf *= v;
                if (beta[(size_t)j] > 0) {
                    const uint64_t lp = i - drop;
                    sslot[s0 + (size_t)d * S + i] = (uint32_t)(lp << 5) | (uint32_t)j;
                    d++;
                }
                if (j + 2 <= w) drop += blk_binom((int64_t)after[(size_t)j + 1] + w - j - 3, w - j - 2);
            }
            spf[p0 + i] = pf;
            // next state in descending lexicographic order
            if (i + 1 < S) {
                const int tail = beta[(size_t)w - 1];
                beta[(size_t)w - 1] = 0;
                int t = w - 2;
                while (beta[(size_t)t] == 0) t--;
                beta[(size_t)t]--;
                beta[(size_t)t + 1] = tail + 1;
            }
        }
    }
    // groups per layer: each sub-layer, suffix photons r = 0..kf; the blocks
    // of whole-wave groups get their walk precomputed (records)
    lay.cmax = std::max(1, std::min(h, nmax));
    std::vector<BlkRec> recs;
    auto add_records = [&](int K, int r) {
        std::vector<int> alpha((size_t)h, 0);
        alpha[0] = K;
        const uint64_t nb = nstates(h, K);
        std::vector<int> after((size_t)h + 1, 0);
        for (uint64_t i = 0; i < nb; i++) {
            BlkRec rec{};
            uint64_t Bk = 0, PE = 0;
            double pf = 1.0;
            for (int q = h - 1; q >= 0; q--) after[(size_t)q] = after[(size_t)q + 1] + alpha[(size_t)q];
            int cnt = 0;
            for (int q = 0; q < h; q++) {
                const uint64_t tev = te[(size_t)q * wt + r + after[(size_t)q + 1]];
                Bk += tev & 0xffffffffull;
                if (alpha[(size_t)q] > 0) {
                    rec.w[4 + cnt / 4] |= (uint32_t)q << (8 * (cnt % 4));
                    rec.w[10 + cnt] = (uint32_t)PE;
                    cnt++;
                    for (int v = 2; v <= alpha[(size_t)q]; v++) pf *= v;
                }
                PE += tev >> 32;
            }
            rec.w[0] = (uint32_t)Bk;
            rec.w[1] = (uint32_t)PE;
            rec.w[2] = (uint32_t)cnt;
            std::memcpy(&rec.w[8], &pf, sizeof(double));
            recs.push_back(rec);
            if (i + 1 < nb) {   // next alpha in descending lexicographic order
                const int tail = alpha[(size_t)h - 1];
                alpha[(size_t)h - 1] = 0;
                int t = h - 2;
                while (alpha[(size_t)t] == 0) t--;
                alpha[(size_t)t]--;
                alpha[(size_t)t + 1] = tail + 1;
            }
        }
    };
    // records hold at most 16 occupied prefix modes and mode indices below 256
    const bool use_records = lay.cmax <= BLK_REC_MAX && h <= 256;
    std::vector<uint32_t> hert;   // (herald index, parent sub-layer offset) pairs
    std::map<std::pair<int, int>, uint32_t> rec_of;   // (K, r) -> first record
    lay.groups.assign((size_t)lay.nlayers + 1, {});
    lay.items.assign((size_t)lay.nlayers + 1, 0);
    for (int k = 1; k <= lay.nlayers; k++) {
        uint64_t items = 0;
        for (const BlkSub &sub : subs[(size_t)k]) {
            if (sub.kf < 0 || sub.kf > nmax) { set_error("block SLOS layout: sub-layer photons outside the tables"); return ERR_ARG; }
            const uint32_t her_off = (uint32_t)(hert.size() / 2);
            for (const auto &e : sub.her) {
                if ((int)e.first >= lay.mloc - m) {
                    set_error("block SLOS layout: herald index outside the row map");
                    return ERR_ARG;
                }
                hert.push_back(e.first);
                hert.push_back(e.second);
            }
            for (int r = 0; r <= sub.kf; r++) {
                BlkGroup g{};
                std::memcpy(&g.hf_lo, &sub.hfact, sizeof(double));   // hf_lo, hf_hi
                g.nb = (uint32_t)nstates(h, sub.kf - r);
                g.S = (uint32_t)nstates(w, r);
                // a block's record depends on (K, r) only: sub-layers share
                // them.  Records take at most max_records entries (256 MiB
                // by default); groups past that walk per lane (regime B).
                bool rec = false;
                if (g.S >= BLK_A_MIN && use_records) {
                    const auto key = std::make_pair(sub.kf - r, r);
                    auto rit = rec_of.find(key);
                    if (rit == rec_of.end() && recs.size() + g.nb <= max_records) {
                        rit = rec_of.emplace(key, (uint32_t)recs.size()).first;
                        add_records(sub.kf - r, r);
                    }
                    if (rit != rec_of.end()) {
                        rec = true;
                        g.rec0 = rit->second;
                    }
                }
                if (rec) {
                    g.nch = (g.S + 63) / 64;
                    g.nitems = g.nb * g.nch;
                } else {
                    g.nch = 0;
                    g.nitems = (uint32_t)(((uint64_t)g.nb * g.S + 63) / 64);
                }
                g.tpf = tpf[(size_t)r];
                g.tslot = tslot[(size_t)r];
                g.r = (uint32_t)r;
                g.K = (uint32_t)(sub.kf - r);
                g.out_base = sub.out_base;
                g.par_base = sub.par_base;
                g.her_off = her_off;
                g.nher = (uint32_t)sub.her.size();
                items += g.nitems;
                lay.groups[(size_t)k].push_back(g);
            }
        }
        lay.items[(size_t)k] = items;
    }
    // one device allocation
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t off = 0;
    lay.off_th = off; off += al(sizeof(uint32_t) * th.size());
    lay.off_te = off; off += al(sizeof(uint64_t) * te.size());
    lay.off_cte = off; off += al(sizeof(uint64_t) * cte.size());
    lay.off_f32 = off; off += al(sizeof(uint32_t) * f32.size());
    lay.off_fd = off; off += al(sizeof(double) * fd.size());
    lay.off_row = off; off += al(sizeof(int32_t) * rowmap.size());
    lay.off_her = off; off += al(sizeof(uint32_t) * std::max<size_t>(hert.size(), 2));
    lay.group_off.assign((size_t)lay.nlayers + 1, 0);
    for (int k = 1; k <= lay.nlayers; k++) {
        lay.group_off[(size_t)k] = off;
        off += al(sizeof(BlkGroup) * lay.groups[(size_t)k].size());
    }
    lay.off_pf = off; off += al(sizeof(double) * spf.size());
    lay.off_slot = off; off += al(sizeof(uint32_t) * std::max<size_t>(sslot.size(), 1));
    lay.nrec = recs.size();
    lay.off_rec = off; off += al(sizeof(BlkRec) * std::max<size_t>(recs.size(), 1));
    lay.bytes = off;
    if (blk_lds_bytes(lay, lay.nlayers) > 160 * 1024) {
        set_error("block SLOS layout: the walk tables and lists outgrow the LDS (too many modes and photons)");
        return ERR_RANGE;
    }
    std::vector<char> &host = *image;
    host.assign(off, 0);
    std::memcpy(host.data() + lay.off_th, th.data(), sizeof(uint32_t) * th.size());
    std::memcpy(host.data() + lay.off_te, te.data(), sizeof(uint64_t) * te.size());
    std::memcpy(host.data() + lay.off_cte, cte.data(), sizeof(uint64_t) * cte.size());
    std::memcpy(host.data() + lay.off_f32, f32.data(), sizeof(uint32_t) * f32.size());
    std::memcpy(host.data() + lay.off_fd, fd.data(), sizeof(double) * fd.size());
    std::memcpy(host.data() + lay.off_row, rowmap.data(), sizeof(int32_t) * rowmap.size());
    if (!hert.empty()) std::memcpy(host.data() + lay.off_her, hert.data(), sizeof(uint32_t) * hert.size());
    for (int k = 1; k <= lay.nlayers; k++)
        std::memcpy(host.data() + lay.group_off[(size_t)k], lay.groups[(size_t)k].data(),
                    sizeof(BlkGroup) * lay.groups[(size_t)k].size());
    std::memcpy(host.data() + lay.off_pf, spf.data(), sizeof(double) * spf.size());
    if (!sslot.empty()) std::memcpy(host.data() + lay.off_slot, sslot.data(), sizeof(uint32_t) * sslot.size());
    if (!recs.empty()) std::memcpy(host.data() + lay.off_rec, recs.data(), sizeof(BlkRec) * recs.size());
    *out = std::move(lay);
    return OK;
}

int blk_layout_build_host(int m, int nmax, int w, BlkLayout *out, std::vector<char> *image, size_t max_records)
{
    if (!blk_layout_ok(m, nmax)) { set_error("block SLOS layout: layers must hold fewer than 2^32 states"); return ERR_ARG; }
    std::vector<std::vector<BlkSub>> subs((size_t)nmax + 1);
    std::vector<uint64_t> counts;
    for (int k = 0; k <= nmax; k++) {
        counts.push_back(nstates(m, k));
        if (k >= 1) {
            BlkSub sub;
            sub.kf = k;
            subs[(size_t)k].push_back(sub);
        }
    }
    std::vector<int32_t> rowmap((size_t)m);
    for (int i = 0; i < m; i++) rowmap[(size_t)i] = i;
    return blk_layout_build_host_ex(m, nmax, w, subs, counts, rowmap, out, image, max_records);
}

// LDS of layer k's launch: the U column, then the per-thread lists of waves
// whose lanes walk their own blocks.
static size_t blk_lds_bytes(const BlkLayout &lay, int k)
{
    return sizeof(BlkGroup) * BLK_GROUP_LDS + sizeof(cd) * lay.mloc + (lay.off_row - lay.off_th) +
           (sizeof(uint32_t) + 1) * (size_t)k * BLK_THREADS;
}

int blk_masked_layers(int m, int n, const int32_t *cond, BlkMasked *out)
{
    return blk_masked_layers_ex(m, n, n, 1, cond, nullptr, out);
}

int blk_masked_layers_ex(int m, int n, int n_mask, int nrows, const int32_t *conds, const int32_t *at_least,
                         BlkMasked *out)
{
    if (m < 1 || n < 0 || n_mask < n || nrows < 1 || !conds) { set_error("masked SLOS: bad mask arguments"); return ERR_ARG; }
    BlkMasked mk;
    std::vector<int32_t> H;
    for (int j = 0; j < m; j++) {
        bool constrained = false;
        for (int r = 0; r < nrows; r++) constrained = constrained || conds[(size_t)r * m + j] >= 0;
        if (constrained) H.push_back(j);
        else mk.rowmap.push_back(j);
    }
    mk.m_free = (int)mk.rowmap.size();
    mk.nher = (int)H.size();
    // a free mode takes any photons beyond a row's missing ones, which makes
    // "missing <= photons to come" the exact test for a row still reachable
    if (mk.m_free < 1) { set_error("masked SLOS: needs a mode that no mask row constrains"); return ERR_ARG; }
    for (int32_t j : H) mk.rowmap.push_back(j);
    const int nh = (int)H.size();
    // per row, over the herald modes: the count asked (-1 any) and whether it is a floor
    std::vector<int32_t> rc((size_t)nrows * nh), rl((size_t)nrows * nh);
    for (int r = 0; r < nrows; r++)
        for (int i = 0; i < nh; i++) {
            rc[(size_t)r * nh + i] = conds[(size_t)r * m + H[(size_t)i]];
            rl[(size_t)r * nh + i] = at_least && at_least[H[(size_t)i]] ? 1 : 0;
        }
    // largest useful value of each herald count
    std::vector<int32_t> cap((size_t)nh, 0);
    for (int i = 0; i < nh; i++)
        for (int r = 0; r < nrows; r++) {
            const int32_t c = rc[(size_t)r * nh + i];
            cap[(size_t)i] = std::max(cap[(size_t)i], (c < 0 || rl[(size_t)r * nh + i]) ? n : std::min(c, n));
        }
    // the fewest photons configuration a still misses for some row (-1: a has
    // more than every row allows on some exact mode)
    auto dmin_of = [&](const std::vector<int32_t> &a) {
        int best = -1;
        for (int r = 0; r < nrows; r++) {
            int d = 0;
            bool fits = true;
            for (int i = 0; i < nh && fits; i++) {
                const int32_t c = rc[(size_t)r * nh + i];
                if (c < 0) continue;
                if (rl[(size_t)r * nh + i]) d += std::max(0, c - a[(size_t)i]);
                else if (a[(size_t)i] > c) fits = false;
                else d += c - a[(size_t)i];
            }
            if (fits && (best < 0 || d < best)) best = d;
        }
        return best;
    };
    // every configuration a <= cap with |a| <= n, in descending lexicographic
    // order: the successor takes one from the last nonzero count and refills
    // the counts after it greedily.  The count is bounded so that the tables
    // below stay within about 2^28 steps of host work.
    const size_t max_conf = std::min<size_t>(1000000, ((size_t)1 << 28) / ((size_t)std::max(1, nh) * (nrows + n + 2)));
    std::vector<std::vector<int32_t>> all;
    std::vector<int> asum, dmin;
    {
        std::vector<int32_t> a((size_t)nh, 0);
        int left = n;
        for (int i = 0; i < nh; i++) { a[(size_t)i] = std::min(cap[(size_t)i], left); left -= a[(size_t)i]; }
        while (true) {
            int sa = 0;
            for (int32_t v : a) sa += v;
            all.push_back(a);
            asum.push_back(sa);
            dmin.push_back(dmin_of(a));
            if (all.size() > max_conf) {
                set_error("masked SLOS: more than " + std::to_string(max_conf) + " herald configurations");
                return ERR_RANGE;
            }
            int i = nh - 1;
            while (i >= 0 && a[(size_t)i] == 0) i--;
            if (i < 0) break;
            a[(size_t)i]--;
            left = n - (sa - 1);
            for (int t = i + 1; t < nh; t++) { a[(size_t)t] = std::min(cap[(size_t)t], left); left -= a[(size_t)t]; }
        }
    }
    const size_t NA = all.size();
    // down[t * nh + i]: the index of all[t] - e_i (every such vector is itself a configuration)
    std::vector<int32_t> down(NA * (size_t)nh, -1);
    {
        struct VecHash {
            size_t operator()(const std::vector<int32_t> &v) const {
                uint64_t h = 1469598103934665603ull;
                for (int32_t x : v) h = (h ^ (uint32_t)x) * 1099511628211ull;
                return (size_t)h;
            }
        };
        std::unordered_map<std::vector<int32_t>, int32_t, VecHash> idx_of;
        idx_of.reserve(NA * 2);
        for (size_t t = 0; t < NA; t++) idx_of.emplace(all[t], (int32_t)t);
        for (size_t t = 0; t < NA; t++)
            for (int i = 0; i < nh; i++) {
                if (all[t][(size_t)i] == 0) continue;
                auto b = all[t];
                b[(size_t)i]--;
                auto it = idx_of.find(b);
                if (it == idx_of.end()) { set_error("masked SLOS: internal: configuration missing"); return ERR_RANGE; }
                down[t * (size_t)nh + i] = it->second;
            }
    }
    mk.subs.assign((size_t)n + 1, {});
    mk.counts.assign((size_t)n + 1, 0);
    // pos[t]: offset of configuration t's sub-layer in the layer (-1: not kept);
    // layer k keeps a when |a| <= k and some row still fits in the n_mask - k
    // photons to come (n_mask > n: the output layer itself holds the states
    // that photons from elsewhere can complete)
    std::vector<int64_t> pos_prev(NA, -1), pos(NA, -1);
    for (int k = 0; k <= n; k++) {
        uint64_t off = 0;
        for (size_t t = 0; t < NA; t++) {
            pos[t] = -1;
            if (asum[t] > k || dmin[t] < 0 || dmin[t] > n_mask - k) continue;
            pos[t] = (int64_t)off;
            mk.nmax_free = std::max(mk.nmax_free, k - asum[t]);
            off += nstates(mk.m_free, k - asum[t]);
            if (off >= (1ull << 32)) { set_error("masked SLOS: a layer holds 2^32 states or more"); return ERR_RANGE; }
        }
        mk.counts[(size_t)k] = off;
        if (k > 0)
            for (size_t t = 0; t < NA; t++) {
                if (pos[t] < 0) continue;
                BlkSub sub;
                sub.out_base = (uint32_t)pos[t];
                sub.kf = k - asum[t];
                if (sub.kf >= 1) {
                    if (pos_prev[t] < 0) { set_error("masked SLOS: internal: parent sub-layer missing"); return ERR_RANGE; }
                    sub.par_base = (uint32_t)pos_prev[t];
                }
                for (int i = 0; i < nh; i++) {
                    if (all[t][(size_t)i] == 0) continue;
                    const int32_t b = down[t * (size_t)nh + i];
                    if (pos_prev[(size_t)b] < 0) { set_error("masked SLOS: internal: herald parent missing"); return ERR_RANGE; }
                    sub.her.push_back({(uint32_t)i, (uint32_t)pos_prev[(size_t)b]});
                }
                if (k == n)
                    for (int32_t v : all[t])
                        for (int f = 2; f <= v; f++) sub.hfact *= f;
                mk.subs[(size_t)k].push_back(std::move(sub));
            }
        if (k < n) std::swap(pos_prev, pos);
    }
    if (mk.counts[(size_t)n] == 0) {
        mk.empty = true;
        *out = std::move(mk);
        return OK;
    }
    // several final sub-layers: the FSArray order of the kept states, by a
    // walk over the original modes (each mode's count from the photons left
    // down to 0) that keeps only the final configurations still reachable:
    // the same herald values so far, their remaining herald photons within
    // the photons left (exactly, once no free mode is left)
    std::vector<std::pair<std::vector<int32_t>, uint32_t>> finals;   // (configuration, sub-layer offset)
    for (size_t t = 0; t < NA; t++)
        if (pos[t] >= 0) finals.push_back({all[t], (uint32_t)pos[t]});
    if (finals.size() > 1) {
        if (mk.counts[(size_t)n] > (1ull << 28)) { set_error("masked SLOS: more than 2^28 kept outputs"); return ERR_RANGE; }
        // the walk below filters the candidate configurations at every node
        if ((double)mk.counts[(size_t)n] * (double)(finals.size() + (size_t)m) > 0x1p34) {
            set_error("masked SLOS: too many outputs times final herald configurations to put in FSArray order");
            return ERR_RANGE;
        }
        std::vector<int32_t> hidx((size_t)m, -1), fidx((size_t)m, -1);
        for (int i = 0; i < nh; i++) hidx[(size_t)H[(size_t)i]] = i;
        for (int e = 0; e < mk.m_free; e++) fidx[(size_t)mk.rowmap[(size_t)e]] = e;
        const size_t nf = finals.size();
        // suf[t * (m + 1) + q]: herald photons of final configuration t on modes q..m-1
        std::vector<int> suf(nf * ((size_t)m + 1), 0), free_after((size_t)m + 1, 0);
        for (size_t t = 0; t < nf; t++)
            for (int q = m - 1; q >= 0; q--)
                suf[t * ((size_t)m + 1) + q] = suf[t * ((size_t)m + 1) + q + 1] +
                                               (hidx[(size_t)q] >= 0 ? finals[t].first[(size_t)hidx[(size_t)q]] : 0);
        for (int q = m - 1; q >= 0; q--) free_after[(size_t)q] = free_after[(size_t)q + 1] + (fidx[(size_t)q] >= 0);
        std::vector<int32_t> f((size_t)mk.m_free, 0);
        std::vector<std::vector<int>> cands((size_t)m + 1);
        cands[0].resize(nf);
        for (size_t t = 0; t < nf; t++) cands[0][t] = (int)t;
        mk.perm.reserve((size_t)mk.counts[(size_t)n]);
        std::vector<int> val((size_t)m, 0), left((size_t)m + 1, 0);
        left[0] = n;
        int q = 0;
        bool start = true;
        while (q >= 0) {
            if (start) {
                val[(size_t)q] = left[(size_t)q];
                start = false;
            } else {
                if (q == m - 1 || val[(size_t)q] == 0) { q--; continue; }
                val[(size_t)q]--;
            }
            const int v = val[(size_t)q];
            if (q == m - 1 && v != left[(size_t)q]) { q--; continue; }
            const int rest = left[(size_t)q] - v;
            auto &next = cands[(size_t)q + 1];
            next.clear();
            for (int t : cands[(size_t)q]) {
                if (hidx[(size_t)q] >= 0 && finals[(size_t)t].first[(size_t)hidx[(size_t)q]] != v) continue;
                const int need = suf[(size_t)t * ((size_t)m + 1) + q + 1];
                if (need > rest || (free_after[(size_t)q + 1] == 0 && need != rest)) continue;
                next.push_back(t);
            }
            if (next.empty()) continue;   // the next smaller value
            if (fidx[(size_t)q] >= 0) f[(size_t)fidx[(size_t)q]] = v;
            if (q == m - 1) {
                // complete: the herald part is exactly one final configuration
                const auto &fc = finals[(size_t)next[0]];
                std::vector<int> fv(f.begin(), f.end());
                mk.perm.push_back(fc.second + (uint32_t)local_rank(fv));
                continue;   // the last mode has a single value: this pops
            }
            left[(size_t)q + 1] = rest;
            q++;
            start = true;
        }
        if (mk.perm.size() != mk.counts[(size_t)n]) { set_error("masked SLOS: internal: output order walk"); return ERR_RANGE; }
    }
    *out = std::move(mk);
    return OK;
}

int blk_layout_upload(BlkLayout *lay, const std::vector<char> &image)
{
    // The persistent launch needs every workgroup resident at once (each
    // takes an equal share of the items): size it by the kernel's occupancy
    // at the largest LDS any layer of the layout asks for.
    int dev = 0, cus = 0, per_cu = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds = blk_lds_bytes(*lay, lay->nlayers);
    int worst = 1 << 30;
    for (const void *kp : {reinterpret_cast<const void *>(&slos_block_kernel<0, true>),
                           reinterpret_cast<const void *>(&slos_block_kernel<0, false>),
                           reinterpret_cast<const void *>(&slos_block_kernel<1, true>),
                           reinterpret_cast<const void *>(&slos_block_kernel<1, false>)}) {
        PCVL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, BLK_THREADS, lds));
        worst = std::min(worst, per_cu);
    }
    worst = std::max(1, std::min(worst, BLK_WG_PER_CU));
    lay->grid_cap = std::max(8, cus * worst / 8 * 8);
    hipError_t e = hipMalloc(&lay->d_mem, lay->bytes);
    if (e != hipSuccess) {
        lay->d_mem = nullptr;
        set_error(std::string("hipMalloc(block SLOS tables): ") + hipGetErrorString(e));
        return ERR_ALLOC;
    }
    e = hipMemcpy(lay->d_mem, image.data(), lay->bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(lay->d_mem);
        lay->d_mem = nullptr;
        set_error(std::string("hipMemcpy(block SLOS tables): ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    return OK;
}

int blk_layout_create(int m, int nmax, int w, BlkLayout *out)
{
    BlkLayout lay;
    std::vector<char> host;
    int rc = blk_layout_build_host(m, nmax, w, &lay, &host);
    if (rc == OK) rc = blk_layout_upload(&lay, host);
    if (rc != OK) return rc;
    *out = std::move(lay);
    return OK;
}

void blk_layout_destroy(BlkLayout *lay)
{
    if (lay->d_mem) (void)hipFree(lay->d_mem);
    lay->d_mem = nullptr;
}

int blk_layer_launch(const BlkLayout &lay, const cd *parent, void *out, const cd *U, int m, int mk, int k, bool probs,
                     double inv_prod_in, bool sentinel, hipStream_t stream)
{
    if (k < 1 || k > lay.nlayers || !lay.d_mem) { set_error("block SLOS: layer outside the layout"); return ERR_ARG; }
    if (lay.items[(size_t)k] == 0) return OK;
    // persistent grid: a multiple of 8 workgroups (the XCD split), at most grid_cap
    const uint64_t want = (lay.items[(size_t)k] + (BLK_THREADS / 64) - 1) / (BLK_THREADS / 64);
    const uint32_t grid = (uint32_t)((std::min<uint64_t>(want, (uint64_t)lay.grid_cap) + 7) / 8 * 8);
    const char *base = lay.d_mem;
    const BlkGroup *groups = reinterpret_cast<const BlkGroup *>(base + lay.group_off[(size_t)k]);
    const uint32_t nchild = (uint32_t)lay.counts[(size_t)k];
    const size_t lds = blk_lds_bytes(lay, k);
    const bool pfi = lay.nmax <= 12;
    const BlkTabOffsets to = {(uint32_t)(lay.off_te - lay.off_th), (uint32_t)(lay.off_cte - lay.off_th),
                              (uint32_t)(lay.off_f32 - lay.off_th), (uint32_t)(lay.off_fd - lay.off_th),
                              (uint32_t)(lay.off_row - lay.off_th)};
    auto kp = probs ? (pfi ? slos_block_kernel<1, true> : slos_block_kernel<1, false>)
                    : (pfi ? slos_block_kernel<0, true> : slos_block_kernel<0, false>);
    hipLaunchKernelGGL(kp, dim3(grid), dim3(BLK_THREADS), lds, stream, parent, out, U, m, mk, k, nchild,
                       lay.h, lay.w, lay.m, lay.mloc,
                       lay.row_identity ? nullptr : reinterpret_cast<const int32_t *>(base + lay.off_row), groups,
                       (int)lay.groups[(size_t)k].size(), reinterpret_cast<const uint2 *>(base + lay.off_her),
                       base + lay.off_th, to, lay.wt,
                       reinterpret_cast<const double *>(base + lay.off_pf),
                       reinterpret_cast<const uint32_t *>(base + lay.off_slot),
                       reinterpret_cast<const BlkRec *>(base + lay.off_rec), inv_prod_in, sentinel ? 1 : 0);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

} // namespace pcvl
