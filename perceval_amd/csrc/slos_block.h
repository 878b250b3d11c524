// Prefix/suffix block layout of the SLOS layers (internal to the library).
//
// A layer F(m, k) in descending lexicographic order is cut at mode h = m - w:
// a child c = (alpha, beta) has alpha on modes 0..h-1 and beta on the w
// suffix modes, |beta| = r.  The children sharing alpha form one contiguous
// block of S(r) = N(w, r) states, ordered by the rank of beta in F(w, r):
//     rank_k(c) = B_k(alpha) + rank(beta).
// Parents, in increasing mode order (the summation order of
// FSMap.compute_slos_layer, perceval/backends/_slos.py:88-102):
//   prefix mode j (alpha_j > 0): rank_k(c) - PE_j(alpha)   (the same beta)
//   suffix mode j (beta_j > 0):  B_k(alpha) - PE_h(alpha) + rank(beta - e_j)
// with PE_j(alpha) = sum_{q<j} E_q(R_{q+1}) as in the per-state kernels.  The
// prefix part is the same for every child of a block and the suffix part the
// same for every block with the same r, so one walk per block and one table
// per r (shared by every layer) replace the per-child unranking.
//
// Masked layers (heralds).  With a mask fixing the photon count c_j of some
// modes H (the masked FSArray of _slos.py:161-164), layer k only needs the
// states whose herald part a = s_H can still reach c: a <= c and
// k - |a| <= n - |c| (partial match).  Every parent of such a state is again
// one.  With the modes reordered herald modes first, each needed layer is a
// concatenation, over the admissible a in descending lexicographic order, of
// full free-mode spaces F(m_F, k - |a|) ("sub-layers"); a child (a, f) takes
// its parents (a - e_j, f) from sub-layer a - e_j of layer k-1 at the same
// position, and (a, f - e_j) from sub-layer a of layer k-1 by the block
// recursion over the free modes.  The unmasked layer is the one sub-layer
// with no herald modes.
#pragma once

#include "common.h"

#include <utility>
#include <vector>

namespace pcvl {

// One group of a layer: the blocks of one sub-layer with r photons in the
// suffix modes.  64 bytes.
struct BlkGroup {
    uint32_t nitems;    // wave items: (block, 64-child chunk), or 64 consecutive children
    uint32_t nb;        // blocks: N(h, K)
    uint32_t S;         // children per block: N(w, r)
    uint32_t nch;       // 64-child chunks per block (one wave each); 0: waves span several blocks
    uint32_t tpf;       // offset of r's factorial products in the suffix pf table
    uint32_t tslot;     // offset of r's slot words in the suffix slot table
    uint32_t r;         // photons in the suffix modes
    uint32_t K;         // photons in the prefix (free) modes
    uint32_t out_base;  // offset of the group's sub-layer in layer k
    uint32_t par_base;  // offset of the same herald configuration's sub-layer in layer k - 1
    uint32_t her_off;   // herald parents: entries [her_off, her_off + nher) of the herald table
    uint32_t nher;
    uint32_t rec0;      // record of the group's block 0 (groups with nch != 0)
    uint32_t hf_lo, hf_hi;   // the sub-layer's prod a_j! (a double), factor of its probabilities
    uint32_t pad0;
};
static_assert(sizeof(BlkGroup) == 64, "BlkGroup is one 64-byte row");

// Precomputed walk of one block (groups with whole waves per block), 128
// bytes so that one scalar fetch brings all of it:
//   w[0] B_k(alpha), w[1] PE_h(alpha), w[2] occupied prefix modes (<= 16),
//   w[4..7] those modes (one byte each, increasing), w[8..9] prod alpha!
//   (double), w[10..25] PE_j(alpha) of each occupied prefix mode j.
constexpr int BLK_REC_WORDS = 32;
constexpr int BLK_REC_MAX = 16;
struct BlkRec {
    uint32_t w[BLK_REC_WORDS];
};

// One sub-layer of layer k: herald configuration a.
struct BlkSub {
    uint32_t out_base = 0;   // offset in layer k
    uint32_t par_base = 0;   // offset of sub-layer a in layer k - 1 (used when kf >= 1)
    int kf = 0;              // photons in the free modes: k - |a|
    // herald parents, in herald mode order: (herald index i, offset of
    // sub-layer a - e_{H_i} in layer k - 1)
    std::vector<std::pair<uint32_t, uint32_t>> her;
    double hfact = 1.0;      // prod a_j! (probabilities of the final layer)
};

struct BlkLayout {
    int m = 0;                // free modes (all modes when unmasked)
    int nmax = 0;             // most photons in the free modes
    int w = 0, h = 0, wt = 0;
    int nlayers = 0;          // layers 1..nlayers
    int mloc = 0;             // U rows used: free modes then herald modes
    bool row_identity = false; // local mode i is U row i (no herald modes)
    std::vector<uint64_t> counts;                // states of layer k (compact when masked), k = 0..nlayers
    std::vector<std::vector<BlkGroup>> groups;   // per layer k (index k)
    std::vector<uint64_t> items;                 // wave items per layer
    int grid_cap = 0;                            // workgroups of the persistent launch (a multiple of 8)
    std::vector<size_t> group_off;               // byte offset of layer k's groups in d_mem
    size_t off_th = 0, off_te = 0, off_cte = 0, off_f32 = 0, off_fd = 0, off_pf = 0, off_slot = 0, off_row = 0,
           off_her = 0, off_rec = 0;
    int cmax = 1;             // most occupied prefix modes of a block: min(h, nmax)
    uint64_t nrec = 0;        // block records
    char *d_mem = nullptr;
    size_t bytes = 0;
};

// Suffix width used when the caller does not choose one (0): min(16, m - 1,
// max(12, m / 2)), narrowed until the suffix tables stay under 512 MiB.
int blk_default_width(int m, int nmax);

// The layout of layers 1..n of the full space F(m, k), suffix width w
// (1 <= w < m), with its device tables.  Every layer must hold fewer than
// 2^32 states.
int blk_layout_create(int m, int nmax, int w, BlkLayout *out);
// Block records kept by default: 256 MiB of 128-byte records.
constexpr size_t BLK_MAX_RECORDS = (256ull << 20) / 128;
// The same without the device upload: fills `image` with the bytes that
// blk_layout_create copies to the device (offsets in the layout).
int blk_layout_build_host(int m, int nmax, int w, BlkLayout *out, std::vector<char> *image,
                          size_t max_records = BLK_MAX_RECORDS);
// General form: m_free free modes holding at most nmax_free photons, the
// sub-layers of each layer (subs[k], k = 1..subs.size()-1, subs[0] unused),
// their state counts (counts[k], k = 0..), and the U row of each local mode
// (rowmap: free modes, then herald modes).  Block records take at most
// max_records entries; the groups past that walk per lane.
int blk_layout_build_host_ex(int m_free, int nmax_free, int w, const std::vector<std::vector<BlkSub>> &subs,
                             const std::vector<uint64_t> &counts, const std::vector<int32_t> &rowmap,
                             BlkLayout *out, std::vector<char> *image, size_t max_records = BLK_MAX_RECORDS);
// The masked layers of an n-photon plan over m modes for a mask of `nrows`
// condition rows (FSMask, perceval/backends/_abstract_backends.py:103-146):
// conds[r * m + j] >= 0 asks for that many photons in mode j (at least that
// many when at_least[j] != 0), < 0 for any count; a state is kept when it
// matches some row.  Herald modes are the modes some row constrains; layer k
// keeps the states whose herald part a can still reach some row (partial
// match, tests/utils/test_mask.py:33-38): for that row, a_j <= c_j on its
// exact modes and its missing photons (c_j - a_j, at-least modes counting
// max(0, c_j - a_j)) fit in the n - k photons still to come.  Every parent
// of a kept state is kept (same row).  The final layer is the concatenation
// of its sub-layers (herald configurations in descending lexicographic
// order); when it has more than one, `perm` lists, in FSArray order of the
// kept states (descending lexicographic over the original modes), each
// one's position in the final layer.
struct BlkMasked {
    int m_free = 0, nher = 0;
    int nmax_free = 0;                          // most photons in the free modes of any sub-layer
    bool empty = false;                         // no state of the final layer matches
    std::vector<int32_t> rowmap;                // free modes, then herald modes (original indices)
    std::vector<std::vector<BlkSub>> subs;      // per layer k = 1..n
    std::vector<uint64_t> counts;               // states per layer k = 0..n
    std::vector<uint32_t> perm;                 // final layer in FSArray order (empty: already in order)
};
int blk_masked_layers(int m, int n, const int32_t *cond, BlkMasked *out);
int blk_masked_layers_ex(int m, int n, int n_mask, int nrows, const int32_t *conds, const int32_t *at_least,
                         BlkMasked *out);

// Uploads a host-built layout (allocates d_mem).
int blk_layout_upload(BlkLayout *lay, const std::vector<char> &image);
void blk_layout_destroy(BlkLayout *lay);
bool blk_layout_ok(int m, int nmax);

// Enqueues layer k of the layout: children from the parent layer k - 1,
// column mk of the m x m unitary U.  probs: write |c|^2 prod s! *
// inv_prod_in (doubles) instead of complex coefficients (prod s! over the
// free modes: herald factorials are the caller's constant, folded into
// inv_prod_in).  sentinel: write one zero coefficient after the last child
// (read by the successor kernels of the next layer).
int blk_layer_launch(const BlkLayout &lay, const cd *parent, void *out, const cd *U, int m, int mk, int k, bool probs,
                     double inv_prod_in, bool sentinel, hipStream_t stream);

} // namespace pcvl
