// Host-side check of the block SLOS layout (perceval_amd/csrc/slos_block.hip),
// built with AddressSanitizer and UndefinedBehaviorSanitizer on the host code.
//
// For many (m, n, w) it builds the layout's tables with blk_layout_build_host
// and replays slos_block_kernel's indexing on the CPU for every workgroup,
// wave and lane (the group search, the XCD-contiguous tile order, both work
// regimes, the prefix walk, the suffix slot tables), then checks:
//   * every table, parent and output index is in bounds;
//   * every child of every layer is written exactly once;
//   * each child's parent list, in the order it is summed, equals brute force
//     (rank of c - e_j for the occupied modes j, increasing);
//   * the factorial product of each child equals prod s!;
//   * each precomputed block record equals the block's walk.
// The replay restates the kernel's index arithmetic line for line; it must be
// kept in step with the kernel.  Exit status 0 when every check holds.
#include "../../perceval_amd/csrc/slos_block.hip"
#include "../../perceval_amd/csrc/capi.cpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <set>

using namespace pcvl;

static std::vector<std::vector<int>> space(int m, int n)
{
    std::vector<std::vector<int>> out;
    std::vector<int> s((size_t)m, 0);
    s[0] = n;
    while (true) {
        out.push_back(s);
        if (m == 1) break;
        int tail = s[(size_t)m - 1];
        if (tail == n) break;
        s[(size_t)m - 1] = 0;
        int t = m - 2;
        while (s[(size_t)t] == 0) t--;
        s[(size_t)t]--;
        s[(size_t)t + 1] = tail + 1;
    }
    return out;
}

static uint64_t rank_of(const std::vector<int> &s)
{
    const int m = (int)s.size();
    int R = 0;
    for (int v : s) R += v;
    uint64_t rk = 0;
    for (int q = 0; q + 1 < m; q++) {
        R -= s[(size_t)q];
        rk += blk_binom((int64_t)R + m - q - 2, m - q - 1);
    }
    return rk;
}

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { if (fails++ < 20) { printf("FAIL %s:%d ", __FILE__, __LINE__); printf(__VA_ARGS__); printf("\n"); } } } while (0)

// Replays every layer of a layout.  `layers[k]` lists the states of layer k in
// the layout's order as occupation vectors over the ORIGINAL modes; `perm`
// gives the local mode order (free modes, then herald modes) as original
// mode indices.  Expected parents of a child: its occupied modes in local
// order, each parent's position in layers[k-1].
static void replay(const BlkLayout &lay, const std::vector<char> &img, const std::vector<std::vector<std::vector<int>>> &layers,
                   const std::vector<int32_t> &perm, const char *what, int grid_cap)
{
    const int h = lay.h, wt = lay.wt, w = lay.w, mfree = lay.m;
    const uint32_t *Th = (const uint32_t *)(img.data() + lay.off_th);
    const uint64_t *TE = (const uint64_t *)(img.data() + lay.off_te);
    const uint64_t *CTE = (const uint64_t *)(img.data() + lay.off_cte);
    const uint32_t *f32 = (const uint32_t *)(img.data() + lay.off_f32);
    const double *fd = (const double *)(img.data() + lay.off_fd);
    const double *spf = (const double *)(img.data() + lay.off_pf);
    const uint32_t *sslot = (const uint32_t *)(img.data() + lay.off_slot);
    const int32_t *rowmap = (const int32_t *)(img.data() + lay.off_row);
    const uint32_t *her = (const uint32_t *)(img.data() + lay.off_her);
    const BlkRec *recs = (const BlkRec *)(img.data() + lay.off_rec);
    const size_t n_th = (lay.off_te - lay.off_th) / 4, n_te = (lay.off_cte - lay.off_te) / 8;
    const size_t n_pf = (lay.off_slot - lay.off_pf) / 8, n_slot = (lay.bytes - lay.off_slot) / 4;
    const size_t n_cte = (lay.off_f32 - lay.off_cte) / 8, n_her = (lay.group_off[1] - lay.off_her) / 8;
    for (int e = 0; e < lay.mloc; e++) CHECK(rowmap[e] == perm[(size_t)e], "%s: rowmap[%d]", what, e);
    for (int k = 1; k <= lay.nlayers; k++) {
        const auto &children = layers[(size_t)k];
        const auto &parents = layers[(size_t)k - 1];
        std::map<std::vector<int>, uint64_t> pidx;
        for (size_t i = 0; i < parents.size(); i++) pidx[parents[i]] = i;
        const uint64_t nchild = children.size(), nparent = parents.size();
        CHECK(lay.counts[(size_t)k] == nchild, "%s: count of layer %d", what, k);
        std::vector<int> written(nchild, 0);
        const BlkGroup *groups = (const BlkGroup *)(img.data() + lay.group_off[(size_t)k]);
        const int ngroups = (int)lay.groups[(size_t)k].size();
        // the persistent launch of blk_layer_launch with this grid cap
        const uint64_t want = (lay.items[(size_t)k] + 3) / 4;
        const uint32_t grid = (uint32_t)((std::min<uint64_t>(want, (uint64_t)grid_cap) + 7) / 8 * 8);
        const uint32_t per_xcd = grid >> 3, stride = per_xcd * 4;
        std::vector<std::vector<uint8_t>> seen((size_t)ngroups);
        for (int gi = 0; gi < ngroups; gi++) seen[(size_t)gi].assign(groups[gi].nitems, 0);
        for (uint32_t wg = 0; wg < grid; wg++) {
          uint32_t cum = 0;   // the kernel's rotation: XCD items of the previous groups, mod stride
          for (int gi = 0; gi < ngroups; gi++) {
            const BlkGroup g = groups[gi];
            const uint32_t x = wg & 7u, j = wg >> 3;
            const uint32_t share = (g.nitems + 7u) >> 3;
            const uint32_t lo = std::min(x * share, g.nitems), hi = std::min(lo + share, g.nitems);
            const uint32_t cum_g = cum;
            cum = (cum + (hi - lo)) % stride;
            for (uint32_t wave = 0; wave < 4; wave++)
              for (uint32_t item = lo + (j * 4 + wave + stride - cum_g) % stride; item < hi; item += stride) {
                CHECK(seen[(size_t)gi][item] == 0, "%s: item %u of group %d taken twice", what, item, gi);
                seen[(size_t)gi][item] = 1;
                const int r = (int)g.r, K = (int)g.K, smax = std::min(r, w);
                const uint32_t S = g.S;
                for (uint32_t lane = 0; lane < 64; lane++) {
                    uint32_t b, beta;
                    if (g.nch != 0) {
                        b = item / g.nch;
                        beta = (item - b * g.nch) * 64u + lane;
                    } else {
                        const uint32_t i = item * 64u + lane;
                        if (i >= g.nb * S) continue;
                        b = i / S;
                        beta = i - b * S;
                    }
                    CHECK(b < g.nb, "%s: block %u >= nb %u", what, b, g.nb);
                    uint32_t rem = b, PE = 0, Bk = 0;
                    int P = K;
                    double pfp = 1;
                    std::vector<std::pair<uint32_t, int>> lst;
                    for (int q = 0; q < h; q++) {
                        int x = 0;
                        if (q < h - 1) {
                            x = P;
                            CHECK((size_t)(q * wt + x) < n_th, "%s: Th index", what);
                            uint32_t t = Th[q * wt + x];
                            while (t > rem) { --x; CHECK(x >= 0, "%s: walk below 0", what); t = Th[q * wt + x]; }
                            rem -= t;
                        }
                        CHECK((size_t)(q * wt + x + r) < n_te, "%s: TE index", what);
                        const uint64_t te = TE[q * wt + x + r];
                        Bk += (uint32_t)te;
                        if (P > x) {
                            lst.push_back({PE, q});
                            pfp *= (lay.nmax <= 12) ? (double)f32[P - x] : fd[P - x];
                        }
                        PE += (uint32_t)(te >> 32);
                        P = x;
                        if (P == 0) {
                            CHECK((size_t)(r * (h + 1) + q + 1) < n_cte, "%s: CTE index", what);
                            const uint64_t ce = CTE[r * (h + 1) + q + 1];
                            Bk += (uint32_t)ce;
                            PE += (uint32_t)(ce >> 32);
                            break;
                        }
                    }
                    if (g.nch != 0) {
                        // whole-wave blocks read their walk from a record
                        const uint64_t rb = (uint64_t)g.rec0 + b;
                        CHECK(rb < lay.nrec, "%s: record index", what);
                        if (rb >= lay.nrec) continue;
                        const uint32_t *rw = recs[rb].w;
                        double rpf;
                        std::memcpy(&rpf, &rw[8], sizeof(double));
                        CHECK(rw[0] == Bk && rw[1] == PE && rw[2] == lst.size() && rpf == pfp,
                              "%s: record of block %u differs from its walk", what, b);
                        CHECK(rw[2] <= (uint32_t)BLK_REC_MAX, "%s: record list longer than 16", what);
                        for (size_t d = 0; d < lst.size() && d < (size_t)BLK_REC_MAX; d++)
                            CHECK(rw[10 + d] == lst[d].first && (int)((rw[4 + d / 4] >> (8 * (d % 4))) & 255u) == lst[d].second,
                                  "%s: record list of block %u", what, b);
                    }
                    if (beta >= S) continue;
                    const uint64_t loc = (uint64_t)Bk + beta;
                    const uint64_t o = (uint64_t)g.out_base + loc;
                    CHECK(o < nchild, "%s: out %lu >= %lu", what, (unsigned long)o, (unsigned long)nchild);
                    if (o >= nchild) continue;
                    written[o]++;
                    std::vector<std::pair<uint64_t, int>> got;   // (parent index, original mode)
                    for (uint32_t d = 0; d < g.nher; d++) {
                        CHECK(g.her_off + d < n_her, "%s: herald index", what);
                        const uint32_t hi2 = her[2 * (g.her_off + d)], hb = her[2 * (g.her_off + d) + 1];
                        CHECK(mfree + (int)hi2 < lay.mloc, "%s: herald row", what);
                        got.push_back({(uint64_t)hb + loc, rowmap[mfree + (int)hi2]});
                    }
                    for (auto &e : lst) got.push_back({(uint64_t)g.par_base + loc - e.first, rowmap[e.second]});
                    for (int d = 0; d < smax; d++) {
                        const size_t ti = (size_t)g.tslot + (size_t)d * S + beta;
                        CHECK(ti < n_slot, "%s: slot index", what);
                        const uint32_t sw = sslot[ti];
                        if (sw == 0xffffffffu) continue;
                        got.push_back({(uint64_t)g.par_base + (Bk - PE) + (sw >> 5), rowmap[h + (int)(sw & 31u)]});
                    }
                    CHECK((size_t)g.tpf + beta < n_pf, "%s: pf index", what);
                    const double pf = pfp * spf[g.tpf + beta];
                    // brute force, local mode order
                    const auto &c = children[o];
                    std::vector<std::pair<uint64_t, int>> exp;
                    double pfe = 1;
                    for (int e = mfree; e < lay.mloc; e++) {
                        const int j = perm[(size_t)e];
                        if (c[(size_t)j] > 0) {
                            auto cc = c;
                            cc[(size_t)j]--;
                            auto it = pidx.find(cc);
                            CHECK(it != pidx.end(), "%s: parent missing from layer %d", what, k - 1);
                            exp.push_back({it == pidx.end() ? ~0ull : it->second, j});
                        }
                    }
                    for (int e = 0; e < mfree; e++) {
                        const int j = perm[(size_t)e];
                        for (int v = 2; v <= c[(size_t)j]; v++) pfe *= v;
                        if (c[(size_t)j] > 0) {
                            auto cc = c;
                            cc[(size_t)j]--;
                            auto it = pidx.find(cc);
                            CHECK(it != pidx.end(), "%s: parent missing from layer %d", what, k - 1);
                            exp.push_back({it == pidx.end() ? ~0ull : it->second, j});
                        }
                    }
                    CHECK(got == exp, "%s: k=%d r=%d b=%u beta=%u: parent list differs", what, k, r, b, beta);
                    for (auto &e : got) CHECK(e.first < nparent, "%s: parent %lu >= %lu", what, (unsigned long)e.first, (unsigned long)nparent);
                    CHECK(pf == pfe, "%s: pf %g vs %g", what, pf, pfe);
                    // the herald modes' factorials ride on the final layer's groups
                    double hfe = 1, hfg;
                    if (k == lay.nlayers)
                        for (int e = mfree; e < lay.mloc; e++)
                            for (int v = 2; v <= c[(size_t)perm[(size_t)e]]; v++) hfe *= v;
                    const uint32_t hw[2] = {g.hf_lo, g.hf_hi};
                    std::memcpy(&hfg, hw, sizeof(double));
                    CHECK(hfg == hfe, "%s: k=%d herald factorials %g vs %g", what, k, hfg, hfe);
                }
              }
          }
        }
        for (int gi = 0; gi < ngroups; gi++)
            for (uint32_t it = 0; it < groups[gi].nitems; it++)
                CHECK(seen[(size_t)gi][it] == 1, "%s: item %u of group %d not taken", what, it, gi);
        for (uint64_t i = 0; i < nchild; i++)
            CHECK(written[i] == 1, "%s: layer %d child %lu written %d times", what, k, (unsigned long)i, written[i]);
    }
}

static void emulate(int m, int nmax, int w, size_t max_records = BLK_MAX_RECORDS)
{
    BlkLayout lay;
    std::vector<char> img;
    if (blk_layout_build_host(m, nmax, w, &lay, &img, max_records) != OK) {
        printf("build failed %s\n", pcvl_last_error());
        exit(1);
    }
    std::vector<std::vector<std::vector<int>>> layers;
    for (int k = 0; k <= nmax; k++) layers.push_back(space(m, k));
    std::vector<int32_t> perm((size_t)m);
    for (int j = 0; j < m; j++) perm[(size_t)j] = j;
    char what[96];
    snprintf(what, sizeof what, "m=%d n=%d w=%d records<=%zu", m, nmax, w, max_records);
    CHECK(lay.nrec <= max_records, "%s: %lu records", what, (unsigned long)lay.nrec);
    replay(lay, img, layers, perm, what, 16);
    replay(lay, img, layers, perm, what, 2048);
}

// Masked: cond[j] >= 0 fixes mode j.  The expected layer k holds the states
// with s_H <= c and k - |s_H| <= n - |c|, in descending lexicographic order
// of (herald modes, free modes).
static void emulate_masked(int m, int n, const std::vector<int32_t> &cond, int w)
{
    BlkMasked mk;
    if (blk_masked_layers(m, n, cond.data(), &mk) != OK) { printf("masked failed %s\n", pcvl_last_error()); exit(1); }
    char what[96];
    snprintf(what, sizeof what, "masked m=%d n=%d w=%d", m, n, w);
    int C = 0;
    for (int32_t v : cond) if (v >= 0) C += v;
    CHECK(mk.empty == (C > n), "%s: empty flag", what);
    if (mk.empty) return;
    std::vector<int32_t> perm = mk.rowmap;
    std::vector<std::vector<std::vector<int>>> layers;
    for (int k = 0; k <= n; k++) {
        std::vector<std::pair<std::vector<int>, std::vector<int>>> kept;   // (local order key, state)
        for (const auto &s : space(m, k)) {
            int sa = 0;
            bool ok = true;
            for (int j = 0; j < m; j++)
                if (cond[(size_t)j] >= 0) {
                    sa += s[(size_t)j];
                    if (s[(size_t)j] > cond[(size_t)j]) ok = false;
                }
            if (!ok || k - sa > n - C) continue;
            std::vector<int> key;
            for (int e = mk.m_free; e < (int)perm.size(); e++) key.push_back(s[(size_t)perm[(size_t)e]]);
            for (int e = 0; e < mk.m_free; e++) key.push_back(s[(size_t)perm[(size_t)e]]);
            kept.push_back({key, s});
        }
        std::sort(kept.begin(), kept.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        std::vector<std::vector<int>> L;
        for (auto &e : kept) L.push_back(e.second);
        CHECK(mk.counts[(size_t)k] == L.size(), "%s: count of layer %d: %lu vs %zu", what, k,
              (unsigned long)mk.counts[(size_t)k], L.size());
        layers.push_back(L);
    }
    BlkLayout lay;
    std::vector<char> img;
    if (blk_layout_build_host_ex(mk.m_free, mk.nmax_free, w, mk.subs, mk.counts, mk.rowmap, &lay, &img) != OK) {
        printf("masked build failed %s\n", pcvl_last_error());
        exit(1);
    }
    replay(lay, img, layers, perm, what, 24);
    replay(lay, img, layers, perm, what, 2048);
}

// General masks: several condition rows (a state matches when it matches
// one), at-least modes.  The expected final layer is the brute-force set of
// matching n-photon states; layer k is its closure under photon removal (the
// states some matching output extends), ordered by (constrained modes, free
// modes) descending; the output order is the FSArray order of the final set.
static void emulate_masks(int m, int n, const std::vector<std::vector<int32_t>> &rows,
                          const std::vector<int32_t> &at_least, int w, int n_mask = -1)
{
    if (n_mask < 0) n_mask = n;
    const int nrows = (int)rows.size();
    std::vector<int32_t> flat;
    for (const auto &r : rows) flat.insert(flat.end(), r.begin(), r.end());
    BlkMasked mk;
    if (blk_masked_layers_ex(m, n, n_mask, nrows, flat.data(), at_least.empty() ? nullptr : at_least.data(), &mk) != OK) {
        printf("masks failed %s\n", pcvl_last_error());
        exit(1);
    }
    char what[96];
    snprintf(what, sizeof what, "masks m=%d n=%d n_mask=%d rows=%d w=%d", m, n, n_mask, nrows, w);
    std::vector<std::vector<int>> full;   // the n_mask-photon states matching a row
    for (const auto &s : space(m, n_mask)) {
        bool any = false;
        for (const auto &r : rows) {
            bool ok = true;
            for (int j = 0; j < m && ok; j++) {
                if (r[(size_t)j] < 0) continue;
                const bool floor = !at_least.empty() && at_least[(size_t)j];
                ok = floor ? s[(size_t)j] >= r[(size_t)j] : s[(size_t)j] == r[(size_t)j];
            }
            any = any || ok;
        }
        if (any) full.push_back(s);
    }
    // every layer, the output layer included, is the closure of those under photon removal
    std::vector<std::set<std::vector<int>>> closure((size_t)n_mask + 1);
    closure[(size_t)n_mask].insert(full.begin(), full.end());
    for (int k = n_mask; k >= 1; k--)
        for (const auto &t : closure[(size_t)k])
            for (int j = 0; j < m; j++)
                if (t[(size_t)j] > 0) {
                    auto u = t;
                    u[(size_t)j]--;
                    closure[(size_t)k - 1].insert(u);
                }
    std::vector<std::vector<int>> fin;   // the kept outputs, FSArray order
    for (const auto &s : space(m, n))
        if (closure[(size_t)n].count(s)) fin.push_back(s);
    CHECK(mk.empty == fin.empty(), "%s: empty flag", what);
    if (mk.empty) return;
    const std::vector<int32_t> perm = mk.rowmap;
    std::vector<std::vector<std::vector<int>>> layers;
    for (int k = 0; k <= n; k++) {
        std::vector<std::pair<std::vector<int>, std::vector<int>>> kept;
        for (const auto &s : closure[(size_t)k]) {
            std::vector<int> key;
            for (int e = mk.m_free; e < (int)perm.size(); e++) key.push_back(s[(size_t)perm[(size_t)e]]);
            for (int e = 0; e < mk.m_free; e++) key.push_back(s[(size_t)perm[(size_t)e]]);
            kept.push_back({key, s});
        }
        std::sort(kept.begin(), kept.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        std::vector<std::vector<int>> L;
        for (auto &e : kept) L.push_back(e.second);
        CHECK(mk.counts[(size_t)k] == L.size(), "%s: count of layer %d: %lu vs %zu", what, k,
              (unsigned long)mk.counts[(size_t)k], L.size());
        layers.push_back(L);
    }
    if (fails) return;
    // output order: perm[i] is the layer position of the i-th state in FSArray order
    std::map<std::vector<int>, uint32_t> pos;
    for (size_t i = 0; i < layers[(size_t)n].size(); i++) pos[layers[(size_t)n][i]] = (uint32_t)i;
    if (mk.perm.empty()) {
        CHECK(layers[(size_t)n] == fin, "%s: single configuration but the layer is not in FSArray order", what);
    } else {
        CHECK(mk.perm.size() == fin.size(), "%s: output order holds %zu entries, %zu expected", what, mk.perm.size(),
              fin.size());
        for (size_t i = 0; i < fin.size() && i < mk.perm.size(); i++)
            CHECK(mk.perm[i] == pos[fin[i]], "%s: output %zu: order %u vs %u", what, i, mk.perm[i], pos[fin[i]]);
    }
    BlkLayout lay;
    std::vector<char> img;
    if (blk_layout_build_host_ex(mk.m_free, mk.nmax_free, w, mk.subs, mk.counts, mk.rowmap, &lay, &img) != OK) {
        printf("masks build failed %s\n", pcvl_last_error());
        exit(1);
    }
    replay(lay, img, layers, perm, what, 24);
    replay(lay, img, layers, perm, what, 2048);
}

int main()
{
    int cases = 0;
    for (int m = 2; m <= 9; m++)
        for (int n = 1; n <= 7; n++)
            for (int w = 1; w < m && w <= 16; w++) {
                if (nstates(m, n) > 200000) continue;
                emulate(m, n, w);
                cases++;
            }
    emulate(12, 6, 10); cases++;
    emulate(14, 9, 10); cases++;
    emulate(24, 5, 10); cases++;
    emulate(24, 6, 8); cases++;
    emulate(20, 6, 12); cases++;
    emulate(3, 14, 2); cases++;
    emulate(8, 13, 7); cases++;
    // a record budget that runs out part way: later groups walk per lane
    for (size_t cap : {0, 40, 300, 2000}) {
        emulate(12, 6, 4, cap); cases++;
        emulate(10, 7, 3, cap); cases++;
    }
    // masks: heralds anywhere among the modes, several counts, empty masks
    const std::vector<std::pair<int, std::vector<int32_t>>> masks = {
        {6, {-1, -1, -1, -1, 1, 1}}, {6, {1, -1, 0, -1, -1, -1}}, {7, {-1, 2, -1, -1, 0, -1, 1}},
        {8, {0, -1, -1, 1, -1, -1, -1, -1}}, {5, {-1, -1, -1, 3, -1}}, {9, {-1, 1, -1, 1, -1, 1, -1, -1, -1}},
        {6, {2, 2, -1, -1, -1, -1}}, {7, {-1, -1, 4, -1, -1, -1, -1}}};
    for (const auto &mm : masks) {
        const int m = mm.first;
        int nfree = 0;
        for (int32_t v : mm.second) nfree += v < 0;
        for (int n = 0; n <= 6; n++)
            for (int w = 1; w < nfree && w <= 16; w++) {
                if (n == 0) continue;
                emulate_masked(m, n, mm.second, w);
                cases++;
            }
    }
    emulate_masked(16, 7, {1, -1, -1, -1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0}, 8); cases++;
    // several rows, at-least modes, rows that leave a herald mode free
    struct MaskCase { int m; std::vector<std::vector<int32_t>> rows; std::vector<int32_t> at; };
    const std::vector<MaskCase> general = {
        {6, {{-1, -1, -1, 0, 1, 1}, {-1, -1, -1, 1, 1, 0}}, {}},
        {6, {{-1, -1, 1, -1, -1, 1}}, {0, 0, 1, 0, 0, 0}},
        {7, {{1, -1, -1, -1, -1, -1, 0}, {-1, -1, 2, -1, -1, -1, -1}}, {}},
        {7, {{1, -1, -1, 2, -1, -1, -1}, {0, -1, -1, 0, -1, -1, -1}}, {1, 0, 0, 1, 0, 0, 0}},
        {8, {{0, 1, -1, -1, -1, -1, -1, 1}, {1, 0, -1, -1, -1, -1, -1, 1}, {2, 2, -1, -1, -1, -1, -1, 0}}, {0, 0, 0, 0, 0, 0, 0, 1}},
        {6, {{3, -1, -1, -1, -1, 3}}, {}},
        {5, {{-1, 1, -1, -1, 1}, {-1, 1, -1, -1, 1}}, {}}};
    for (const auto &mc : general) {
        int nfree = 0;
        for (int j = 0; j < mc.m; j++) {
            bool fr = true;
            for (const auto &r : mc.rows) fr = fr && r[(size_t)j] < 0;
            nfree += fr;
        }
        for (int n = 1; n <= 6; n++)
            for (int w = 1; w < nfree && w <= 16; w++) {
                emulate_masks(mc.m, n, mc.rows, mc.at, w);
                cases++;
            }
    }
    emulate_masks(14, 6, {{1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0}, {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1}},
                  {}, 8);
    cases++;
    // masks set for more photons than the input holds (a part of a separated
    // input): the output keeps the states that the missing photons complete
    for (const auto &mc : general) {
        int nfree = 0;
        for (int j = 0; j < mc.m; j++) {
            bool fr = true;
            for (const auto &r : mc.rows) fr = fr && r[(size_t)j] < 0;
            nfree += fr;
        }
        for (int n = 1; n <= 4; n++)
            for (int extra = 1; extra <= 3; extra++) {
                emulate_masks(mc.m, n, mc.rows, mc.at, std::min(3, nfree - 1), n + extra);
                cases++;
            }
    }
    {   // random masks: 2..4 free modes among 5..8, 1..3 rows, floors, counts 0..3, a mask
        // photon count up to 2 above the input's (fixed seed)
        uint64_t st = 0x9e3779b97f4a7c15ull;
        auto rnd = [&](int lo, int hi) {   // uniform in [lo, hi]
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            return lo + (int)(st % (uint64_t)(hi - lo + 1));
        };
        for (int trial = 0; trial < 150; trial++) {
            const int m = rnd(5, 8), nfree = rnd(2, 4), nrows = rnd(1, 3);
            std::vector<int32_t> free_mode((size_t)m, 0), at((size_t)m, 0);
            for (int f = 0; f < nfree; f++) {
                int j = rnd(0, m - 1);
                while (free_mode[(size_t)j]) j = (j + 1) % m;
                free_mode[(size_t)j] = 1;
            }
            std::vector<std::vector<int32_t>> rows((size_t)nrows, std::vector<int32_t>((size_t)m, -1));
            for (auto &r : rows)
                for (int j = 0; j < m; j++)
                    if (!free_mode[(size_t)j]) r[(size_t)j] = rnd(0, 4) == 0 ? -1 : rnd(0, 3);
            for (int j = 0; j < m; j++) at[(size_t)j] = !free_mode[(size_t)j] && rnd(0, 3) == 0;
            const int n = rnd(1, 5);
            emulate_masks(m, n, rows, at, rnd(1, nfree - 1), n + rnd(0, 2));
            cases++;
        }
    }
    {   // floors of 0 on 24 modes: C(34, 10) configurations, refused after the cap, not enumerated
        std::vector<int32_t> row(26, 0), at(26, 1);
        row[0] = row[1] = -1;
        BlkMasked mk;
        CHECK(blk_masked_layers_ex(26, 10, 10, 1, row.data(), at.data(), &mk) == ERR_RANGE, "configuration cap");
        cases++;
    }
    printf("%d cases, %d failures\n", cases, fails);
    return fails ? 1 : 0;
}
