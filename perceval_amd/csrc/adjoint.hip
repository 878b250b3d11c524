// Gradients of output probabilities with respect to circuit phases.
//
// The reference has no analytic gradient: its QML loop differentiates
// numerically (benchmark/benchmark_QML-DE-solver.py:140-142) through repeated
// SLOSBackend.set_circuit/all_prob calls (perceval/backends/_slos.py:130-141).
// This kernel returns, for a batch of phase vectors of one mesh, the loss
//     L(phi) = sum_{s in targets} P(s),  P(s) = |perm(U[s, in])|^2 / (prod s! prod t!)
// and dL/dphi for every phase, in one pass per parameter set:
//   1. U(phi) = L_K ... L_1 from the mesh program (phase shifters and fixed
//      2x2 blocks, circuit order as perceval/components/linear_circuit.py:487-503);
//   2. per target, the amplitude A_s and all n*n cofactor permanents
//      d perm / d a_rc by one Glynn pass (2^{n-1} terms; per term the column
//      sums, prod_{c' != c} by prefix/suffix, and delta_r); they accumulate
//      G_ij = sum_s conj(A_s) dA_s/dU_ij;
//   3. one backward sweep over the components: with F = L_p^H ... (peeling
//      F_K = U) and Z = G^T L_K ... L_{p+1}, a phase phi on mode q of
//      component p contributes dL/dphi = 2 Re(i e^{i phi} (F Z)_qq).
// One workgroup per parameter set; all arithmetic complex FP64.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

namespace pcvl {

constexpr int ADJ_MAX_M = 32;
constexpr int ADJ_MAX_N = 8;
constexpr int ADJ_MAX_TARGETS = 512;     // targets of one plan (its LDS decides what fits)
constexpr int ADJ_BLOCK_TARGETS = 64;    // targets whose cofactors are in LDS at once
constexpr int ADJ_THREADS = 512;

struct MeshOp {
    int32_t kind;    // 0: phase shifter on mode q; 1: fixed 2x2 block on (q, q+1)
    int32_t q;
    int32_t param;   // phase index (kind 0), -1 for a constant phase
    int32_t pad;
    double c[8];     // kind 0: c[0] = constant phase; kind 1: 2x2 block row-major (re, im)
};

__device__ __forceinline__ cd cexpi(double ph) { return {cos(ph), sin(ph)}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.r, -a.i}; }

// Sum over the 64 lanes of a wave, every lane receiving it (butterfly: a
// fixed order, so the result does not depend on timing).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Glynn with derivatives for target row r of M = U[rows, cols] (n = N; U in
// LDS, row-major m x m, rows the target's modes, cols the input's): writes
// nrm * d perm / d M[r][c] to cof[c] for c < N, and nrm * perm(M) to *amp
// when r == 0.  The matrix is read from LDS as the Gray walk flips its rows,
// so no N x N copy sits in registers.
template <int N>
__device__ void glynn_with_row_derivs(const cd *U, int m, const int32_t *rows, const int32_t *cols, int r, double nrm,
                                      cd *cof, cd *amp)
{
    cd perm, dr[N];
    int col[N];
#pragma unroll
    for (int c = 0; c < N; c++) col[c] = cols[c];
    cd cs[N];
#pragma unroll
    for (int c = 0; c < N; c++) cs[c] = {0.0, 0.0};
    for (int i = 0; i < N; i++) {
        const cd *row = U + (size_t)rows[i] * m;
#pragma unroll
        for (int c = 0; c < N; c++) cs[c] = cadd(cs[c], row[col[c]]);
    }
    perm = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < N; c++) dr[c] = {0.0, 0.0};
    uint32_t g = 0;
    for (uint32_t t = 0; t < (1u << (N - 1)); t++) {
        if (t) {
            const int b = __builtin_ctz(t);
            g ^= 1u << b;
            const double sgn = ((g >> b) & 1) ? -2.0 : 2.0;
            const cd *row = U + (size_t)rows[b + 1] * m;
#pragma unroll
            for (int c = 0; c < N; c++) {
                const cd v = row[col[c]];
                cs[c].r = fma(sgn, v.r, cs[c].r);
                cs[c].i = fma(sgn, v.i, cs[c].i);
            }
        }
        const double sign = (t & 1) ? -1.0 : 1.0;
        // delta_r of this term (row 0 always +1)
        const double dr_sign = (r == 0) ? 1.0 : (((g >> (r - 1)) & 1) ? -1.0 : 1.0);
        cd P[N];
        P[0] = {1.0, 0.0};
#pragma unroll
        for (int c = 1; c < N; c++) P[c] = cmul(P[c - 1], cs[c - 1]);
        cd S = {sign, 0.0};
#pragma unroll
        for (int c = N - 1; c >= 0; c--) {
            const cd ex = cmul(P[c], S);      // sign * prod_{c' != c} cs_c'
            dr[c].r = fma(dr_sign, ex.r, dr[c].r);
            dr[c].i = fma(dr_sign, ex.i, dr[c].i);
            S = cmul(S, cs[c]);
        }
        perm = cadd(perm, S);                 // S = sign * prod_c cs_c
    }
    const double sc = ldexp(1.0, -(N - 1)) * nrm;
#pragma unroll
    for (int c = 0; c < N; c++) cof[c] = {dr[c].r * sc, dr[c].i * sc};
    if (r == 0) *amp = {perm.r * sc, perm.i * sc};
}

template <int N>
__device__ __forceinline__ void dispatch_glynn(int n, const cd *U, int m, const int32_t *rows, const int32_t *cols,
                                               int r, double nrm, cd *cof, cd *amp)
{
    if constexpr (N <= ADJ_MAX_N) {
        if (n == N) { glynn_with_row_derivs<N>(U, m, rows, cols, r, nrm, cof, amp); return; }
        dispatch_glynn<N + 1>(n, U, m, rows, cols, r, nrm, cof, amp);
    }
}

// Index tables of one call, one int32 array (offsets in entries):
//   lop[nops]        ops in level order: a level's ops act on disjoint modes,
//                    so they commute and may run side by side
//   lstart[nlev+1]   first entry of each level in lop
//   rstart[m+1], rlist[targets*n]   per mode i, the (s*n + r) with rows_s[r] = i
//   cstart[m+1], clist[n]           per mode j, the c with in[c] = j
//   tm[targets*n], in[n]            target modes, input modes
struct AdjIdx {
    int lop, lstart, rstart, rlist, cstart, clist, tm, in, total;
    int nlev;
    int glog;   // lanes per op slot in the sweeps: 2^glog >= m
};

// Dynamic LDS of the kernel, in order (byte offsets rounded to 16): the mesh
// program (when staged), e^{i phi} of the set's phases, U (m x m), Z (m x m),
// amplitudes, cofactors (targets x n x n), target norms, a reduction buffer of
// ADJ_THREADS doubles, the index tables.
struct AdjLds {
    size_t ops, zph, U, Z, amp, cof, tn, tw, red, idx, bytes;
};

__host__ __device__ inline size_t adj_al16(size_t b) { return (b + 15) & ~(size_t)15; }

__host__ __device__ inline AdjLds adj_lds(int nops, int nparams, int m, int n, int ntargets, int nidx)
{
    AdjLds L;
    size_t o = 0;
    L.ops = o; o = adj_al16(o + sizeof(MeshOp) * (size_t)(nops > 0 ? nops : 1));
    L.zph = o; o = adj_al16(o + sizeof(cd) * (size_t)(nparams > 0 ? nparams : 1));
    L.U = o; o = adj_al16(o + sizeof(cd) * (size_t)m * m);
    L.Z = o; o = adj_al16(o + sizeof(cd) * (size_t)m * m);
    const int nb = ntargets < ADJ_BLOCK_TARGETS ? ntargets : ADJ_BLOCK_TARGETS;
    L.amp = o; o = adj_al16(o + sizeof(cd) * (size_t)nb);
    L.cof = o; o = adj_al16(o + sizeof(cd) * (size_t)nb * n * n);
    L.tn = o; o = adj_al16(o + sizeof(double) * (size_t)ntargets);
    L.tw = o; o = adj_al16(o + sizeof(double) * (size_t)ntargets);
    L.red = o; o = adj_al16(o + sizeof(double) * ADJ_THREADS);
    L.idx = o; o = adj_al16(o + sizeof(int32_t) * (size_t)nidx);
    L.bytes = o;
    return L;
}

// One DPP step of a lane-group sum: v + v[lane ^ pattern] on the two 32-bit
// halves of a double (the DPP lane moves are 32-bit).
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    return v + __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over each group of 2^glog lanes, every lane of the group receiving it,
// in a fixed order: pairs within quads (quad_perm 1,0,3,2 then 2,3,0,1), then
// the two quads of each 8 (row_half_mirror), the two 8s of each row of 16
// (row_mirror), and across the two rows of a 32-lane group by a lane shuffle.
// The DPP moves stay in registers; only the last step goes through the LDS
// crossbar.
__device__ __forceinline__ double group_sum(double v, int glog)
{
    if (glog >= 1) v = dpp_add<0xB1>(v);
    if (glog >= 2) v = dpp_add<0x4E>(v);
    if (glog >= 3) v = dpp_add<0x141>(v);
    if (glog >= 4) v = dpp_add<0x140>(v);
    if (glog >= 5) v += __shfl_xor(v, 16, 64);
    return v;
}

// One workgroup per parameter set (grid-stride).  The forward product and the
// backward sweep run in wave 0, level by level: the wave holds 64 >> glog op
// slots of 2^glog lanes, lane `col` of a slot owning column col of U / F and
// row col of Z.  An op touches one or two rows of F and the same columns of
// Z, which each lane updates in its own column / row, and the ops of a level
// act on disjoint modes, so the sweeps need no barrier; the one cross-lane
// quantity, (F Z)_qq, is a reduction within the slot.
__global__ void __launch_bounds__(ADJ_THREADS) mesh_adjoint_kernel(
    const MeshOp *__restrict__ ops_g, int nops, int m, const double *__restrict__ phases, int nparams, int nsets,
    int n, const int32_t *__restrict__ idx_g, AdjIdx X, const double *__restrict__ tn_g,
    const double *__restrict__ tw_g, int ntargets, double *__restrict__ loss_out, double *__restrict__ grad_out,
    double *__restrict__ probs_out, int stage_ops)
{
    extern __shared__ uint64_t s_adj[];
    const AdjLds L = adj_lds(stage_ops ? nops : 0, nparams, m, n, ntargets, X.total);
    char *base = reinterpret_cast<char *>(s_adj);
    cd *zph = reinterpret_cast<cd *>(base + L.zph);
    MeshOp *ops_l = reinterpret_cast<MeshOp *>(base + L.ops);
    const MeshOp *ops = stage_ops ? ops_l : ops_g;   // long programs are read from global memory
    cd *U = reinterpret_cast<cd *>(base + L.U);
    cd *Z = reinterpret_cast<cd *>(base + L.Z);
    cd *amp = reinterpret_cast<cd *>(base + L.amp);
    cd *cof = reinterpret_cast<cd *>(base + L.cof);
    double *tn = reinterpret_cast<double *>(base + L.tn);
    double *tw = reinterpret_cast<double *>(base + L.tw);
    double *red = reinterpret_cast<double *>(base + L.red);
    int32_t *idx = reinterpret_cast<int32_t *>(base + L.idx);
    const int32_t *lop = idx + X.lop, *lstart = idx + X.lstart, *rstart = idx + X.rstart, *rlist = idx + X.rlist;
    const int32_t *cstart = idx + X.cstart, *clist = idx + X.clist, *tm = idx + X.tm, *in_modes = idx + X.in;
    const int tid = threadIdx.x;
    // the program and index tables, once per workgroup
    if (stage_ops)
        for (int e = tid; e < nops * (int)(sizeof(MeshOp) / 8); e += ADJ_THREADS)
            reinterpret_cast<uint64_t *>(ops_l)[e] = reinterpret_cast<const uint64_t *>(ops_g)[e];
    for (int e = tid; e < X.total; e += ADJ_THREADS) idx[e] = idx_g[e];
    for (int e = tid; e < ntargets; e += ADJ_THREADS) {
        tn[e] = tn_g[e];
        tw[e] = tw_g[e];
    }
    __syncthreads();
    const bool w0 = tid < 64;
    const int glog = X.glog, slots = 64 >> glog;
    const int slot = tid >> glog, col = tid & ((1 << glog) - 1);   // in wave 0
    for (int set = blockIdx.x; set < nsets; set += gridDim.x) {
        const double *ph = phases + (size_t)set * nparams;
        for (int e = tid; e < nparams; e += ADJ_THREADS) zph[e] = cexpi(ph[e]);
        __syncthreads();
        // 1. U = L_K ... L_1: start from the identity, apply the levels in circuit order
        if (w0) {
            if (slot == 0 && col < m)
                for (int i = 0; i < m; i++) U[i * m + col] = {i == col ? 1.0 : 0.0, 0.0};
            for (int lv = 0; lv < X.nlev; lv++) {
                const int e1 = lstart[lv + 1];
                for (int k = lstart[lv] + slot; k - slot < e1; k += slots) {
                    if (k >= e1 || col >= m) continue;
                    const MeshOp &op = ops[lop[k]];
                    const int q = op.q;
                    if (op.kind == 0) {
                        const cd z = op.param >= 0 ? zph[op.param] : cexpi(op.c[0]);
                        U[q * m + col] = cmul(z, U[q * m + col]);
                    } else {
                        const cd a = {op.c[0], op.c[1]}, b = {op.c[2], op.c[3]}, c = {op.c[4], op.c[5]},
                                 d = {op.c[6], op.c[7]};
                        const cd x = U[q * m + col], y = U[(q + 1) * m + col];
                        U[q * m + col] = cadd(cmul(a, x), cmul(b, y));
                        U[(q + 1) * m + col] = cadd(cmul(c, x), cmul(d, y));
                    }
                }
            }
        }
        __syncthreads();
        // 2. per block of up to 64 targets: amplitudes and cofactors (thread (target s, row r)), the block's
        // share of the loss, and its share of G^T, added into Z in block order:
        // Z[j][i] = G[i][j] = sum over (s, r) with rows_s[r] = i (increasing s, r) of
        // w_s conj(A_s) sum over c with in[c] = j of cof[s][r][c]
        double lacc = 0.0;   // thread 0: the loss, in target order
        for (int s0 = 0, blk = 0; s0 < ntargets; s0 += ADJ_BLOCK_TARGETS, blk++) {
            const int nt = ntargets - s0 < ADJ_BLOCK_TARGETS ? ntargets - s0 : ADJ_BLOCK_TARGETS;
            for (int w = tid; w < nt * n; w += ADJ_THREADS) {
                const int s = w / n, r = w - (w / n) * n;
                // scaled by 1 / sqrt(prod s! prod t!)
                dispatch_glynn<1>(n, U, m, tm + (s0 + s) * n, in_modes, r, tn[s0 + s], cof + (size_t)(s * n + r) * n,
                                  amp + s);
            }
            __syncthreads();
            // loss and per-target probabilities
            if (tid < nt) {
                const cd a = amp[tid];
                const double pr = a.r * a.r + a.i * a.i;
                if (probs_out) probs_out[(size_t)set * ntargets + s0 + tid] = pr;
                red[tid] = tw[s0 + tid] * pr;
            }
            __syncthreads();
            if (tid == 0)
                for (int s = 0; s < nt; s++) lacc += red[s];
            const int *rs = rstart + blk * (m + 1);
            for (int e = tid; e < m * m; e += ADJ_THREADS) {
                const int j = e / m, i = e - (e / m) * m;
                cd g = {0.0, 0.0};
                const int c0 = cstart[j], c1 = cstart[j + 1];
                if (c0 < c1)
                    for (int u = rs[i]; u < rs[i + 1]; u++) {
                        const int sr = rlist[u] - s0 * n;   // (s, r) within the block
                        const cd a = amp[sr / n];
                        const double ws = tw[s0 + sr / n];
                        const cd ca = {ws * a.r, -(ws * a.i)};
                        for (int v = c0; v < c1; v++) cfma_acc(g, ca, cof[(size_t)sr * n + clist[v]]);
                    }
                if (blk == 0) Z[j * m + i] = g;
                else Z[j * m + i] = cadd(Z[j * m + i], g);
            }
            __syncthreads();   // amp, cof and red are rewritten by the next block
        }
        if (tid == 0) loss_out[set] = lacc;
        // 3. backward sweep (wave 0), levels in reverse: F (in U) <- L_p^H F, grad, Z <- Z L_p
        if (w0) {
            for (int lv = X.nlev - 1; lv >= 0; lv--) {
                const int e0 = lstart[lv], e1 = lstart[lv + 1];
                for (int k0 = e0; k0 < e1; k0 += slots) {
                    const int k = k0 + slot;
                    const int pidx = k < e1 ? lop[k] : -1;
                    const int kind = pidx >= 0 ? ops[pidx].kind : -1;
                    const bool mine = pidx >= 0 && col < m;
                    bool graded = false;
                    cd z = {1.0, 0.0}, f = {0.0, 0.0}, zq = {0.0, 0.0};
                    int q = 0, prm = -1;
                    if (kind == 0) {
                        const MeshOp &op = ops[pidx];
                        q = op.q;
                        prm = op.param;
                        z = prm >= 0 ? zph[prm] : cexpi(op.c[0]);
                        graded = prm >= 0;
                        if (mine) {
                            f = cmul(cconj(z), U[q * m + col]);   // F <- L^H F
                            U[q * m + col] = f;
                            zq = Z[col * m + q];
                        }
                    }
                    // (F Z)_qq = sum_k F[q][k] Z[k][q]; dL/dphi = 2 Re(i z (F Z)_qq)
                    if (__ballot(graded) != 0) {   // uniform: every lane takes part in the reduction
                        const cd v = cmul(f, zq);
                        const double fr = group_sum(v.r, glog), fi = group_sum(v.i, glog);
                        if (graded && col == 0) grad_out[(size_t)set * nparams + prm] = -2.0 * (z.r * fi + z.i * fr);
                    }
                    if (kind == 0 && mine) Z[col * m + q] = cmul(zq, z);   // Z <- Z L
                    if (kind == 1 && mine) {
                        const MeshOp &op = ops[pidx];
                        const int q1 = op.q;
                        const cd a = {op.c[0], op.c[1]}, b = {op.c[2], op.c[3]}, c = {op.c[4], op.c[5]},
                                 d = {op.c[6], op.c[7]};
                        // F <- L^H F on rows q, q+1: L^H = [[a*, c*], [b*, d*]]
                        const cd x = U[q1 * m + col], y = U[(q1 + 1) * m + col];
                        U[q1 * m + col] = cadd(cmul(cconj(a), x), cmul(cconj(c), y));
                        U[(q1 + 1) * m + col] = cadd(cmul(cconj(b), x), cmul(cconj(d), y));
                        // Z <- Z L on columns q, q+1
                        const cd u = Z[col * m + q1], v = Z[col * m + q1 + 1];
                        Z[col * m + q1] = cadd(cmul(u, a), cmul(v, c));
                        Z[col * m + q1 + 1] = cadd(cmul(u, b), cmul(v, d));
                    }
                }
            }
        }
        __syncthreads();   // U and Z are rewritten by the next set
    }
}

} // namespace pcvl

using namespace pcvl;

struct pcvl_mesh_plan_s {
    int m = 0, n = 0, nops = 0, nparams = 0, ntargets = 0;
    AdjIdx X{};
    bool stage_ops = false;
    size_t lds = 0;
    char *d_blk = nullptr;            // program, index tables, target norms
    const MeshOp *d_ops = nullptr;
    const int32_t *d_idx = nullptr;
    const double *d_tn = nullptr;
    double *d_tw = nullptr;           // per-target weights of the loss (ones unless set)
    int grid_cap = 0;
};

extern "C" int pcvl_mesh_plan_create(const int32_t *h_kind, const int32_t *h_mode, const int32_t *h_param,
                                     const double *h_const, int nops, int m, int nparams, const int32_t *h_in_occ,
                                     const int32_t *h_targets, int ntargets, pcvl_mesh_plan *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out plan pointer is null");
    *out = nullptr;
    PCVL_ARG_CHECK(m >= 1 && m <= ADJ_MAX_M, "mode count outside [1, 32]");
    PCVL_ARG_CHECK(nops >= 0 && (nops == 0 || (h_kind && h_mode && h_param && h_const)), "bad mesh program");
    PCVL_ARG_CHECK(nparams >= 0, "negative sizes");
    PCVL_ARG_CHECK(ntargets >= 1 && ntargets <= ADJ_MAX_TARGETS, "target count outside [1, 512]");
    PCVL_ARG_CHECK(h_in_occ && h_targets, "null input/target states");
    std::vector<MeshOp> ops(nops);
    for (int p = 0; p < nops; p++) {
        MeshOp &o = ops[p];
        o.kind = h_kind[p];
        o.q = h_mode[p];
        o.param = h_param[p];
        o.pad = 0;
        for (int k = 0; k < 8; k++) o.c[k] = h_const[p * 8 + k];
        PCVL_ARG_CHECK(o.kind == 0 || o.kind == 1, "op kind must be 0 (phase) or 1 (2x2 block)");
        PCVL_ARG_CHECK(o.q >= 0 && o.q + (o.kind == 1 ? 1 : 0) < m, "op mode out of range");
        PCVL_ARG_CHECK(o.kind == 1 ? o.param == -1 : (o.param >= -1 && o.param < nparams), "op parameter out of range");
    }
    std::vector<int32_t> in_modes, tmodes;
    double pin = 1.0;
    for (int i = 0; i < m; i++) {
        PCVL_ARG_CHECK(h_in_occ[i] >= 0, "negative occupation");
        for (int v = 0; v < h_in_occ[i]; v++) in_modes.push_back(i);
        for (int v = 2; v <= h_in_occ[i]; v++) pin *= v;
    }
    const int n = (int)in_modes.size();
    PCVL_ARG_CHECK(n >= 1 && n <= ADJ_MAX_N, "photon count outside [1, 8]");
    std::vector<double> tnorm(ntargets);
    for (int s = 0; s < ntargets; s++) {
        int ns = 0;
        double ps = 1.0;
        for (int i = 0; i < m; i++) {
            const int v = h_targets[s * m + i];
            PCVL_ARG_CHECK(v >= 0, "negative target occupation");
            ns += v;
            for (int k = 0; k < v; k++) tmodes.push_back(i);
            for (int k = 2; k <= v; k++) ps *= k;
        }
        PCVL_ARG_CHECK(ns == n, "target photon number differs from the input's");
        tnorm[s] = 1.0 / sqrt(ps * pin);
    }
    // levels: an op goes one level above the last op on any of its modes
    std::vector<int> last(m, 0), level(nops);
    int nlev = 0;
    for (int p = 0; p < nops; p++) {
        const int q = ops[p].q, q2 = ops[p].kind == 1 ? q + 1 : q;
        const int lv = std::max(last[q], last[q2]);
        level[p] = lv;
        last[q] = last[q2] = lv + 1;
        nlev = std::max(nlev, lv + 1);
    }
    AdjIdx X;
    X.nlev = nlev;
    X.glog = 0;
    while ((1 << X.glog) < m) X.glog++;
    std::vector<int32_t> idx;
    X.lop = (int)idx.size();
    std::vector<int32_t> lstart(nlev + 1, 0);
    for (int p = 0; p < nops; p++) lstart[level[p] + 1]++;
    for (int lv = 0; lv < nlev; lv++) lstart[lv + 1] += lstart[lv];
    {
        std::vector<int32_t> lop(nops), fill(lstart.begin(), lstart.end() - 1);
        for (int p = 0; p < nops; p++) lop[fill[level[p]]++] = p;   // circuit order within a level
        idx.insert(idx.end(), lop.begin(), lop.end());
    }
    X.lstart = (int)idx.size();
    idx.insert(idx.end(), lstart.begin(), lstart.end());
    // per block of 64 targets and mode i the (s, r) whose target row is i (rstart[blk * (m + 1) + i]),
    // per mode j the c whose input column is j
    X.rstart = (int)idx.size();
    const int nblk = (ntargets + ADJ_BLOCK_TARGETS - 1) / ADJ_BLOCK_TARGETS;
    std::vector<int32_t> rstart((size_t)nblk * (m + 1), 0), rlist;
    for (int b = 0; b < nblk; b++) {
        const int e0 = b * ADJ_BLOCK_TARGETS * n, e1 = std::min(ntargets, (b + 1) * ADJ_BLOCK_TARGETS) * n;
        for (int i = 0; i < m; i++) {
            rstart[(size_t)b * (m + 1) + i] = (int32_t)rlist.size();
            for (int e = e0; e < e1; e++)
                if (tmodes[(size_t)e] == i) rlist.push_back(e);
        }
        rstart[(size_t)b * (m + 1) + m] = (int32_t)rlist.size();
    }
    idx.insert(idx.end(), rstart.begin(), rstart.end());
    X.rlist = (int)idx.size();
    idx.insert(idx.end(), rlist.begin(), rlist.end());
    X.cstart = (int)idx.size();
    std::vector<int32_t> cstart(m + 1, 0), clist;
    for (int j = 0; j < m; j++) {
        cstart[j] = (int32_t)clist.size();
        for (int c = 0; c < n; c++)
            if (in_modes[(size_t)c] == j) clist.push_back(c);
    }
    cstart[m] = (int32_t)clist.size();
    idx.insert(idx.end(), cstart.begin(), cstart.end());
    X.clist = (int)idx.size();
    idx.insert(idx.end(), clist.begin(), clist.end());
    X.tm = (int)idx.size();
    idx.insert(idx.end(), tmodes.begin(), tmodes.end());
    X.in = (int)idx.size();
    idx.insert(idx.end(), in_modes.begin(), in_modes.end());
    X.total = (int)idx.size();
    // the program goes to LDS when it fits beside the rest (64 KiB budget), else
    // it is read from global memory
    const bool stage_ops = adj_lds(nops, nparams, m, n, ntargets, X.total).bytes <= 64 * 1024;
    const AdjLds L = adj_lds(stage_ops ? nops : 0, nparams, m, n, ntargets, X.total);
    if (L.bytes > 160 * 1024) {   // a caller may split the targets over several plans
        set_error("too many phases or targets for the kernel's LDS (" + std::to_string(L.bytes) + " bytes)");
        return ERR_RANGE;
    }
    auto *plan = new (std::nothrow) pcvl_mesh_plan_s();
    if (!plan) { set_error("host allocation failed"); return ERR_ALLOC; }
    plan->m = m; plan->n = n; plan->nops = nops; plan->nparams = nparams; plan->ntargets = ntargets;
    plan->X = X;
    plan->stage_ops = stage_ops;
    plan->lds = L.bytes;
    // one device block: program, index tables, norms
    const size_t b_ops = adj_al16(sizeof(MeshOp) * std::max(1, nops));
    const size_t b_idx = adj_al16(sizeof(int32_t) * idx.size()), b_n = adj_al16(sizeof(double) * ntargets);
    hipError_t e = hipMalloc((void **)&plan->d_blk, b_ops + b_idx + 2 * b_n);
    if (e != hipSuccess) {
        delete plan;
        set_error(std::string("hipMalloc(mesh plan): ") + hipGetErrorString(e));
        return ERR_ALLOC;
    }
    plan->d_ops = reinterpret_cast<const MeshOp *>(plan->d_blk);
    plan->d_idx = reinterpret_cast<const int32_t *>(plan->d_blk + b_ops);
    plan->d_tn = reinterpret_cast<const double *>(plan->d_blk + b_ops + b_idx);
    plan->d_tw = reinterpret_cast<double *>(plan->d_blk + b_ops + b_idx + b_n);
    const std::vector<double> ones((size_t)ntargets, 1.0);
    if (nops) e = hipMemcpy(plan->d_blk, ops.data(), sizeof(MeshOp) * nops, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(plan->d_blk + b_ops, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(plan->d_blk + b_ops + b_idx, tnorm.data(), sizeof(double) * ntargets,
                                       hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(plan->d_tw, ones.data(), sizeof(double) * ntargets, hipMemcpyHostToDevice);
    int dev = 0, cus = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) {
        (void)hipFree(plan->d_blk);
        delete plan;
        set_error(std::string("mesh plan upload: ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    plan->grid_cap = cus * 4;
    *out = plan;
    return OK;
}

extern "C" int pcvl_mesh_plan_execute(pcvl_mesh_plan plan, const double *d_phases, int nsets, double *d_loss,
                                      double *d_grad, double *d_probs, hipStream_t stream)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    PCVL_ARG_CHECK(nsets >= 0, "negative set count");
    if (nsets == 0) return OK;
    PCVL_ARG_CHECK(d_phases && d_loss && d_grad, "null device buffer");
    const pcvl_mesh_plan_s &P = *plan;
    const int grid = std::min(nsets, P.grid_cap);
    hipLaunchKernelGGL(mesh_adjoint_kernel, dim3(grid), dim3(ADJ_THREADS), P.lds, stream, P.d_ops, P.nops, P.m,
                       d_phases, P.nparams, nsets, P.n, P.d_idx, P.X, P.d_tn, P.d_tw, P.ntargets, d_loss, d_grad,
                       d_probs, P.stage_ops ? 1 : 0);
    PCVL_HIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int pcvl_mesh_plan_set_weights(pcvl_mesh_plan plan, const double *h_weights)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    std::vector<double> w((size_t)plan->ntargets, 1.0);
    if (h_weights)
        for (int s = 0; s < plan->ntargets; s++) {
            PCVL_ARG_CHECK(std::isfinite(h_weights[s]), "target weights must be finite");
            w[(size_t)s] = h_weights[s];
        }
    PCVL_HIP_CHECK(hipMemcpy(plan->d_tw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    return OK;
}

extern "C" int pcvl_mesh_plan_destroy(pcvl_mesh_plan plan)
{
    if (!plan) return OK;
    hipError_t e = plan->d_blk ? hipFree(plan->d_blk) : hipSuccess;
    delete plan;
    if (e != hipSuccess) {
        set_error(std::string("hipFree(mesh plan): ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    return OK;
}

extern "C" int pcvl_mesh_adjoint(const int32_t *h_kind, const int32_t *h_mode, const int32_t *h_param,
                                 const double *h_const, int nops, int m, const double *d_phases, int nparams,
                                 int nsets, const int32_t *h_in_occ, const int32_t *h_targets, int ntargets,
                                 double *d_loss, double *d_grad, double *d_probs, hipStream_t stream)
{
    PCVL_ARG_CHECK(nsets >= 0 && nparams >= 0, "negative sizes");
    pcvl_mesh_plan plan = nullptr;
    int rc = pcvl_mesh_plan_create(h_kind, h_mode, h_param, h_const, nops, m, nparams, h_in_occ, h_targets, ntargets,
                                   &plan);
    if (rc != OK) return rc;
    rc = pcvl_mesh_plan_execute(plan, d_phases, nsets, d_loss, d_grad, d_probs, stream);
    if (rc == OK) {
        const hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
            rc = ERR_HIP;
        }
    }
    const int rc2 = pcvl_mesh_plan_destroy(plan);
    return rc != OK ? rc : rc2;
}
