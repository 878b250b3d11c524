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

#include <vector>

namespace pcvl {

constexpr int ADJ_MAX_M = 32;
constexpr int ADJ_MAX_N = 8;
constexpr int ADJ_MAX_TARGETS = 64;
constexpr int ADJ_THREADS = 256;

struct MeshOp {
    int32_t kind;    // 0: phase shifter on mode q; 1: fixed 2x2 block on (q, q+1)
    int32_t q;
    int32_t param;   // phase index (kind 0), -1 for a constant phase
    int32_t pad;
    double c[8];     // kind 0: c[0] = constant phase; kind 1: 2x2 block row-major (re, im)
};

struct AdjShared {
    cd U[ADJ_MAX_M * ADJ_MAX_M];
    cd Z[ADJ_MAX_M * ADJ_MAX_M];
    cd amp[ADJ_MAX_TARGETS];
    cd cof[ADJ_MAX_TARGETS * ADJ_MAX_N * ADJ_MAX_N];
    double red[ADJ_THREADS];
};

__device__ __forceinline__ cd cexpi(double ph) { return {cos(ph), sin(ph)}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.r, -a.i}; }

// Glynn with derivatives for one target row r of an n x n matrix M (row-major
// in registers of the caller): returns perm(M) (from row r == 0's thread) and
// d perm / d M[r][c] for c < n.
template <int N>
__device__ void glynn_with_row_derivs(const cd *M, int r, cd &perm, cd *dr)
{
    cd cs[N];
    for (int c = 0; c < N; c++) {
        cd s = {0.0, 0.0};
        for (int i = 0; i < N; i++) s = cadd(s, M[i * N + c]);
        cs[c] = s;
    }
    perm = {0.0, 0.0};
    for (int c = 0; c < N; c++) dr[c] = {0.0, 0.0};
    uint32_t g = 0;
    for (uint32_t t = 0; t < (1u << (N - 1)); t++) {
        if (t) {
            const int b = __builtin_ctz(t);
            g ^= 1u << b;
            const double s = ((g >> b) & 1) ? -2.0 : 2.0;
            for (int c = 0; c < N; c++) {
                cs[c].r = fma(s, M[(b + 1) * N + c].r, cs[c].r);
                cs[c].i = fma(s, M[(b + 1) * N + c].i, cs[c].i);
            }
        }
        const double sign = (t & 1) ? -1.0 : 1.0;
        // delta_r of this term (row 0 always +1)
        const double dr_sign = (r == 0) ? 1.0 : (((g >> (r - 1)) & 1) ? -1.0 : 1.0);
        cd P[N];
        P[0] = {1.0, 0.0};
        for (int c = 1; c < N; c++) P[c] = cmul(P[c - 1], cs[c - 1]);
        cd S = {sign, 0.0};
        for (int c = N - 1; c >= 0; c--) {
            const cd ex = cmul(P[c], S);      // sign * prod_{c' != c} cs_c'
            dr[c].r = fma(dr_sign, ex.r, dr[c].r);
            dr[c].i = fma(dr_sign, ex.i, dr[c].i);
            S = cmul(S, cs[c]);
        }
        perm = cadd(perm, S);                 // S = sign * prod_c cs_c
    }
    const double sc = ldexp(1.0, -(N - 1));
    perm = {perm.r * sc, perm.i * sc};
    for (int c = 0; c < N; c++) dr[c] = {dr[c].r * sc, dr[c].i * sc};
}

template <int N>
__device__ __forceinline__ void dispatch_glynn(int n, const cd *M, int r, cd &perm, cd *dr)
{
    if constexpr (N <= ADJ_MAX_N) {
        if (n == N) { glynn_with_row_derivs<N>(M, r, perm, dr); return; }
        dispatch_glynn<N + 1>(n, M, r, perm, dr);
    }
}

__global__ void __launch_bounds__(ADJ_THREADS) mesh_adjoint_kernel(
    const MeshOp *__restrict__ ops, int nops, int m, const double *__restrict__ phases, int nparams, int nsets,
    const int32_t *__restrict__ in_modes, int n, const int32_t *__restrict__ target_modes,
    const double *__restrict__ target_norm, int ntargets, double *__restrict__ loss_out, double *__restrict__ grad_out,
    double *__restrict__ probs_out)
{
    __shared__ AdjShared sh;
    const int tid = threadIdx.x;
    for (int set = blockIdx.x; set < nsets; set += gridDim.x) {
        const double *ph = phases + (size_t)set * nparams;
        // 1. U = L_K ... L_1: start from the identity, apply ops in circuit order
        for (int e = tid; e < m * m; e += ADJ_THREADS) sh.U[e] = {(e / m == e % m) ? 1.0 : 0.0, 0.0};
        __syncthreads();
        for (int p = 0; p < nops; p++) {
            const MeshOp op = ops[p];
            if (op.kind == 0) {
                const cd z = cexpi(op.param >= 0 ? ph[op.param] : op.c[0]);
                if (tid < m) sh.U[op.q * m + tid] = cmul(z, sh.U[op.q * m + tid]);
            } else if (tid < m) {
                const cd a = {op.c[0], op.c[1]}, b = {op.c[2], op.c[3]}, c = {op.c[4], op.c[5]}, d = {op.c[6], op.c[7]};
                const cd x = sh.U[op.q * m + tid], y = sh.U[(op.q + 1) * m + tid];
                sh.U[op.q * m + tid] = cadd(cmul(a, x), cmul(b, y));
                sh.U[(op.q + 1) * m + tid] = cadd(cmul(c, x), cmul(d, y));
            }
            __syncthreads();
        }
        // 2. amplitudes and cofactors: thread (target s, row r)
        for (int w = tid; w < ntargets * n; w += ADJ_THREADS) {
            const int s = w / n, r = w % n;
            cd M[ADJ_MAX_N * ADJ_MAX_N];
            for (int i = 0; i < n; i++)
                for (int c = 0; c < n; c++) M[i * n + c] = sh.U[target_modes[s * n + i] * m + in_modes[c]];
            cd perm, dr[ADJ_MAX_N];
            dispatch_glynn<1>(n, M, r, perm, dr);
            const double nrm = target_norm[s];   // 1 / sqrt(prod s! prod t!)
            for (int c = 0; c < n; c++) sh.cof[(s * n + r) * n + c] = {dr[c].r * nrm, dr[c].i * nrm};
            if (r == 0) sh.amp[s] = {perm.r * nrm, perm.i * nrm};
        }
        __syncthreads();
        // loss and per-target probabilities
        if (tid < ntargets) {
            const cd a = sh.amp[tid];
            const double pr = a.r * a.r + a.i * a.i;
            if (probs_out) probs_out[(size_t)set * ntargets + tid] = pr;
            sh.red[tid] = pr;
        }
        __syncthreads();
        if (tid == 0) {
            double l = 0.0;
            for (int s = 0; s < ntargets; s++) l += sh.red[s];
            loss_out[set] = l;
        }
        // G^T into Z: Z[j][i] = G[i][j] = sum_s conj(A_s) sum_{(r,c): rows_s[r]=i, in[c]=j} cof_src
        for (int e = tid; e < m * m; e += ADJ_THREADS) {
            const int j = e / m, i = e % m;
            cd g = {0.0, 0.0};
            for (int c = 0; c < n; c++) {
                if (in_modes[c] != j) continue;
                for (int s = 0; s < ntargets; s++) {
                    const cd ca = cconj(sh.amp[s]);
                    for (int r = 0; r < n; r++)
                        if (target_modes[s * n + r] == i) cfma_acc(g, ca, sh.cof[(s * n + r) * n + c]);
                }
            }
            sh.Z[j * m + i] = g;
        }
        __syncthreads();
        // 3. backward sweep: F (in sh.U) <- L_p^H F, grad, Z <- Z L_p
        for (int p = nops - 1; p >= 0; p--) {
            const MeshOp op = ops[p];
            if (op.kind == 0) {
                const double phi = op.param >= 0 ? ph[op.param] : op.c[0];
                const cd z = cexpi(phi);
                if (tid < m) sh.U[op.q * m + tid] = cmul(cconj(z), sh.U[op.q * m + tid]);   // F <- L^H F
                __syncthreads();
                if (op.param >= 0) {
                    // (F Z)_qq = sum_k F[q][k] Z[k][q]
                    double re = 0.0, im = 0.0;
                    if (tid < m) {
                        const cd v = cmul(sh.U[op.q * m + tid], sh.Z[tid * m + op.q]);
                        re = v.r; im = v.i;
                    }
                    sh.red[tid] = re;
                    __syncthreads();
                    if (tid == 0) {
                        double sr = 0.0, si = 0.0;
                        for (int k = 0; k < m; k++) sr += sh.red[k];
                        sh.red[ADJ_THREADS - 1] = sr;
                        (void)si;
                    }
                    __syncthreads();
                    const double fr = sh.red[ADJ_THREADS - 1];
                    __syncthreads();
                    sh.red[tid] = im;
                    __syncthreads();
                    if (tid == 0) {
                        double si = 0.0;
                        for (int k = 0; k < m; k++) si += sh.red[k];
                        // dL/dphi = 2 Re(i z (F Z)_qq) = -2 (z.r * si + z.i * fr)
                        grad_out[(size_t)set * nparams + op.param] = -2.0 * (z.r * si + z.i * fr);
                    }
                    __syncthreads();
                }
                if (tid < m) sh.Z[tid * m + op.q] = cmul(sh.Z[tid * m + op.q], z);             // Z <- Z L
            } else {
                const cd a = {op.c[0], op.c[1]}, b = {op.c[2], op.c[3]}, c = {op.c[4], op.c[5]}, d = {op.c[6], op.c[7]};
                if (tid < m) {
                    // F <- L^H F on rows q, q+1: L^H = [[a*, c*], [b*, d*]]
                    const cd x = sh.U[op.q * m + tid], y = sh.U[(op.q + 1) * m + tid];
                    sh.U[op.q * m + tid] = cadd(cmul(cconj(a), x), cmul(cconj(c), y));
                    sh.U[(op.q + 1) * m + tid] = cadd(cmul(cconj(b), x), cmul(cconj(d), y));
                    // Z <- Z L on columns q, q+1
                    const cd u = sh.Z[tid * m + op.q], v = sh.Z[tid * m + op.q + 1];
                    sh.Z[tid * m + op.q] = cadd(cmul(u, a), cmul(v, c));
                    sh.Z[tid * m + op.q + 1] = cadd(cmul(u, b), cmul(v, d));
                }
            }
            __syncthreads();
        }
    }
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_mesh_adjoint(const int32_t *h_kind, const int32_t *h_mode, const int32_t *h_param,
                                 const double *h_const, int nops, int m, const double *d_phases, int nparams,
                                 int nsets, const int32_t *h_in_occ, const int32_t *h_targets, int ntargets,
                                 double *d_loss, double *d_grad, double *d_probs, hipStream_t stream)
{
    PCVL_ARG_CHECK(m >= 1 && m <= ADJ_MAX_M, "mode count outside [1, 32]");
    PCVL_ARG_CHECK(nops >= 0 && (nops == 0 || (h_kind && h_mode && h_param && h_const)), "bad mesh program");
    PCVL_ARG_CHECK(nsets >= 0 && nparams >= 0, "negative sizes");
    PCVL_ARG_CHECK(ntargets >= 1 && ntargets <= ADJ_MAX_TARGETS, "target count outside [1, 64]");
    PCVL_ARG_CHECK(h_in_occ && h_targets, "null input/target states");
    if (nsets == 0) return OK;
    PCVL_ARG_CHECK(d_phases && d_loss && d_grad, "null device buffer");
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
    // small device copies of the program and index tables
    const size_t b_ops = sizeof(MeshOp) * std::max(1, nops);
    const size_t b_in = sizeof(int32_t) * n, b_t = sizeof(int32_t) * tmodes.size(), b_n = sizeof(double) * ntargets;
    char *blk = nullptr;
    PCVL_HIP_CHECK(hipMallocAsync((void **)&blk, b_ops + b_in + b_t + b_n + 64, stream));
    MeshOp *d_ops = reinterpret_cast<MeshOp *>(blk);
    int32_t *d_in = reinterpret_cast<int32_t *>(blk + b_ops);
    int32_t *d_tm = reinterpret_cast<int32_t *>(blk + b_ops + b_in);
    double *d_tn = reinterpret_cast<double *>(blk + ((b_ops + b_in + b_t + 7) & ~size_t(7)));
    if (nops) PCVL_HIP_CHECK(hipMemcpyAsync(d_ops, ops.data(), sizeof(MeshOp) * nops, hipMemcpyHostToDevice, stream));
    PCVL_HIP_CHECK(hipMemcpyAsync(d_in, in_modes.data(), b_in, hipMemcpyHostToDevice, stream));
    PCVL_HIP_CHECK(hipMemcpyAsync(d_tm, tmodes.data(), b_t, hipMemcpyHostToDevice, stream));
    PCVL_HIP_CHECK(hipMemcpyAsync(d_tn, tnorm.data(), b_n, hipMemcpyHostToDevice, stream));
    int dev = 0, cus = 0;
    PCVL_HIP_CHECK(hipGetDevice(&dev));
    PCVL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::min(nsets, cus * 4);
    hipLaunchKernelGGL(mesh_adjoint_kernel, dim3(grid), dim3(ADJ_THREADS), 0, stream, d_ops, nops, m, d_phases,
                       nparams, nsets, d_in, n, d_tm, d_tn, ntargets, d_loss, d_grad, d_probs);
    PCVL_HIP_CHECK(hipGetLastError());
    // host vectors above are the sources of the async copies
    PCVL_HIP_CHECK(hipStreamSynchronize(stream));
    PCVL_HIP_CHECK(hipFreeAsync(blk, stream));
    return OK;
}
