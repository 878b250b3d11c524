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
#include "../../include/pcvl_hip.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

namespace pcvl {

constexpr int SLOS_MAX_MODES = 64;
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
            reinterpret_cast<double *>(out)[idx] = (acc.r * acc.r + acc.i * acc.i) * pfact * inv_prod_in;
        }
    }
}

// Writes the vacuum layer: coefficient (1, 0), or probability 1.
__global__ void slos_vacuum_kernel(double *dst, int as_prob)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dst[0] = 1.0;
        if (!as_prob) dst[1] = 0.0;
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

struct SlosPlanImpl {
    int m = 0, n = 0, w = 0;
    std::vector<int32_t> mk;          // input mode of each added photon
    std::vector<uint64_t> counts;     // layer sizes N_0..N_n
    double inv_prod_in = 1.0;         // 1 / prod t!
    uint64_t max_inter = 1;           // largest intermediate layer (k < n)
    uint64_t *d_T = nullptr, *d_E = nullptr;
    int grid_cap = 0;
};

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
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 64]");
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
    std::vector<uint64_t> T, E;
    build_tables(m, n, T, E);
    size_t bytes = sizeof(uint64_t) * T.size();
    hipError_t e = hipMalloc(&p.d_T, 2 * bytes);
    if (e != hipSuccess) { delete plan; set_error(std::string("hipMalloc(slos tables): ") + hipGetErrorString(e)); return ERR_ALLOC; }
    p.d_E = p.d_T + T.size();
    e = hipMemcpy(p.d_T, T.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.d_E, E.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p.d_T);
        delete plan;
        set_error(std::string("hipMemcpy(slos tables): ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    int rc = device_grid_cap(&p.grid_cap);
    if (rc != OK) { (void)hipFree(p.d_T); delete plan; return rc; }
    *out = plan;
    return OK;
}

extern "C" int pcvl_slos_plan_info(pcvl_slos_plan plan, int *n, int64_t *out_count, size_t *work_bytes,
                                   int32_t *h_path)
{
    PCVL_ARG_CHECK(plan != nullptr, "plan is null");
    const SlosPlanImpl &p = plan->impl;
    if (n) *n = p.n;
    if (out_count) *out_count = (int64_t)p.counts[p.n];
    // two ping-pong buffers of the largest intermediate layer (complex)
    if (work_bytes) *work_bytes = p.n >= 2 ? 2 * p.max_inter * sizeof(cd) : sizeof(cd) * 2;
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
    size_t need = 0;
    pcvl_slos_plan_info(plan, nullptr, nullptr, &need, nullptr);
    PCVL_ARG_CHECK(d_work != nullptr && work_bytes >= need, "workspace missing or smaller than pcvl_slos_plan_info reports");
    const cd *U = reinterpret_cast<const cd *>(d_U);
    cd *bufA = reinterpret_cast<cd *>(d_work);
    cd *bufB = bufA + p.max_inter;
    const size_t lds = 2 * sizeof(uint64_t) * (size_t)p.m * p.w;

    if (p.n == 0) {
        // Vacuum: one output state with coefficient 1 (probability 1).
        hipLaunchKernelGGL(slos_vacuum_kernel, dim3(1), dim3(64), 0, stream, d_out, mode == PCVL_SLOS_PROBS ? 1 : 0);
        PCVL_HIP_CHECK(hipGetLastError());
        return OK;
    }
    // Layer 0 is the vacuum coefficient 1; layer 1 reads it from bufA[0].
    hipLaunchKernelGGL(slos_vacuum_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<double *>(bufA), 0);
    PCVL_HIP_CHECK(hipGetLastError());
    const cd *prev = bufA;
    for (int k = 1; k <= p.n; k++) {
        const uint64_t nc = p.counts[k];
        const int grid = (int)std::min<uint64_t>((nc + SLOS_THREADS - 1) / SLOS_THREADS, (uint64_t)p.grid_cap);
        const bool last = (k == p.n);
        void *dst = last ? static_cast<void *>(d_out) : static_cast<void *>(prev == bufA ? bufB : bufA);
        if (last && mode == PCVL_SLOS_PROBS) {
            hipLaunchKernelGGL(slos_layer_kernel<1>, dim3(grid), dim3(SLOS_THREADS), lds, stream, prev, p.counts[k - 1],
                               dst, U, p.m, p.mk[k - 1], k, nc, p.d_T, p.d_E, p.w, p.inv_prod_in);
        } else {
            hipLaunchKernelGGL(slos_layer_kernel<0>, dim3(grid), dim3(SLOS_THREADS), lds, stream, prev, p.counts[k - 1],
                               dst, U, p.m, p.mk[k - 1], k, nc, p.d_T, p.d_E, p.w, p.inv_prod_in);
        }
        PCVL_HIP_CHECK(hipGetLastError());
        prev = reinterpret_cast<const cd *>(dst);
    }
    return OK;
}

extern "C" int pcvl_slos_plan_destroy(pcvl_slos_plan plan)
{
    if (!plan) return OK;
    hipError_t e = hipSuccess;
    if (plan->impl.d_T) e = hipFree(plan->impl.d_T);
    delete plan;
    if (e != hipSuccess) { set_error(std::string("hipFree(slos plan): ") + hipGetErrorString(e)); return ERR_HIP; }
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
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 64]");
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
    PCVL_ARG_CHECK(m >= 1 && m <= SLOS_MAX_MODES, "mode count outside [1, 64]");
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
