/*
 * pcvl_oracle.c -- CPU oracle for the Perceval simulation-core hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (perceval_amd) never calls
 * into it.
 *
 * The reference keeps every piece of arithmetic of this path inside the closed
 * exqalibur~=1.1.0 binary (setup.py:77), which is absent, so each function
 * below restates the published algorithm the reference calls, and cites the
 * reference call site it stands in for:
 *
 *   oracle_perm_glynn_ld / _d, oracle_perm_ryser_ld
 *       xq.permanent_cx (perceval/backends/_naive.py:70-71); Glynn 2010 and
 *       Ryser 1963 as cited in docs/source/reference/backends/naive.rst.
 *   oracle_fock_count / _rank / _unrank / _enumerate
 *       xq.FSArray(m, n) iteration order and FSArray.find
 *       (perceval/utils/states.py:255-298, perceval/backends/_slos.py:190);
 *       descending-lexicographic order as pinned by
 *       tests/utils/test_statevector.py:430-438.
 *   oracle_slos_scatter / oracle_slos_gather / oracle_slos_path
 *       _Path._decompose + FSMap.compute_slos_layer
 *       (perceval/backends/_slos.py:61-102; symbolic restatement :91-97) and
 *       xq.all_prob_normalize_output (_slos.py:211-214).
 *   oracle_cc_samples
 *       xq.Clifford2017.samples (perceval/backends/_clifford2017.py:52-57);
 *       algorithm A of Clifford & Clifford 2017 (docs/source/backends.rst:108-116).
 *
 * Parallel loops use plain POSIX threads (or_parallel), never OpenMP.
 *
 * Parity pins for each function live in tests/test_oracle.py (golden vectors
 * from the reference's own tests and docs, mpmath goldens, closed forms and
 * brute force).
 *
 * Conventions: complex numbers are interleaved (re, im) doubles; matrices are
 * row-major; unitaries are indexed [output mode][input mode] as _umat[ok, ik]
 * in perceval/backends/_naive.py:65.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include <pthread.h>
#include <unistd.h>

typedef long double ld;

#define OR_MAXN 40    /* largest permanent order the oracle accepts */
#define OR_MAXM 128   /* largest mode count the oracle accepts */


/* ------------------------------------------------------------------------ */
/* Fork-join over plain POSIX threads (no OpenMP runtime, so the oracle can  */
/* be loaded into processes that already carry another OpenMP runtime).     */
/* ------------------------------------------------------------------------ */

typedef void (*or_task_fn)(void *ctx, int64_t lo, int64_t hi);

typedef struct {
    or_task_fn fn;
    void *ctx;
    int64_t total, grain;
    int64_t next;           /* next unclaimed index, taken under `lock` */
    pthread_mutex_t lock;
} or_pool;

static void *or_worker(void *arg)
{
    or_pool *p = (or_pool *)arg;
    for (;;) {
        pthread_mutex_lock(&p->lock);
        int64_t lo = p->next;
        p->next = lo + p->grain;
        pthread_mutex_unlock(&p->lock);
        if (lo >= p->total) break;
        int64_t hi = lo + p->grain < p->total ? lo + p->grain : p->total;
        p->fn(p->ctx, lo, hi);
    }
    return NULL;
}

static int or_default_threads(void)
{
    const char *e = getenv("OMP_NUM_THREADS");
    int t = e ? atoi(e) : 0;
    if (t <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        t = c > 0 ? (int)c : 1;
    }
    return t > 256 ? 256 : t;
}

/* Run fn over [0, total) in chunks of `grain`, dynamically claimed by
 * `threads` workers (the caller is one of them). */
static void or_parallel(or_task_fn fn, void *ctx, int64_t total, int64_t grain, int threads)
{
    if (threads <= 0) threads = or_default_threads();
    if (grain < 1) grain = 1;
    if ((int64_t)threads > (total + grain - 1) / grain) threads = (int)((total + grain - 1) / grain);
    if (threads <= 1) { if (total > 0) fn(ctx, 0, total); return; }
    or_pool p;
    p.fn = fn; p.ctx = ctx; p.total = total; p.grain = grain; p.next = 0;
    pthread_mutex_init(&p.lock, NULL);
    pthread_t tid[256];
    int started = 0;
    for (int t = 1; t < threads; t++) {
        if (pthread_create(&tid[started], NULL, or_worker, &p) == 0) started++;
    }
    or_worker(&p);
    for (int t = 0; t < started; t++) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&p.lock);
}

/* ------------------------------------------------------------------------ */
/* Permanents                                                                */
/* ------------------------------------------------------------------------ */

/* Glynn's formula, perm(A) = 2^{-(n-1)} sum_delta (prod delta) prod_j sum_i delta_i a_ij
 * with delta_0 = +1, summed in binary-reflected Gray-code order, long double
 * throughout (64-bit mantissa).  Step t flips row 1 + ctz(t); the sign of the
 * term of step t is (-1)^t. */
static void glynn_ld(const double *a, int n, ld *out_re, ld *out_im)
{
    if (n == 0) { *out_re = 1; *out_im = 0; return; }
    ld cr[OR_MAXN], ci[OR_MAXN];
    for (int j = 0; j < n; j++) {
        ld sr = 0, si = 0;
        for (int i = 0; i < n; i++) { sr += a[2 * (i * n + j)]; si += a[2 * (i * n + j) + 1]; }
        cr[j] = sr; ci[j] = si;
    }
    ld accr = 0, acci = 0;
    uint64_t steps = 1ull << (n - 1), g = 0;
    for (uint64_t t = 0; t < steps; t++) {
        if (t) {
            int b = __builtin_ctzll(t);
            int r = b + 1;
            g ^= 1ull << b;
            ld s = ((g >> b) & 1) ? -2.0L : 2.0L;
            const double *row = a + 2 * r * n;
            for (int j = 0; j < n; j++) { cr[j] += s * row[2 * j]; ci[j] += s * row[2 * j + 1]; }
        }
        ld pr = cr[0], pi = ci[0];
        for (int j = 1; j < n; j++) {
            ld nr = pr * cr[j] - pi * ci[j];
            ld ni = pr * ci[j] + pi * cr[j];
            pr = nr; pi = ni;
        }
        if (t & 1) { accr -= pr; acci -= pi; } else { accr += pr; acci += pi; }
    }
    ld scale = ldexpl(1.0L, -(n - 1));
    *out_re = accr * scale; *out_im = acci * scale;
}

/* Ryser's formula, perm(A) = (-1)^n sum_{S subset cols} (-1)^{|S|} prod_i sum_{j in S} a_ij,
 * Gray-code over column subsets, long double.  Independent of Glynn; used to
 * cross-check it. */
static void ryser_ld(const double *a, int n, ld *out_re, ld *out_im)
{
    if (n == 0) { *out_re = 1; *out_im = 0; return; }
    ld rr[OR_MAXN], ri[OR_MAXN];
    for (int i = 0; i < n; i++) { rr[i] = 0; ri[i] = 0; }
    ld accr = 0, acci = 0;
    uint64_t total = 1ull << n, g = 0;
    for (uint64_t t = 1; t < total; t++) {
        int b = __builtin_ctzll(t);
        g ^= 1ull << b;
        ld s = ((g >> b) & 1) ? 1.0L : -1.0L;
        for (int i = 0; i < n; i++) { rr[i] += s * a[2 * (i * n + b)]; ri[i] += s * a[2 * (i * n + b) + 1]; }
        ld pr = rr[0], pi = ri[0];
        for (int i = 1; i < n; i++) {
            ld nr = pr * rr[i] - pi * ri[i];
            ld ni = pr * ri[i] + pi * rr[i];
            pr = nr; pi = ni;
        }
        if (__builtin_popcountll(g) & 1) { accr -= pr; acci -= pi; } else { accr += pr; acci += pi; }
    }
    if (n & 1) { accr = -accr; acci = -acci; }
    *out_re = accr; *out_im = acci;
}

/* Glynn in double precision, split into independent chunks of 2^c Gray-code
 * steps whose column sums are recomputed from scratch at the chunk start.
 * This is the CPU restatement timed as the baseline; the chunking bounds the
 * drift of the running column sums. */
static void glynn_d_chunk(const double *a, int n, uint64_t base, uint64_t len, double *sr_out, double *si_out)
{
    double cr[OR_MAXN], ci[OR_MAXN];
    uint64_t g = base ^ (base >> 1);
    for (int j = 0; j < n; j++) {
        double sr = a[2 * j], si = a[2 * j + 1];
        for (int i = 1; i < n; i++) {
            double d = ((g >> (i - 1)) & 1) ? -1.0 : 1.0;
            sr += d * a[2 * (i * n + j)]; si += d * a[2 * (i * n + j) + 1];
        }
        cr[j] = sr; ci[j] = si;
    }
    double accr = 0, acci = 0;
    for (uint64_t u = 0; u < len; u++) {
        uint64_t t = base + u;
        if (u) {
            int b = __builtin_ctzll(t);
            g ^= 1ull << b;
            double s = ((g >> b) & 1) ? -2.0 : 2.0;
            const double *row = a + 2 * (b + 1) * n;
            for (int j = 0; j < n; j++) { cr[j] += s * row[2 * j]; ci[j] += s * row[2 * j + 1]; }
        }
        double pr = cr[0], pi = ci[0];
        for (int j = 1; j < n; j++) {
            double nr = pr * cr[j] - pi * ci[j];
            double ni = pr * ci[j] + pi * cr[j];
            pr = nr; pi = ni;
        }
        if (t & 1) { accr -= pr; acci -= pi; } else { accr += pr; acci += pi; }
    }
    *sr_out = accr; *si_out = acci;
}

static void glynn_d(const double *a, int n, double *out)
{
    if (n == 0) { out[0] = 1; out[1] = 0; return; }
    uint64_t steps = 1ull << (n - 1);
    uint64_t chunk = steps < (1ull << 14) ? steps : (1ull << 14);
    double sr = 0, si = 0;
    for (uint64_t base = 0; base < steps; base += chunk) {
        double cr, ci;
        glynn_d_chunk(a, n, base, chunk, &cr, &ci);
        sr += cr; si += ci;
    }
    double scale = ldexp(1.0, -(n - 1));
    out[0] = sr * scale; out[1] = si * scale;
}

/* Single permanents.  Return 0 on success, -1 on a bad order. */
int oracle_perm_glynn_ld(const double *a, int n, double *out)
{
    if (n < 0 || n > OR_MAXN) return -1;
    ld r, i;
    glynn_ld(a, n, &r, &i);
    out[0] = (double)r; out[1] = (double)i;
    return 0;
}

int oracle_perm_ryser_ld(const double *a, int n, double *out)
{
    if (n < 0 || n > OR_MAXN) return -1;
    ld r, i;
    ryser_ld(a, n, &r, &i);
    out[0] = (double)r; out[1] = (double)i;
    return 0;
}

int oracle_perm_glynn_d(const double *a, int n, double *out)
{
    if (n < 0 || n > OR_MAXN) return -1;
    glynn_d(a, n, out);
    return 0;
}

/* Long-double Glynn of one large matrix split over threads: the 2^(n-1)
 * Gray-code steps are cut into chunks of 2^16 steps (fewer for small n); a
 * chunk rebuilds its column sums in long double from its first Gray code,
 * then steps as glynn_ld does.  Chunk partials are summed in chunk order, so
 * the result does not depend on the thread count.  Same formula as glynn_ld
 * (_naive.py:70-71's xq.permanent_cx); for orders 25..28, where the
 * single-threaded form takes minutes. */
typedef struct { const double *a; int n; uint64_t chunk; ld *part; } glynn_mt_ctx;

static void glynn_ld_chunks(void *vctx, int64_t lo, int64_t hi)
{
    glynn_mt_ctx *c = (glynn_mt_ctx *)vctx;
    const double *a = c->a;
    int n = c->n;
    for (int64_t k = lo; k < hi; k++) {
        uint64_t base = (uint64_t)k * c->chunk;
        uint64_t g = base ^ (base >> 1);
        ld cr[OR_MAXN], ci[OR_MAXN];
        for (int j = 0; j < n; j++) {
            ld sr = a[2 * j], si = a[2 * j + 1];
            for (int i = 1; i < n; i++) {
                ld d = ((g >> (i - 1)) & 1) ? -1.0L : 1.0L;
                sr += d * a[2 * (i * n + j)]; si += d * a[2 * (i * n + j) + 1];
            }
            cr[j] = sr; ci[j] = si;
        }
        ld accr = 0, acci = 0;
        for (uint64_t u = 0; u < c->chunk; u++) {
            uint64_t t = base + u;
            if (u) {
                int b = __builtin_ctzll(t);
                g ^= 1ull << b;
                ld s = ((g >> b) & 1) ? -2.0L : 2.0L;
                const double *row = a + 2 * (b + 1) * n;
                for (int j = 0; j < n; j++) { cr[j] += s * row[2 * j]; ci[j] += s * row[2 * j + 1]; }
            }
            ld pr = cr[0], pi = ci[0];
            for (int j = 1; j < n; j++) {
                ld nr = pr * cr[j] - pi * ci[j];
                ld ni = pr * ci[j] + pi * cr[j];
                pr = nr; pi = ni;
            }
            if (t & 1) { accr -= pr; acci -= pi; } else { accr += pr; acci += pi; }
        }
        c->part[2 * k] = accr; c->part[2 * k + 1] = acci;
    }
}

int oracle_perm_glynn_ld_mt(const double *a, int n, double *out, int threads)
{
    if (n < 0 || n > OR_MAXN) return -1;
    if (n == 0) { out[0] = 1; out[1] = 0; return 0; }
    uint64_t steps = 1ull << (n - 1);
    uint64_t chunk = steps < (1ull << 16) ? steps : (1ull << 16);
    uint64_t nchunks = steps / chunk;
    ld *part = (ld *)malloc(sizeof(ld) * 2 * nchunks);
    if (!part) return -3;
    glynn_mt_ctx c = {a, n, chunk, part};
    or_parallel(glynn_ld_chunks, &c, (int64_t)nchunks, 1, threads);
    ld sr = 0, si = 0;
    for (uint64_t k = 0; k < nchunks; k++) { sr += part[2 * k]; si += part[2 * k + 1]; }
    free(part);
    ld scale = ldexpl(1.0L, -(n - 1));
    out[0] = (double)(sr * scale); out[1] = (double)(si * scale);
    return 0;
}

/* Batched permanents over a packed array: matrix b has order n[b] and starts
 * at complex element offsets[b].  precise=1 uses the long-double Glynn,
 * precise=0 the chunked double Glynn.  Threads over matrices. */
typedef struct {
    const double *mats; const int32_t *n; const int64_t *offsets; double *out; int precise;
} perm_batch_ctx;

static void perm_batch_task(void *vctx, int64_t lo, int64_t hi)
{
    perm_batch_ctx *c = (perm_batch_ctx *)vctx;
    for (int64_t b = lo; b < hi; b++) {
        const double *a = c->mats + 2 * c->offsets[b];
        if (c->precise) {
            ld r, i;
            glynn_ld(a, c->n[b], &r, &i);
            c->out[2 * b] = (double)r; c->out[2 * b + 1] = (double)i;
        } else {
            glynn_d(a, c->n[b], c->out + 2 * b);
        }
    }
}

int oracle_perm_batch(const double *mats, const int32_t *n, const int64_t *offsets,
                      int64_t count, double *out, int precise, int threads)
{
    for (int64_t b = 0; b < count; b++)
        if (n[b] < 0 || n[b] > OR_MAXN) return -1;
    perm_batch_ctx c = {mats, n, offsets, out, precise};
    or_parallel(perm_batch_task, &c, count, 1, threads);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fock-space enumeration in descending lexicographic order                  */
/* ------------------------------------------------------------------------ */

/* Binomial coefficient with C(a,b)=0 outside 0<=b<=a; saturates at UINT64_MAX. */
static uint64_t binom(int64_t a, int64_t b)
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

uint64_t oracle_binom(int64_t a, int64_t b) { return binom(a, b); }

/* Number of n-photon states over m modes: C(m+n-1, n). */
uint64_t oracle_fock_count(int m, int n)
{
    if (m <= 0 || n < 0) return (m == 0 && n == 0) ? 1 : 0;
    return binom((int64_t)m + n - 1, n);
}

/* rank(s) = sum_{i=0}^{m-2} C(r_i - s_i + m - i - 2, m - i - 1), r_i = n - sum_{j<i} s_j.
 * The i-th term counts the states that agree with s before mode i and hold
 * more photons than s in mode i, i.e. those that precede s in descending
 * lexicographic order. */
int64_t oracle_fock_rank(const int32_t *occ, int m)
{
    int64_t n = 0;
    for (int i = 0; i < m; i++) { if (occ[i] < 0) return -1; n += occ[i]; }
    int64_t r = n, rank = 0;
    for (int i = 0; i + 1 < m; i++) {
        rank += (int64_t)binom(r - occ[i] + m - i - 2, m - i - 1);
        r -= occ[i];
    }
    return rank;
}

/* Inverse of oracle_fock_rank for n-photon states over m modes. */
int oracle_fock_unrank(int64_t rank, int m, int n, int32_t *occ)
{
    if (m <= 0 || n < 0 || rank < 0 || (uint64_t)rank >= oracle_fock_count(m, n)) return -1;
    int64_t r = n, R = rank;
    for (int i = 0; i + 1 < m; i++) {
        int64_t v;
        for (v = r; v > 0; v--) {
            int64_t cnt = (int64_t)binom(r - v + m - i - 2, m - i - 2);
            if (R < cnt) break;
            R -= cnt;
        }
        occ[i] = (int32_t)v;
        r -= v;
    }
    occ[m - 1] = (int32_t)r;
    return 0;
}

/* Successor in descending lexicographic order: move one photon from the
 * rightmost occupied mode i < m-1 to mode i+1 and gather the tail there.
 * Returns 0 when s was the last state. */
static int fock_next(int32_t *s, int m)
{
    int i = m - 2;
    while (i >= 0 && s[i] == 0) i--;
    if (i < 0) return 0;
    int32_t tail = 0;
    for (int j = i + 1; j < m; j++) { tail += s[j]; s[j] = 0; }
    s[i] -= 1;
    s[i + 1] = tail + 1;
    return 1;
}

/* Write all states (count x m int32, row per state) in FSArray order. */
int oracle_fock_enumerate(int m, int n, int32_t *out, int64_t out_rows)
{
    uint64_t cnt = oracle_fock_count(m, n);
    if ((uint64_t)out_rows < cnt) return -1;
    int32_t s[OR_MAXM];
    if (m > OR_MAXM) return -1;
    for (int j = 0; j < m; j++) s[j] = 0;
    s[0] = n;
    int64_t k = 0;
    do {
        memcpy(out + k * m, s, sizeof(int32_t) * m);
        k++;
    } while (fock_next(s, m));
    return (uint64_t)k == cnt ? 0 : -2;
}

/* ------------------------------------------------------------------------ */
/* SLOS                                                                      */
/* ------------------------------------------------------------------------ */

/* Order in which the photons of a single input state are added, following
 * _Path._decompose (perceval/backends/_slos.py:61-86) for one target: at each
 * layer take the mode holding the most remaining photons, first such mode on
 * ties. */
int oracle_slos_path(const int32_t *in_occ, int m, int32_t *mk_out)
{
    int32_t rem[OR_MAXM];
    if (m > OR_MAXM) return -1;
    int n = 0;
    for (int i = 0; i < m; i++) { rem[i] = in_occ[i]; n += in_occ[i]; }
    for (int k = 0; k < n; k++) {
        int best = 0;
        for (int i = 1; i < m; i++) if (rem[i] > rem[best]) best = i;
        mk_out[k] = best;
        rem[best]--;
    }
    return n;
}

static inline void cfma(double *acc, const double *x, const double *y)
{
    double xr = x[0], xi = x[1], yr = y[0], yi = y[1];
    acc[0] += xr * yr - xi * yi;
    acc[1] += xr * yi + xi * yr;
}

/* One layer in the scatter form of FSMap.compute_slos_layer:
 *   coefs[idx(p + e_j)] += parent[p] * U[j, mk]  for all parents p, modes j,
 * parents visited in rank order, j ascending.  The child rank is
 * rank(p) + sum_{i<j} C(r_i - p_i + m - i - 2, m - i - 2). */
static void slos_layer_scatter(const double *U, int m, int mk, int kparent,
                               const double *parent, double *child, uint64_t nchild)
{
    memset(child, 0, sizeof(double) * 2 * nchild);
    int32_t p[OR_MAXM];
    for (int j = 0; j < m; j++) p[j] = 0;
    p[0] = kparent;
    uint64_t prank = 0;
    do {
        uint64_t off = 0;
        int64_t r = kparent;
        for (int j = 0; j < m; j++) {
            cfma(child + 2 * (prank + off), parent + 2 * prank, U + 2 * (j * m + mk));
            if (j + 1 < m) {
                off += binom(r - p[j] + m - j - 2, m - j - 2);
                r -= p[j];
            }
        }
        prank++;
    } while (fock_next(p, m));
}

/* One layer in the gather form: child c takes parent c - e_j for every
 * occupied mode j, whose rank is rank(c) - sum_{i<j} C(r_i - c_i + m - i - 3, m - i - 2).
 * Modes are visited j = m-1 .. 0 so that each child sums its parents in
 * increasing parent rank, the same order as the scatter form.  Threads over
 * contiguous child ranges. */
typedef struct {
    const double *U; int m, mk, kchild; const double *parent; double *child;
} gather_ctx;

static void gather_task(void *vctx, int64_t lo, int64_t hi)
{
    gather_ctx *g = (gather_ctx *)vctx;
    const int m = g->m;
    int32_t c[OR_MAXM];
    uint64_t E[OR_MAXM];
    oracle_fock_unrank(lo, m, g->kchild, c);
    for (int64_t idx = lo; idx < hi; idx++) {
        int64_t r = g->kchild;
        uint64_t pre = 0;
        for (int i = 0; i < m; i++) {
            E[i] = pre;
            if (i + 1 < m) {
                pre += binom(r - c[i] + m - i - 3, m - i - 2);
                r -= c[i];
            }
        }
        double acc[2] = {0.0, 0.0};
        for (int j = m - 1; j >= 0; j--) {
            if (c[j] > 0)
                cfma(acc, g->parent + 2 * ((uint64_t)idx - E[j]), g->U + 2 * (j * m + g->mk));
        }
        g->child[2 * idx] = acc[0]; g->child[2 * idx + 1] = acc[1];
        fock_next(c, m);
    }
}

static void slos_layer_gather(const double *U, int m, int mk, int kchild,
                              const double *parent, double *child, uint64_t nchild, int threads)
{
    gather_ctx g = {U, m, mk, kchild, parent, child};
    int64_t grain = (int64_t)(nchild / 64) + 1;
    or_parallel(gather_task, &g, (int64_t)nchild, grain, threads);
}

static double prodnfact(const int32_t *s, int m)
{
    double p = 1.0;
    for (int i = 0; i < m; i++) for (int v = 2; v <= s[i]; v++) p *= v;
    return p;
}

static int slos_run(const double *U, int m, const int32_t *in_occ, double *out_coefs,
                    int gather, int threads)
{
    if (m <= 0 || m > OR_MAXM) return -1;
    int32_t mk[256];
    int n = 0;
    for (int i = 0; i < m; i++) { if (in_occ[i] < 0) return -1; n += in_occ[i]; }
    if (n > 255) return -1;
    oracle_slos_path(in_occ, m, mk);
    uint64_t maxcnt = oracle_fock_count(m, n);
    double *bufA = (double *)malloc(sizeof(double) * 2 * maxcnt);
    double *bufB = (double *)malloc(sizeof(double) * 2 * maxcnt);
    if (!bufA || !bufB) { free(bufA); free(bufB); return -3; }
    bufA[0] = 1.0; bufA[1] = 0.0;
    double *prev = bufA, *cur = bufB;
    for (int k = 1; k <= n; k++) {
        uint64_t nc = oracle_fock_count(m, k);
        if (gather) slos_layer_gather(U, m, mk[k - 1], k, prev, cur, nc, threads);
        else slos_layer_scatter(U, m, mk[k - 1], k - 1, prev, cur, nc);
        double *t = prev; prev = cur; cur = t;
    }
    memcpy(out_coefs, prev, sizeof(double) * 2 * maxcnt);
    free(bufA); free(bufB);
    return 0;
}

/* Final-layer SLOS coefficients c_n(s) (amplitude = c * sqrt(prod s!/prod t!),
 * _slos.py:192), complex, FSArray order.  Scatter form, single thread. */
int oracle_slos_scatter(const double *U, int m, const int32_t *in_occ, double *out_coefs)
{
    return slos_run(U, m, in_occ, out_coefs, 0, 1);
}

/* Same coefficients, gather form, OpenMP. */
int oracle_slos_gather(const double *U, int m, const int32_t *in_occ, double *out_coefs, int threads)
{
    return slos_run(U, m, in_occ, out_coefs, 1, threads);
}

/* Probabilities |c|^2 * prod s! / prod t! over the final layer
 * (_slos.py:211-213 with xq.all_prob_normalize_output). */
int oracle_slos_probs(const double *coefs, int m, int n, const int32_t *in_occ, double *out_probs)
{
    double pin = prodnfact(in_occ, m);
    int32_t s[OR_MAXM];
    if (m > OR_MAXM) return -1;
    for (int j = 0; j < m; j++) s[j] = 0;
    s[0] = n;
    uint64_t k = 0;
    do {
        double re = coefs[2 * k], im = coefs[2 * k + 1];
        out_probs[k] = (re * re + im * im) / pin * prodnfact(s, m);
        k++;
    } while (fock_next(s, m));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Counter-based random numbers (SplitMix64 mixing, Steele et al. 2014)      */
/* ------------------------------------------------------------------------ */

static inline uint64_t smix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* 64 random bits for (seed, stream, counter). */
static inline uint64_t rand64(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    uint64_t key = smix(seed ^ smix(stream));
    return smix(key + ctr * 0x9E3779B97F4A7C15ull);
}

/* Uniform double in [0,1) for (seed, stream, counter); identical definition
 * on the device side, so a sampler driven by the same seed draws the same
 * uniforms. */
double oracle_uniform(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    return (double)(rand64(seed, stream, ctr) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------------ */
/* Gurvits' randomised permanent estimate                                    */
/* ------------------------------------------------------------------------ */

/* Restates xq.estimate_permanent_cx(M, iterations) as called from
 * NaiveApproxBackend (perceval/backends/_naive_approx.py:52-62): the mean of
 * X(x) = (prod_j x_j) prod_i sum_j a_ij x_j over `iters` sign vectors, sample
 * s taking x_j = -1 where bit j of rand64(seed, stream, s) is set, and the 99%
 * normal error bound 2.5758 sqrt((mean |X|^2 - |mean X|^2) / iters).
 * out = {Re mean, Im mean, error}.  Samples are summed in order s = 0, 1, ... */
int oracle_gurvits(const double *a, int n, int64_t iters, uint64_t seed, uint64_t stream, double *out)
{
    if (n < 0 || n > 64 || iters < 1) return -1;
    double sr = 0.0, si = 0.0, sq = 0.0;
    double yr[64], yi[64];
    for (int64_t s = 0; s < iters; s++) {
        uint64_t bits = rand64(seed, stream, (uint64_t)s);
        int neg = 0;
        for (int i = 0; i < n; i++) { yr[i] = 0.0; yi[i] = 0.0; }
        for (int j = 0; j < n; j++) {
            double x = ((bits >> j) & 1) ? -1.0 : 1.0;
            neg ^= (int)((bits >> j) & 1);
            for (int i = 0; i < n; i++) {
                yr[i] += x * a[2 * (i * n + j)];
                yi[i] += x * a[2 * (i * n + j) + 1];
            }
        }
        double pr = 1.0, pi = 0.0;
        for (int i = 0; i < n; i++) {
            double tr = pr * yr[i] - pi * yi[i];
            double ti = pr * yi[i] + pi * yr[i];
            pr = tr;
            pi = ti;
        }
        if (neg) { pr = -pr; pi = -pi; }
        sr += pr;
        si += pi;
        sq += pr * pr + pi * pi;
    }
    double inv = 1.0 / (double)iters;
    double mr = sr * inv, mi = si * inv;
    double var = sq * inv - (mr * mr + mi * mi);
    out[0] = mr;
    out[1] = mi;
    out[2] = 2.5758293035489004 * sqrt((var > 0.0 ? var : 0.0) * inv);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Clifford & Clifford 2017, algorithm A                                     */
/* ------------------------------------------------------------------------ */

/* Sub-permanents of the (k-1) x k matrix B with one column removed:
 * sub[l] = perm(B without column l), via one Glynn pass over the rows of B
 * with prefix/suffix products of the shared column sums. */
static void cc_subperms(const double *B, int rows, int k, double *sub)
{
    for (int l = 0; l < 2 * k; l++) sub[l] = 0.0;
    if (rows == 0) { sub[0] = 1.0; return; } /* k == 1: perm of the empty matrix */
    double cr[OR_MAXN], ci[OR_MAXN], pr[OR_MAXN + 1], pi[OR_MAXN + 1];
    for (int l = 0; l < k; l++) {
        double sr = 0, si = 0;
        for (int i = 0; i < rows; i++) { sr += B[2 * (i * k + l)]; si += B[2 * (i * k + l) + 1]; }
        cr[l] = sr; ci[l] = si;
    }
    uint64_t steps = 1ull << (rows - 1), g = 0;
    for (uint64_t t = 0; t < steps; t++) {
        if (t) {
            int b = __builtin_ctzll(t);
            g ^= 1ull << b;
            double s = ((g >> b) & 1) ? -2.0 : 2.0;
            const double *row = B + 2 * (b + 1) * k;
            for (int l = 0; l < k; l++) { cr[l] += s * row[2 * l]; ci[l] += s * row[2 * l + 1]; }
        }
        /* prefix products pr[l] = prod_{l'<l} c_l' */
        pr[0] = 1.0; pi[0] = 0.0;
        for (int l = 0; l + 1 < k; l++) {
            pr[l + 1] = pr[l] * cr[l] - pi[l] * ci[l];
            pi[l + 1] = pr[l] * ci[l] + pi[l] * cr[l];
        }
        double sgn = (t & 1) ? -1.0 : 1.0;
        double qr = sgn, qi = 0.0; /* running suffix product times the term sign */
        for (int l = k - 1; l >= 0; l--) {
            sub[2 * l] += pr[l] * qr - pi[l] * qi;
            sub[2 * l + 1] += pr[l] * qi + pi[l] * qr;
            double nr = qr * cr[l] - qi * ci[l];
            double ni = qr * ci[l] + qi * cr[l];
            qr = nr; qi = ni;
        }
    }
    double scale = ldexp(1.0, -(rows - 1));
    for (int l = 0; l < 2 * k; l++) sub[l] *= scale;
}

/* Draw one sample.  Uniform counters: 0..n-2 for the column permutation
 * (Fisher-Yates, i = n-1 down to 1), then n-1+k for the k-th output photon. */
static void cc_one(const double *U, int m, const int32_t *cols, int n,
                   uint64_t seed, uint64_t sample, int32_t *out_occ)
{
    int32_t perm[OR_MAXN];
    for (int i = 0; i < n; i++) perm[i] = cols[i];
    uint64_t ctr = 0;
    for (int i = n - 1; i >= 1; i--) {
        double u = oracle_uniform(seed, sample, ctr++);
        int j = (int)(u * (i + 1));
        if (j > i) j = i;
        int32_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    ctr = (uint64_t)(n > 0 ? n - 1 : 0);
    /* A[i][l] = U[i][perm[l]] */
    int32_t r[OR_MAXN];
    double B[2 * OR_MAXN * OR_MAXN];
    double sub[2 * OR_MAXN];
    double *w = (double *)malloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) out_occ[i] = 0;
    for (int k = 1; k <= n; k++) {
        /* B = A[r_1..r_{k-1}, 0..k-1] */
        for (int i = 0; i < k - 1; i++)
            for (int l = 0; l < k; l++) {
                B[2 * (i * k + l)] = U[2 * (r[i] * m + perm[l])];
                B[2 * (i * k + l) + 1] = U[2 * (r[i] * m + perm[l]) + 1];
            }
        cc_subperms(B, k - 1, k, sub);
        double total = 0.0;
        for (int i = 0; i < m; i++) {
            double ar = 0, ai = 0;
            for (int l = 0; l < k; l++) {
                double ur = U[2 * (i * m + perm[l])], ui = U[2 * (i * m + perm[l]) + 1];
                ar += ur * sub[2 * l] - ui * sub[2 * l + 1];
                ai += ur * sub[2 * l + 1] + ui * sub[2 * l];
            }
            w[i] = ar * ar + ai * ai;
            total += w[i];
        }
        double target = oracle_uniform(seed, sample, ctr++) * total;
        int pick = m - 1;
        double cum = 0.0;
        for (int i = 0; i < m; i++) {
            cum += w[i];
            if (target < cum) { pick = i; break; }
        }
        r[k - 1] = pick;
        out_occ[pick]++;
    }
    free(w);
}

/* count samples of occupation vectors (count x m int32).  Sample s uses RNG
 * stream first_sample + s.  Threads over samples. */
typedef struct {
    const double *U; int m; const int32_t *cols; int n; uint64_t seed, first; int32_t *out;
} cc_ctx;

static void cc_task(void *vctx, int64_t lo, int64_t hi)
{
    cc_ctx *c = (cc_ctx *)vctx;
    for (int64_t s = lo; s < hi; s++)
        cc_one(c->U, c->m, c->cols, c->n, c->seed, c->first + (uint64_t)s, c->out + s * c->m);
}

int oracle_cc_samples(const double *U, int m, const int32_t *in_occ, int64_t count,
                      uint64_t seed, uint64_t first_sample, int32_t *out_occ, int threads)
{
    if (m <= 0 || m > OR_MAXM) return -1;
    int32_t cols[OR_MAXN];
    int n = 0;
    for (int i = 0; i < m; i++) {
        if (in_occ[i] < 0) return -1;
        for (int v = 0; v < in_occ[i]; v++) { if (n >= OR_MAXN) return -1; cols[n++] = i; }
    }
    cc_ctx c = {U, m, cols, n, seed, first_sample, out_occ};
    or_parallel(cc_task, &c, count, 1, threads);
    return 0;
}

int oracle_max_threads(void)
{
    return or_default_threads();
}
