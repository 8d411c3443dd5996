/*
 * oracle.c -- CPU restatement of PhantomFHE's RNS polynomial hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke entry and the
 * benchmark's CPU-baseline leg as the checker. The product never uses it.
 *
 * Each function restates the reference algorithm and cites the reference
 * file:line it follows (paths relative to the reference repository root).
 * Arithmetic is exact (unsigned __int128); every output is a canonical
 * residue in [0, q), as the reference's final csub steps make it.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;

/* ------------------------------------------------------------------ */
/* threading: limb-parallel loops                                      */
/* ------------------------------------------------------------------ */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

typedef void (*orc_task_fn)(void *arg, size_t i);
typedef struct {
    orc_task_fn fn;
    void *arg;
    size_t n;
    size_t next;
    pthread_mutex_t mu;
} orc_pool;

static void *orc_worker(void *p) {
    orc_pool *pool = (orc_pool *)p;
    for (;;) {
        pthread_mutex_lock(&pool->mu);
        size_t i = pool->next++;
        pthread_mutex_unlock(&pool->mu);
        if (i >= pool->n) break;
        pool->fn(pool->arg, i);
    }
    return NULL;
}

static void parallel_for(size_t n, orc_task_fn fn, void *arg) {
    int nt = g_threads;
    if (nt <= 1 || n <= 1) {
        for (size_t i = 0; i < n; i++) fn(arg, i);
        return;
    }
    if ((size_t)nt > n) nt = (int)n;
    orc_pool pool = {fn, arg, n, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    if (nt > 256) nt = 256;
    for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, orc_worker, &pool);
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------ */
/* scalar modular arithmetic                                           */
/* ------------------------------------------------------------------ */
static inline u64 mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }
static inline u64 addmod(u64 a, u64 b, u64 q) { u64 r = a + b; return r >= q ? r - q : r; }
static inline u64 submod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

static u64 powmod(u64 a, u64 e, u64 q) {
    u64 r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

/* inverse modulo a prime (every modulus on the path is prime) */
static u64 invmod(u64 a, u64 q) { return powmod(a % q, q - 2, q); }

/* compute_shoup: floor(w * 2^64 / q)  (include/host/uintarithsmallmod.h:119-124) */
u64 orc_shoup(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }

/* Harvey/Shoup lazy product in [0, 2q)  (include/uintmodmath.cuh:223-231) */
static inline u64 shoup_lazy(u64 x, u64 w, u64 ws, u64 q) {
    u64 hi = (u64)(((u128)x * ws) >> 64);
    return x * w - hi * q;
}
/* csub_q  (include/uintmodmath.cuh:18-21) */
static inline u64 csub(u64 x, u64 q) { return x >= q ? x - q : x; }

/* ------------------------------------------------------------------ */
/* primes and roots                                                    */
/* ------------------------------------------------------------------ */

/* Deterministic Miller-Rabin for 64-bit inputs. The reference runs 40 random
 * bases (src/host/numth.cu:129-205); for n < 3.3e24 the fixed base set below
 * decides primality exactly, so both agree on every 64-bit value. */
int orc_is_prime(u64 n) {
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n == small[i]) return 1;
        if (n % small[i] == 0) return 0;
    }
    u64 d = n - 1;
    int r = 0;
    while ((d & 1) == 0) { d >>= 1; r++; }
    for (int i = 0; i < 12; i++) {
        u64 x = powmod(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int k = 1; k < r; k++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* get_primes: descending primes p = 1 (mod 2N) starting from 2^bits - 2N + 1
 * and stepping by 2N while p > 2^(bits-1)   (src/host/numth.cu:207-233) */
int orc_get_primes(u64 N, int bits, size_t count, u64 *out) {
    u64 factor = 2 * N;
    u64 value = ((u64)1 << bits) - factor + 1;
    u64 lower = (u64)1 << (bits - 1);
    size_t got = 0;
    while (got < count && value > lower) {
        if (orc_is_prime(value)) out[got++] = value;
        value -= factor;
    }
    return got == count ? 0 : -1;
}

/* CoeffModulus::Create: for each bit size, the primes of that size are drawn
 * from get_primes' descending list by taking .back() each time, so equal sizes
 * come out in ascending order   (src/host/modulus.cu:80-111) */
int orc_coeff_modulus_create(u64 N, const int *bits, size_t n, u64 *out) {
    for (size_t i = 0; i < n; i++) {
        size_t cnt = 0, rank = 0;
        for (size_t j = 0; j < n; j++) {
            if (bits[j] == bits[i]) {
                cnt++;
                if (j < i) rank++;
            }
        }
        u64 *tmp = (u64 *)malloc(cnt * sizeof(u64));
        if (orc_get_primes(N, bits[i], cnt, tmp) != 0) { free(tmp); return -1; }
        out[i] = tmp[cnt - 1 - rank];
        free(tmp);
    }
    return 0;
}

/* try_minimal_primitive_root: smallest primitive degree-th root of unity mod q.
 * The reference starts from a random primitive root and scans all its odd
 * powers keeping the minimum (src/host/numth.cu:309-331); the set of odd
 * powers of any primitive root is the set of all primitive roots, so the
 * minimum does not depend on the start. Here the start is found by a
 * deterministic search instead of random_device. */
u64 orc_minimal_root(u64 q, u64 degree) {
    if ((q - 1) % degree != 0) return 0;
    u64 quot = (q - 1) / degree;
    u64 g = 0;
    for (u64 x = 2; x < q; x++) {
        u64 c = powmod(x, quot, q);
        if (powmod(c, degree >> 1, q) == q - 1) { g = c; break; }
    }
    if (!g) return 0;
    u64 g2 = mulmod(g, g, q), cur = g, best = g;
    for (u64 i = 0; i < degree; i++) {
        if (cur < best) best = cur;
        cur = mulmod(cur, g2, q);
    }
    return best;
}

static inline u64 brv(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* NTT::NTT tables (src/host/ntt.cu:11-56): root_powers[brv(i)] = psi^i,
 * inv_root_powers[brv(i)] = psi^-i, entry [1] of the inverse table
 * pre-multiplied by n^-1; Shoup companions for all. */
void orc_ntt_tables(u64 q, int logN, u64 *rp, u64 *rps, u64 *irp, u64 *irps, u64 *ninv) {
    u64 N = (u64)1 << logN;
    u64 psi = orc_minimal_root(q, 2 * N);
    u64 ipsi = invmod(psi, q);
    u64 p = psi;
    for (u64 i = 1; i < N; i++) {
        rp[brv(i, logN)] = p;
        p = mulmod(p, psi, q);
    }
    rp[0] = 1;
    p = ipsi;
    for (u64 i = 1; i < N; i++) {
        irp[brv(i, logN)] = p;
        p = mulmod(p, ipsi, q);
    }
    irp[0] = 1;
    u64 ni = invmod(N % q, q);
    irp[1] = mulmod(irp[1], ni, q);
    for (u64 i = 0; i < N; i++) {
        rps[i] = orc_shoup(rp[i], q);
        irps[i] = orc_shoup(irp[i], q);
    }
    *ninv = ni;
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
struct orc_ctx {
    int logN;
    size_t N, size_QP, size_P, size_Q;
    u64 *q;                       /* [QP] */
    u64 *rp, *rps, *irp, *irps;   /* [QP][N] */
    u64 *ninv;                    /* [QP] */
};

orc_ctx *orc_ctx_create(int logN, const u64 *moduli, size_t size_QP, size_t size_P) {
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->logN = logN;
    c->N = (size_t)1 << logN;
    c->size_QP = size_QP;
    c->size_P = size_P;
    c->size_Q = size_QP - size_P;
    c->q = (u64 *)malloc(size_QP * sizeof(u64));
    memcpy(c->q, moduli, size_QP * sizeof(u64));
    size_t tot = size_QP * c->N;
    c->rp = (u64 *)malloc(tot * 8);
    c->rps = (u64 *)malloc(tot * 8);
    c->irp = (u64 *)malloc(tot * 8);
    c->irps = (u64 *)malloc(tot * 8);
    c->ninv = (u64 *)malloc(size_QP * 8);
    for (size_t i = 0; i < size_QP; i++)
        orc_ntt_tables(c->q[i], logN, c->rp + i * c->N, c->rps + i * c->N, c->irp + i * c->N,
                       c->irps + i * c->N, c->ninv + i);
    return c;
}

void orc_ctx_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->q); free(c->rp); free(c->rps); free(c->irp); free(c->irps); free(c->ninv);
    free(c);
}

const u64 *orc_ctx_root_powers(const orc_ctx *c, size_t i) { return c->rp + i * c->N; }
const u64 *orc_ctx_inv_root_powers(const orc_ctx *c, size_t i) { return c->irp + i * c->N; }

/* ------------------------------------------------------------------ */
/* NTT / INTT of one limb                                              */
/* ------------------------------------------------------------------ */

/* Forward negacyclic NTT, natural order in, bit-reversed order out: CT
 * butterflies with lazy [0,4q) ranges (include/butterfly.cuh:10-22), stage
 * order of the radix-2 kernel (src/ntt/ntt_1d.cu:17-76), final csub 2q, q. */
static void ntt_fwd_limb(u64 *a, size_t N, u64 q, const u64 *rp, const u64 *rps) {
    u64 q2 = 2 * q;
    for (size_t m = 1, t = N >> 1; m < N; m <<= 1, t >>= 1) {
        for (size_t i = 0; i < m; i++) {
            u64 w = rp[m + i], ws = rps[m + i];
            u64 *x = a + 2 * i * t, *y = x + t;
            for (size_t j = 0; j < t; j++) {
                u64 tw_y = shoup_lazy(y[j], w, ws, q);
                u64 xx = csub(x[j], q2);
                y[j] = xx + q2 - tw_y;
                x[j] = xx + tw_y;
            }
        }
    }
    for (size_t j = 0; j < N; j++) a[j] = csub(csub(a[j], q2), q);
}

/* Inverse: GS butterflies (include/butterfly.cuh:28-37) in the order of
 * src/ntt/ntt_1d.cu:194-256; the last stage's twiddle carries n^-1 for the
 * odd output, the even output is multiplied by n^-1 explicitly. */
static void ntt_inv_limb(u64 *a, size_t N, u64 q, const u64 *irp, const u64 *irps, u64 ninv) {
    u64 q2 = 2 * q;
    u64 ninv_s = orc_shoup(ninv, q);
    for (size_t m = N >> 1, t = 1; m >= 1; m >>= 1, t <<= 1) {
        for (size_t i = 0; i < m; i++) {
            u64 w = irp[m + i], ws = irps[m + i];
            u64 *x = a + 2 * i * t, *y = x + t;
            for (size_t j = 0; j < t; j++) {
                u64 tt = x[j] + q2 - y[j];
                u64 s = csub(x[j] + y[j], q2);
                x[j] = s;
                y[j] = shoup_lazy(tt, w, ws, q);
            }
        }
        if (m == 1) break;
    }
    size_t h = N >> 1;
    for (size_t j = 0; j < h; j++) {
        u64 x = csub(a[j], q);
        u64 y = csub(a[j + h], q);
        a[j] = csub(shoup_lazy(x, ninv, ninv_s, q), q);
        a[j + h] = y;
    }
}

typedef struct {
    const orc_ctx *c;
    u64 *data;
    size_t start;
    int inverse;
} ntt_job;

static void ntt_task(void *arg, size_t i) {
    ntt_job *j = (ntt_job *)arg;
    const orc_ctx *c = j->c;
    size_t p = j->start + i;
    u64 *a = j->data + i * c->N;
    if (j->inverse)
        ntt_inv_limb(a, c->N, c->q[p], c->irp + p * c->N, c->irps + p * c->N, c->ninv[p]);
    else
        ntt_fwd_limb(a, c->N, c->q[p], c->rp + p * c->N, c->rps + p * c->N);
}

void orc_ntt_forward(const orc_ctx *c, u64 *data, size_t L, size_t start) {
    ntt_job j = {c, data, start, 0};
    parallel_for(L, ntt_task, &j);
}

void orc_ntt_inverse(const orc_ctx *c, u64 *data, size_t L, size_t start) {
    ntt_job j = {c, data, start, 1};
    parallel_for(L, ntt_task, &j);
}

/* limb transforms where each limb names its own QP table row */
typedef struct {
    const orc_ctx *c;
    u64 *data;
    const size_t *rows;
    int inverse;
} ntt_rows_job;

static void ntt_rows_task(void *arg, size_t i) {
    ntt_rows_job *j = (ntt_rows_job *)arg;
    const orc_ctx *c = j->c;
    size_t p = j->rows[i];
    u64 *a = j->data + i * c->N;
    if (j->inverse)
        ntt_inv_limb(a, c->N, c->q[p], c->irp + p * c->N, c->irps + p * c->N, c->ninv[p]);
    else
        ntt_fwd_limb(a, c->N, c->q[p], c->rp + p * c->N, c->rps + p * c->N);
}

static void ntt_rows(const orc_ctx *c, u64 *data, const size_t *rows, size_t L, int inverse) {
    ntt_rows_job j = {c, data, rows, inverse};
    parallel_for(L, ntt_rows_task, &j);
}

/* as ntt_rows over limbs that lie anywhere: limb i at ptrs[i] */
typedef struct {
    const orc_ctx *c;
    u64 **ptrs;
    const size_t *rows;
    int inverse;
} ntt_ptrs_job;

static void ntt_ptrs_task(void *arg, size_t i) {
    ntt_ptrs_job *j = (ntt_ptrs_job *)arg;
    const orc_ctx *c = j->c;
    size_t p = j->rows[i];
    if (j->inverse)
        ntt_inv_limb(j->ptrs[i], c->N, c->q[p], c->irp + p * c->N, c->irps + p * c->N, c->ninv[p]);
    else
        ntt_fwd_limb(j->ptrs[i], c->N, c->q[p], c->rp + p * c->N, c->rps + p * c->N);
}

static void ntt_ptrs(const orc_ctx *c, u64 **ptrs, const size_t *rows, size_t L, int inverse) {
    ntt_ptrs_job j = {c, ptrs, rows, inverse};
    parallel_for(L, ntt_ptrs_task, &j);
}

/* ------------------------------------------------------------------ */
/* RNS base conversion                                                 */
/* ------------------------------------------------------------------ */

/* punctured product prod_{k != skip} base[k] reduced mod m; equals the
 * reference's multiprecision punctured product reduced by modulo_uint
 * (src/host/rns.cu:296-311, 447-453) */
static u64 punctured_mod(const u64 *base, size_t n, size_t skip, u64 m) {
    u64 r = 1 % m;
    for (size_t k = 0; k < n; k++)
        if (k != skip) r = mulmod(r, base[k] % m, m);
    return r;
}

static u64 prod_mod(const u64 *base, size_t n, u64 m) {
    u64 r = 1 % m;
    for (size_t k = 0; k < n; k++) r = mulmod(r, base[k] % m, m);
    return r;
}

typedef struct {
    size_t N, ni, no;
    const u64 *ibase, *obase;
    const u64 *y;         /* [ni][N], already multiplied by qHatInv */
    const u64 *mat;       /* [no][ni] qiHat mod pj */
    u64 *out;             /* [no][N] */
    const size_t *out_rows; /* output limb index for each obase prime (NULL = identity) */
} bconv_job;

/* FastBConv phase 2: out_p = (sum_i y_i * [qHat_i]_p) mod p with a 128-bit
 * accumulator and one final reduction (src/rns_bconv.cu:143-168, 455-485;
 * include/rns_bconv.cuh:127-143). No alpha correction. */
static void bconv_task(void *arg, size_t o) {
    bconv_job *j = (bconv_job *)arg;
    u64 p = j->obase[o];
    const u64 *row = j->mat + o * j->ni;
    u64 *dst = j->out + (j->out_rows ? j->out_rows[o] : o) * j->N;
    for (size_t k = 0; k < j->N; k++) {
        u128 acc = 0;
        for (size_t i = 0; i < j->ni; i++) acc += (u128)j->y[i * j->N + k] * row[i];
        dst[k] = (u64)(acc % p);
    }
}

/* FastBConv (phase 1 + phase 2): y_i = [x_i * qHatInv_i]_{q_i}
 * (src/rns_bconv.cu:40-60), then phase 2 as above; DBaseConverter::bConv_BEHZ
 * src/rns_bconv.cu:212-229 */
typedef struct {
    size_t N, ni;
    const u64 *ibase, *in;
    u64 *y;
} bconv_pre_job;

/* FastBConv phase 1 for input i: y_i = [x_i * qHatInv_i]_{q_i} */
static void bconv_pre_task(void *arg, size_t i) {
    bconv_pre_job *j = (bconv_pre_job *)arg;
    u64 q = j->ibase[i];
    u64 hinv = invmod(punctured_mod(j->ibase, j->ni, i, q), q);
    for (size_t k = 0; k < j->N; k++) j->y[i * j->N + k] = mulmod(j->in[i * j->N + k], hinv, q);
}

/* several FastBConv phase-2 jobs as one pool of (job, output) tasks, so that
 * the digits of a modup or the polynomials of a moddown share the threads */
typedef struct {
    bconv_job *jobs;
    size_t njobs;
} bconv_multi;

static void bconv_multi_task(void *arg, size_t t) {
    bconv_multi *m = (bconv_multi *)arg;
    size_t j = 0;
    while (t >= m->jobs[j].no) t -= m->jobs[j++].no;
    bconv_task(&m->jobs[j], t);
}

static void bconv_multi_run(bconv_job *jobs, size_t njobs) {
    size_t n = 0;
    for (size_t j = 0; j < njobs; j++) n += jobs[j].no;
    bconv_multi m = {jobs, njobs};
    parallel_for(n, bconv_multi_task, &m);
}

int orc_fastbconv(size_t N, const u64 *ibase, size_t ni, const u64 *obase, size_t no,
                  const u64 *in, u64 *out) {
    u64 *y = (u64 *)malloc(ni * N * 8);
    u64 *mat = (u64 *)malloc(no * ni * 8);
    bconv_pre_job pj = {N, ni, ibase, in, y};
    parallel_for(ni, bconv_pre_task, &pj);
    for (size_t o = 0; o < no; o++)
        for (size_t i = 0; i < ni; i++) mat[o * ni + i] = punctured_mod(ibase, ni, i, obase[o]);
    bconv_job j = {N, ni, no, ibase, obase, y, mat, out, NULL};
    parallel_for(no, bconv_task, &j);
    free(y);
    free(mat);
    return 0;
}

/* ------------------------------------------------------------------ */
/* hybrid key switching                                                */
/* ------------------------------------------------------------------ */

/* beta = ceil(size_Ql / alpha), alpha = size_P (src/rns.cu:152) */
size_t orc_beta(const orc_ctx *c, size_t size_Ql) {
    return (size_Ql + c->size_P - 1) / c->size_P;
}

/* QlP prime at position pos and the QP table row it uses: data limbs map to
 * themselves, special limbs to size_Q + (pos - size_Ql)
 * (src/ntt/fntt_2d.cu:434-436, src/eval_key_switch.cu:21) */
static size_t qlp_row(const orc_ctx *c, size_t size_Ql, size_t pos) {
    return pos < size_Ql ? pos : c->size_Q + (pos - size_Ql);
}

/* DRNSTool::modup for CKKS (src/rns_bconv.cu:530-628) with the constants of
 * the hybrid-KS part of the DRNSTool ctor (src/rns.cu:136-190):
 *  1. t_cks = INTT(c2) * partQlHatInv  (fused scale, src/ntt/ntt_modup.cu:356-393)
 *  2. per digit j: own limbs copied in NTT form (modup_copy_partQl_kernel,
 *     src/rns_bconv.cu:522-528); complement limbs = FastBConv phase 2 of the
 *     digit, written at their QlP positions (bconv_matmul_padded, 455-485)
 *  3. NTT of the complement limbs (exclude-range NTT, src/ntt/ntt_modup.cu:610-657) */
typedef struct {
    size_t N, alpha, size_Ql, beta;
    const u64 *qlp;
    u64 *t_cks;
} scale_job;

/* limb i of digit j = i / alpha: t_cks[i] *= [(Q_j / q_i)^-1]_{q_i} (partQlHatInv) */
static void modup_scale_task(void *arg, size_t i) {
    scale_job *sj = (scale_job *)arg;
    size_t N = sj->N, j = i / sj->alpha;
    size_t st = sj->alpha * j;
    size_t sz = (j == sj->beta - 1) ? sj->size_Ql - sj->alpha * (sj->beta - 1) : sj->alpha;
    u64 q = sj->qlp[i];
    u64 hinv = invmod(punctured_mod(sj->qlp + st, sz, i - st, q), q);
    for (size_t k = 0; k < N; k++) sj->t_cks[i * N + k] = mulmod(sj->t_cks[i * N + k], hinv, q);
}

int orc_modup(const orc_ctx *c, size_t size_Ql, const u64 *c2, u64 *t_mod_up) {
    size_t N = c->N, alpha = c->size_P, size_QlP = size_Ql + c->size_P;
    size_t beta = orc_beta(c, size_Ql);
    u64 *qlp = (u64 *)malloc(size_QlP * 8);
    size_t *rows = (size_t *)malloc(size_QlP * sizeof(size_t));
    for (size_t p = 0; p < size_QlP; p++) {
        rows[p] = qlp_row(c, size_Ql, p);
        qlp[p] = c->q[rows[p]];
    }
    u64 *t_cks = (u64 *)malloc(size_Ql * N * 8);
    memcpy(t_cks, c2, size_Ql * N * 8);
    orc_ntt_inverse(c, t_cks, size_Ql, 0);
    scale_job sj = {N, alpha, size_Ql, beta, qlp, t_cks};
    parallel_for(size_Ql, modup_scale_task, &sj);
    /* per digit j: own limbs copied, complement limbs converted (all digits'
     * conversions as one task pool), then one NTT over every complement limb */
    bconv_job *jobs = (bconv_job *)malloc(beta * sizeof(bconv_job));
    u64 **obases = (u64 **)malloc(beta * sizeof(u64 *));
    size_t **orows = (size_t **)malloc(beta * sizeof(size_t *));
    u64 **mats = (u64 **)malloc(beta * sizeof(u64 *));
    size_t total = 0;
    for (size_t j = 0; j < beta; j++) {
        size_t st = alpha * j;
        size_t sz = (j == beta - 1) ? size_Ql - alpha * (beta - 1) : alpha;
        u64 *dst = t_mod_up + j * size_QlP * N;
        memcpy(dst + st * N, c2 + st * N, sz * N * 8);
        size_t no = size_QlP - sz;
        obases[j] = (u64 *)malloc(no * 8);
        orows[j] = (size_t *)malloc(no * sizeof(size_t));
        for (size_t o = 0; o < no; o++) {
            size_t pos = o < st ? o : o + sz;
            obases[j][o] = qlp[pos];
            orows[j][o] = pos;
        }
        mats[j] = (u64 *)malloc(no * sz * 8);
        for (size_t o = 0; o < no; o++)
            for (size_t i = 0; i < sz; i++) mats[j][o * sz + i] = punctured_mod(qlp + st, sz, i, obases[j][o]);
        jobs[j] = (bconv_job){N, sz, no, qlp + st, obases[j], t_cks + st * N, mats[j], dst, orows[j]};
        total += no;
    }
    bconv_multi_run(jobs, beta);
    /* forward NTT of every complement limb, in place, with its own table row */
    u64 **ptrs = (u64 **)malloc(total * sizeof(u64 *));
    size_t *ntt_rows_v = (size_t *)malloc(total * sizeof(size_t));
    for (size_t j = 0, t = 0; j < beta; j++)
        for (size_t o = 0; o < jobs[j].no; o++, t++) {
            ptrs[t] = jobs[j].out + orows[j][o] * N;
            ntt_rows_v[t] = rows[orows[j][o]];
        }
    ntt_ptrs(c, ptrs, ntt_rows_v, total, 0);
    for (size_t j = 0; j < beta; j++) { free(obases[j]); free(orows[j]); free(mats[j]); }
    free(jobs); free(obases); free(orows); free(mats); free(ptrs); free(ntt_rows_v);
    free(t_cks); free(qlp); free(rows);
    return 0;
}

typedef struct {
    const orc_ctx *c;
    size_t size_Ql, size_QlP, beta;
    const u64 *t, *key;
    u64 *cx;
} ip_job;

/* key_switch_inner_prod_c2_and_evk (src/eval_key_switch.cu:14-69):
 * cx0[i] = sum_j t[j][i]*evk_j.b[row(i)], cx1[i] = sum_j t[j][i]*evk_j.a[row(i)],
 * 128-bit accumulation (the mid-loop reduction never fires: the threshold
 * computed at src/eval_key_switch.cu:156-157 is nonzero), one final reduction. */
static void ip_task(void *arg, size_t pos) {
    ip_job *j = (ip_job *)arg;
    const orc_ctx *c = j->c;
    size_t N = c->N, QPN = c->size_QP * N;
    size_t row = qlp_row(c, j->size_Ql, pos);
    u64 q = c->q[row];
    for (size_t k = 0; k < N; k++) {
        u128 a0 = 0, a1 = 0;
        for (size_t b = 0; b < j->beta; b++) {
            u64 x = j->t[b * j->size_QlP * N + pos * N + k];
            const u64 *kb = j->key + b * 2 * QPN;
            a0 += (u128)x * kb[row * N + k];
            a1 += (u128)x * kb[QPN + row * N + k];
        }
        j->cx[pos * N + k] = (u64)(a0 % q);
        j->cx[j->size_QlP * N + pos * N + k] = (u64)(a1 % q);
    }
}

int orc_inner_product(const orc_ctx *c, size_t size_Ql, const u64 *t_mod_up, const u64 *key,
                      size_t dnum, u64 *cx) {
    size_t beta = orc_beta(c, size_Ql);
    if (beta > dnum) return -1;
    ip_job j = {c, size_Ql, size_Ql + c->size_P, beta, t_mod_up, key, cx};
    parallel_for(size_Ql + c->size_P, ip_task, &j);
    return 0;
}

/* DRNSTool::moddown_from_NTT for CKKS (src/rns_bconv.cu:776-828):
 *  1. INTT of the P limbs of cx_i (include_special_mod INTT)
 *  2. delta = FastBConv P -> Ql (bConv_BEHZ with base_P_to_Ql_conv, src/rns.cu:193-194)
 *  3. out = (cx_i[Ql] - NTT(delta)) * P^-1 mod q_i
 *     (fused phase 2, src/ntt/ntt_moddown.cu:106-216; sub_negate_const_mult
 *      include/uintmodmath.cuh:234-242) */
typedef struct {
    const orc_ctx *c;
    const u64 *cx_i, *delta;
    u64 *out;
} moddown_tail_job;

/* limb i: out = (cx_i - delta) * P^-1 mod q_i */
static void moddown_tail_task(void *arg, size_t i) {
    moddown_tail_job *tj = (moddown_tail_job *)arg;
    const orc_ctx *c = tj->c;
    size_t N = c->N;
    u64 q = c->q[i];
    u64 pinv = invmod(prod_mod(c->q + c->size_Q, c->size_P, q), q);
    for (size_t k = 0; k < N; k++) {
        u64 d = submod(tj->cx_i[i * N + k], tj->delta[i * N + k], q);
        tj->out[i * N + k] = mulmod(d, pinv, q);
    }
}

typedef struct {
    moddown_tail_job *jobs;
    size_t size_Ql;
} moddown_tail_pair;

static void moddown_tail_pair_task(void *arg, size_t t) {
    moddown_tail_pair *tp = (moddown_tail_pair *)arg;
    moddown_tail_task(&tp->jobs[t / tp->size_Ql], t % tp->size_Ql);
}

int orc_moddown(const orc_ctx *c, size_t size_Ql, const u64 *cx_i, u64 *out) {
    size_t N = c->N, P = c->size_P;
    u64 *xp = (u64 *)malloc(P * N * 8);
    memcpy(xp, cx_i + size_Ql * N, P * N * 8);
    orc_ntt_inverse(c, xp, P, c->size_Q);
    u64 *delta = (u64 *)malloc(size_Ql * N * 8);
    orc_fastbconv(N, c->q + c->size_Q, P, c->q, size_Ql, xp, delta);
    orc_ntt_forward(c, delta, size_Ql, 0);
    moddown_tail_job tj = {c, cx_i, delta, out};
    parallel_for(size_Ql, moddown_tail_task, &tj);
    free(xp);
    free(delta);
    return 0;
}

/* orc_moddown of both halves cx[0], cx[1] (each [QlP][N]) into out2[0], out2[1]
 * (each [Ql][N]), their limbs pooled in every step so that they share the
 * threads: INTT of the 2P special limbs, FastBConv P -> Ql of both, NTT of the
 * 2 Ql delta limbs, the tail of every limb. Same values as two orc_moddown calls. */
static void moddown_pair(const orc_ctx *c, size_t size_Ql, const u64 *cx, u64 *out2) {
    size_t N = c->N, P = c->size_P, size_QlP = size_Ql + P;
    u64 *xp = (u64 *)malloc(2 * P * N * 8);
    size_t *prow = (size_t *)malloc(2 * P * sizeof(size_t));
    for (size_t h = 0; h < 2; h++) {
        memcpy(xp + h * P * N, cx + h * size_QlP * N + size_Ql * N, P * N * 8);
        for (size_t j = 0; j < P; j++) prow[h * P + j] = c->size_Q + j;
    }
    ntt_rows(c, xp, prow, 2 * P, 1);
    /* FastBConv phase 1 (both halves share the scales) and phase 2 */
    const u64 *ibase = c->q + c->size_Q;
    u64 *y = (u64 *)malloc(2 * P * N * 8);
    bconv_pre_job pj = {N, P, ibase, xp, y};
    bconv_pre_job pj2 = {N, P, ibase, xp + P * N, y + P * N};
    parallel_for(P, bconv_pre_task, &pj);
    parallel_for(P, bconv_pre_task, &pj2);
    u64 *mat = (u64 *)malloc(size_Ql * P * 8);
    for (size_t o = 0; o < size_Ql; o++)
        for (size_t i = 0; i < P; i++) mat[o * P + i] = punctured_mod(ibase, P, i, c->q[o]);
    u64 *delta = (u64 *)malloc(2 * size_Ql * N * 8);
    bconv_job jobs[2] = {{N, P, size_Ql, ibase, c->q, y, mat, delta, NULL},
                         {N, P, size_Ql, ibase, c->q, y + P * N, mat, delta + size_Ql * N, NULL}};
    bconv_multi_run(jobs, 2);
    size_t *drow = (size_t *)malloc(2 * size_Ql * sizeof(size_t));
    for (size_t h = 0; h < 2; h++)
        for (size_t i = 0; i < size_Ql; i++) drow[h * size_Ql + i] = i;
    ntt_rows(c, delta, drow, 2 * size_Ql, 0);
    moddown_tail_job tj[2] = {{c, cx, delta, out2},
                              {c, cx + size_QlP * N, delta + size_Ql * N, out2 + size_Ql * N}};
    moddown_tail_pair tp = {tj, size_Ql};
    parallel_for(2 * size_Ql, moddown_tail_pair_task, &tp);
    free(xp); free(prow); free(y); free(mat); free(delta); free(drow);
}

/* keyswitch_inplace core (src/eval_key_switch.cu:95-182): modup, inner
 * product, moddown of both halves. out2 = [2][Ql][N], the values the
 * reference adds into ct[0], ct[1] (add_to_ct_kernel). */
int orc_keyswitch(const orc_ctx *c, size_t size_Ql, const u64 *c2, const u64 *key, size_t dnum,
                  u64 *out2) {
    size_t N = c->N, size_QlP = size_Ql + c->size_P;
    size_t beta = orc_beta(c, size_Ql);
    if (beta > dnum) return -1;
    u64 *t = (u64 *)malloc(beta * size_QlP * N * 8);
    u64 *cx = (u64 *)malloc(2 * size_QlP * N * 8);
    orc_modup(c, size_Ql, c2, t);
    orc_inner_product(c, size_Ql, t, key, dnum, cx);
    moddown_pair(c, size_Ql, cx, out2);
    free(t);
    free(cx);
    return 0;
}

/* ------------------------------------------------------------------ */
/* evaluator                                                           */
/* ------------------------------------------------------------------ */

/* tensor_prod_2x2_rns_poly (src/polymath.cu:463-498): d0 = c0 c0',
 * d2 = c1 c1', d1 = (c0+c1)(c0'+c1') - d0 - d2 (= c0 c1' + c1 c0') */
typedef struct {
    const orc_ctx *c;
    size_t LN;
    const u64 *ct1, *ct2;
    u64 *out3;
} tensor_job;

static void tensor_task(void *arg, size_t i) {
    tensor_job *j = (tensor_job *)arg;
    size_t N = j->c->N, LN = j->LN;
    u64 q = j->c->q[i];
    for (size_t k = 0; k < N; k++) {
        size_t o = i * N + k;
        u64 a0 = j->ct1[o], a1 = j->ct1[LN + o], b0 = j->ct2[o], b1 = j->ct2[LN + o];
        u64 d0 = mulmod(a0, b0, q), d2 = mulmod(a1, b1, q);
        u64 d1 = mulmod(addmod(a0, a1, q), addmod(b0, b1, q), q);
        d1 = submod(submod(d1, d0, q), d2, q);
        j->out3[o] = d0;
        j->out3[LN + o] = d1;
        j->out3[2 * LN + o] = d2;
    }
}

void orc_tensor(const orc_ctx *c, size_t size_Ql, const u64 *ct1, const u64 *ct2, u64 *out3) {
    tensor_job j = {c, size_Ql * c->N, ct1, ct2, out3};
    parallel_for(size_Ql, tensor_task, &j);
}

/* multiply_and_relin_inplace for CKKS (src/evaluate.cu:1061-1104):
 * bgv_ckks_multiply (345-397) then relinearize_inplace (1342-1374) */
typedef struct {
    const orc_ctx *c;
    size_t LN;
    const u64 *d, *ks;
    u64 *out2;
} add_job;

/* limb i of both output polynomials: d + ks (add_to_ct_kernel) */
static void add_ks_task(void *arg, size_t i) {
    add_job *aj = (add_job *)arg;
    size_t N = aj->c->N, LN = aj->LN;
    u64 q = aj->c->q[i];
    for (size_t k = 0; k < N; k++) {
        size_t o = i * N + k;
        aj->out2[o] = addmod(aj->d[o], aj->ks[o], q);
        aj->out2[LN + o] = addmod(aj->d[LN + o], aj->ks[LN + o], q);
    }
}

int orc_multiply_and_relin(const orc_ctx *c, size_t size_Ql, const u64 *ct1, const u64 *ct2,
                           const u64 *key, size_t dnum, u64 *out2) {
    size_t N = c->N, LN = size_Ql * N;
    u64 *d = (u64 *)malloc(3 * LN * 8);
    u64 *ks = (u64 *)malloc(2 * LN * 8);
    orc_tensor(c, size_Ql, ct1, ct2, d);
    int rc = orc_keyswitch(c, size_Ql, d + 2 * LN, key, dnum, ks);
    if (rc == 0) {
        add_job aj = {c, LN, d, ks, out2};
        parallel_for(size_Ql, add_ks_task, &aj);
    }
    free(d);
    free(ks);
    return rc;
}

/* rescale_to_next -> DRNSTool::divide_and_round_q_last_ntt
 * (src/rns.cu:1128-1184): INTT of the last limb, reduce it mod every other
 * q_j, NTT, then (c_j - NTT([c_last]_qj)) * q_last^-1 -- a floor division
 * (no +q_last/2 term). */
int orc_rescale(const orc_ctx *c, size_t size_Ql, const u64 *ct, size_t ct_size, u64 *out) {
    size_t N = c->N, L1 = size_Ql - 1;
    if (size_Ql < 2) return -1;
    u64 *last = (u64 *)malloc(N * 8);
    u64 *tmp = (u64 *)malloc(L1 * N * 8);
    u64 ql = c->q[L1];
    for (size_t s = 0; s < ct_size; s++) {
        const u64 *ci = ct + s * size_Ql * N;
        u64 *co = out + s * L1 * N;
        memcpy(last, ci + L1 * N, N * 8);
        orc_ntt_inverse(c, last, 1, L1);
        for (size_t j = 0; j < L1; j++)
            for (size_t k = 0; k < N; k++) tmp[j * N + k] = last[k] % c->q[j];
        orc_ntt_forward(c, tmp, L1, 0);
        for (size_t j = 0; j < L1; j++) {
            u64 q = c->q[j];
            u64 inv = invmod(ql % q, q);
            for (size_t k = 0; k < N; k++)
                co[j * N + k] = mulmod(submod(ci[j * N + k], tmp[j * N + k], q), inv, q);
        }
    }
    free(last);
    free(tmp);
    return 0;
}

/* get_elt_from_step (include/galois.cuh:16-49): 5^step mod 2N, negative
 * steps map to N/2 - |step|, step 0 to 2N - 1 (conjugation) */
uint32_t orc_galois_elt_from_step(int step, size_t N) {
    uint64_t m = 2 * (uint64_t)N;
    if (step == 0) return (uint32_t)(m - 1);
    int neg = step < 0;
    uint32_t ps = (uint32_t)(neg ? -step : step);
    if (ps >= N / 2) return 0;
    int e = neg ? (int)(N / 2) - (int)ps : (int)ps;
    uint64_t g = 1;
    while (e--) g = (g * 5) & (m - 1);
    return (uint32_t)g;
}

/* NTT-domain permutation table (include/galois.cuh:98-113) */
void orc_galois_perm_table(int logN, uint32_t elt, uint32_t *out) {
    size_t N = (size_t)1 << logN;
    for (size_t i = N; i < 2 * N; i++) {
        uint64_t r = brv(i, logN + 1);
        uint64_t idx = ((uint64_t)elt * r) >> 1;
        idx &= (N - 1);
        out[i - N] = (uint32_t)brv(idx, logN);
    }
}

/* apply_galois_inplace for CKKS (src/evaluate.cu:1567-1630): c0 <- pi(c0),
 * t <- pi(c1), c1 <- 0, then keyswitch(t) added to (c0, c1). */
int orc_apply_galois(const orc_ctx *c, size_t size_Ql, const u64 *ct, uint32_t elt,
                     const u64 *key, size_t dnum, u64 *out2) {
    size_t N = c->N, LN = size_Ql * N;
    uint32_t *perm = (uint32_t *)malloc(N * 4);
    orc_galois_perm_table(c->logN, elt, perm);
    u64 *t = (u64 *)malloc(LN * 8);
    u64 *ks = (u64 *)malloc(2 * LN * 8);
    for (size_t i = 0; i < size_Ql; i++)
        for (size_t k = 0; k < N; k++) {
            out2[i * N + k] = ct[i * N + perm[k]];
            t[i * N + k] = ct[LN + i * N + perm[k]];
        }
    int rc = orc_keyswitch(c, size_Ql, t, key, dnum, ks);
    if (rc == 0) {
        for (size_t i = 0; i < size_Ql; i++) {
            u64 q = c->q[i];
            for (size_t k = 0; k < N; k++) {
                out2[i * N + k] = addmod(out2[i * N + k], ks[i * N + k], q);
                out2[LN + i * N + k] = ks[LN + i * N + k];
            }
        }
    }
    free(perm); free(t); free(ks);
    return rc;
}

/* generate_one_kswitch_key (src/secretkey.cu:297-341) with the digit count
 * generalised to ceil(size_Q / alpha): digit j = (b_j, a_j) at the key level,
 * b_j = -(a_j s + e_j) + P s2 on the limbs of digit j
 * (encrypt_zero_symmetric src/secretkey.cu:232-289;
 *  multiply_temp_mod_and_add_rns_poly src/polymath.cu:318-338). */
int orc_make_kswitch_key(const orc_ctx *c, const u64 *s_ntt, const u64 *s2_ntt, const u64 *a,
                         const u64 *e, size_t dnum, u64 *key) {
    size_t N = c->N, QP = c->size_QP, QPN = QP * N, alpha = c->size_P;
    u64 *en = (u64 *)malloc(QPN * 8);
    for (size_t j = 0; j < dnum; j++) {
        memcpy(en, e + j * QPN, QPN * 8);
        orc_ntt_forward(c, en, QP, 0);
        u64 *b = key + j * 2 * QPN;
        u64 *ak = b + QPN;
        const u64 *aj = a + j * QPN;
        size_t st = alpha * j, en_idx = st + alpha;
        if (en_idx > c->size_Q) en_idx = c->size_Q;
        for (size_t i = 0; i < QP; i++) {
            u64 q = c->q[i];
            u64 pmod = (i >= st && i < en_idx) ? prod_mod(c->q + c->size_Q, alpha, q) : 0;
            for (size_t k = 0; k < N; k++) {
                size_t o = i * N + k;
                u64 v = addmod(mulmod(aj[o], s_ntt[o], q), en[o], q);
                v = v ? q - v : 0;
                if (pmod) v = addmod(v, mulmod(s2_ntt[o], pmod, q), q);
                b[o] = v;
                ak[o] = aj[o];
            }
        }
    }
    free(en);
    return 0;
}

/* ------------------------------------------------------------------ */
/* BFV multiply with BEHZ base extension, BFV relinearisation          */
/* ------------------------------------------------------------------ */

/* BEHZ constants of the top data level (src/rns.cu:392-571): base B of
 * 61-bit primes (size_Q, one more when 32 + bits(t) + bits(Q) >= 61 (size_Q + 1)),
 * m_sk (the first prime of the same search), m_tilde = 2^32, Bsk = B u {m_sk}. */
struct orc_behz {
    size_t size_Q, size_B, size_Bsk;
    u64 t, m_tilde;
    u64 *bsk;                        /* [Bsk]: B then m_sk */
    u64 *rp, *rps, *irp, *irps, *ninv;   /* NTT tables of the Bsk primes */
};

static int bits_of_product(const u64 *q, size_t n) {
    /* significant bits of prod q_i, exact (multi-word) */
    u64 w[64] = {0};
    size_t len = 1;
    w[0] = 1;
    for (size_t i = 0; i < n; i++) {
        u128 carry = 0;
        for (size_t k = 0; k < len; k++) {
            u128 v = (u128)w[k] * q[i] + carry;
            w[k] = (u64)v;
            carry = v >> 64;
        }
        if (carry) w[len++] = (u64)carry;
    }
    int b = 64 * (int)(len - 1);
    u64 top = w[len - 1];
    while (top) { b++; top >>= 1; }
    return b;
}

static int bit_count(u64 v) { int b = 0; while (v) { b++; v >>= 1; } return b; }

orc_behz *orc_behz_create(const orc_ctx *c, u64 t) {
    orc_behz *b = (orc_behz *)calloc(1, sizeof(orc_behz));
    b->size_Q = c->size_Q;
    b->t = t;
    b->m_tilde = (u64)1 << 32;
    int total = bits_of_product(c->q, c->size_Q);
    b->size_B = c->size_Q;
    if (32 + bit_count(t) + total >= 61 * (int)c->size_Q + 61) b->size_B++;
    b->size_Bsk = b->size_B + 1;
    u64 *primes = (u64 *)malloc((b->size_B + 1) * 8);
    orc_get_primes(c->N, 61, b->size_B + 1, primes);
    b->bsk = (u64 *)malloc(b->size_Bsk * 8);
    memcpy(b->bsk, primes + 1, b->size_B * 8);   /* B */
    b->bsk[b->size_B] = primes[0];                /* m_sk */
    free(primes);
    size_t N = c->N, tot = b->size_Bsk * N;
    b->rp = (u64 *)malloc(tot * 8); b->rps = (u64 *)malloc(tot * 8);
    b->irp = (u64 *)malloc(tot * 8); b->irps = (u64 *)malloc(tot * 8);
    b->ninv = (u64 *)malloc(b->size_Bsk * 8);
    for (size_t i = 0; i < b->size_Bsk; i++)
        orc_ntt_tables(b->bsk[i], c->logN, b->rp + i * N, b->rps + i * N, b->irp + i * N, b->irps + i * N,
                       b->ninv + i);
    return b;
}

void orc_behz_destroy(orc_behz *b) {
    if (!b) return;
    free(b->bsk); free(b->rp); free(b->rps); free(b->irp); free(b->irps); free(b->ninv);
    free(b);
}

size_t orc_behz_bsk_size(const orc_behz *b) { return b->size_Bsk; }
void orc_behz_bsk(const orc_behz *b, u64 *out) { memcpy(out, b->bsk, b->size_Bsk * 8); }

static void bsk_ntt(const orc_behz *b, size_t N, u64 *data, int inverse) {
    for (size_t i = 0; i < b->size_Bsk; i++) {
        if (inverse)
            ntt_inv_limb(data + i * N, N, b->bsk[i], b->irp + i * N, b->irps + i * N, b->ninv[i]);
        else
            ntt_fwd_limb(data + i * N, N, b->bsk[i], b->rp + i * N, b->rps + i * N);
    }
}

/* mod of a value below 2^64 by m (m may be 2^32) */
static inline u64 mod64(u64 x, u64 m) { return x % m; }

/* BEHZ steps (1)-(3) for one polynomial (BEHZ_mul_1, src/evaluate.cu:404-438):
 * fastbconv_m_tilde (src/rns.cu:1249-1277) then sm_mrq (1290-1333); out in
 * coefficient form [Bsk][N] */
static void behz_lift(const orc_ctx *c, const orc_behz *b, const u64 *x, u64 *out) {
    size_t N = c->N, k = b->size_Q, nb = b->size_Bsk;
    u64 mt = b->m_tilde;
    u64 *y = (u64 *)malloc(k * N * 8);
    for (size_t i = 0; i < k; i++) {
        u64 q = c->q[i];
        /* m_tilde * QHatInv mod q_i (src/rns.cu:452-459) */
        u64 f = mulmod(mt % q, invmod(punctured_mod(c->q, k, i, q), q), q);
        for (size_t n = 0; n < N; n++) y[i * N + n] = mulmod(x[i * N + n], f, q);
    }
    /* -Q^-1 mod m_tilde (src/rns.cu:546-553); inverse of an odd value mod 2^32 */
    u64 qmod = prod_mod(c->q, k, mt), inv = 1;
    for (int it = 0; it < 6; it++) inv = (inv * (2 - qmod * inv)) & (mt - 1);
    u64 neg_inv_q = (mt - inv) & (mt - 1);
    for (size_t p = 0; p < nb; p++) {
        u64 pm = b->bsk[p];
        u64 qmodp = prod_mod(c->q, k, pm);
        u64 mtinv = invmod(mt % pm, pm);
        for (size_t n = 0; n < N; n++) {
            u128 acc = 0, accm = 0;
            for (size_t i = 0; i < k; i++) {
                acc += (u128)y[i * N + n] * punctured_mod(c->q, k, i, pm);
                accm += (u128)y[i * N + n] * punctured_mod(c->q, k, i, mt);
            }
            u64 yp = (u64)(acc % pm), ym = (u64)(accm % mt);
            u64 r = mod64(ym * neg_inv_q, mt);
            if (r >= mt >> 1) r += pm - mt;
            u128 v = (u128)r * qmodp + yp;
            out[p * N + n] = mulmod((u64)(v % pm), mtinv, pm);
        }
    }
    free(y);
}

/* bfv_multiply_behz (src/evaluate.cu:440-545): ct1, ct2 [2][size_Q][N]
 * coefficient form -> out3 [3][size_Q][N] */
int orc_bfv_multiply(const orc_ctx *c, const orc_behz *b, const u64 *ct1, const u64 *ct2, u64 *out3) {
    size_t N = c->N, k = b->size_Q, nb = b->size_Bsk, QN = k * N, BN = nb * N;
    u64 *aq = (u64 *)malloc(3 * QN * 8), *bq = (u64 *)malloc(2 * QN * 8);
    u64 *ab = (u64 *)malloc(3 * BN * 8), *bb = (u64 *)malloc(2 * BN * 8);
    const u64 *src[2] = {ct1, ct2};
    u64 *dq[2] = {aq, bq}, *db[2] = {ab, bb};
    for (int s = 0; s < 2; s++)
        for (int poly = 0; poly < 2; poly++) {
            memcpy(dq[s] + poly * QN, src[s] + poly * QN, QN * 8);
            orc_ntt_forward(c, dq[s] + poly * QN, k, 0);
            behz_lift(c, b, src[s] + poly * QN, db[s] + poly * BN);
            bsk_ntt(b, N, db[s] + poly * BN, 0);
        }
    /* (4) tensor in both bases (tensor_prod_2x2_rns_poly, src/polymath.cu:463-498) */
    for (int base = 0; base < 2; base++) {
        size_t L = base ? nb : k;
        u64 *A = base ? ab : aq, *B = base ? bb : bq;
        for (size_t i = 0; i < L; i++) {
            u64 q = base ? b->bsk[i] : c->q[i];
            for (size_t n = 0; n < N; n++) {
                size_t o = i * N + n, LN = L * N;
                u64 a0 = A[o], a1 = A[LN + o], b0 = B[o], b1 = B[LN + o];
                u64 r0 = mulmod(a0, b0, q), r2 = mulmod(a1, b1, q);
                u64 r1 = mulmod(addmod(a0, a1, q), addmod(b0, b1, q), q);
                r1 = submod(submod(r1, r0, q), r2, q);
                A[o] = r0; A[LN + o] = r1; A[2 * LN + o] = r2;
            }
        }
    }
    /* (5)-(6) INTT, times t (nwt_2d_radix8_backward_inplace_scale) */
    for (int poly = 0; poly < 3; poly++) {
        orc_ntt_inverse(c, aq + poly * QN, k, 0);
        bsk_ntt(b, N, ab + poly * BN, 1);
        for (size_t i = 0; i < k; i++)
            for (size_t n = 0; n < N; n++) aq[poly * QN + i * N + n] = mulmod(aq[poly * QN + i * N + n], b->t, c->q[i]);
        for (size_t i = 0; i < nb; i++)
            for (size_t n = 0; n < N; n++)
                ab[poly * BN + i * N + n] = mulmod(ab[poly * BN + i * N + n], b->t % b->bsk[i], b->bsk[i]);
    }
    /* (7) fast_floor (src/rns.cu:1379-1407) and (8) fastbconv_sk (1455-1517,
     * multiply_and_negated_add_rns_poly src/polymath.cu:606-634) */
    size_t nB = b->size_B;
    u64 msk = b->bsk[nB];
    u64 *w = (u64 *)malloc(BN * 8);
    u64 *yq = (u64 *)malloc(QN * 8), *u = (u64 *)malloc(nB * N * 8);
    for (int poly = 0; poly < 3; poly++) {
        const u64 *zq = aq + poly * QN, *zb = ab + poly * BN;
        for (size_t i = 0; i < k; i++) {
            u64 q = c->q[i], hinv = invmod(punctured_mod(c->q, k, i, q), q);
            for (size_t n = 0; n < N; n++) yq[i * N + n] = mulmod(zq[i * N + n], hinv, q);
        }
        for (size_t p = 0; p < nb; p++) {
            u64 pm = b->bsk[p], qinv = invmod(prod_mod(c->q, k, pm), pm);
            for (size_t n = 0; n < N; n++) {
                u128 acc = 0;
                for (size_t i = 0; i < k; i++) acc += (u128)yq[i * N + n] * punctured_mod(c->q, k, i, pm);
                u64 cv = (u64)(acc % pm);
                w[p * N + n] = mulmod(addmod(zb[p * N + n], pm - cv == pm ? 0 : pm - cv, pm), qinv, pm);
            }
        }
        for (size_t j = 0; j < nB; j++) {
            u64 bj = b->bsk[j], hinv = invmod(punctured_mod(b->bsk, nB, j, bj), bj);
            for (size_t n = 0; n < N; n++) u[j * N + n] = mulmod(w[j * N + n], hinv, bj);
        }
        u64 binv_msk = invmod(prod_mod(b->bsk, nB, msk), msk);
        u64 *out = out3 + poly * QN;
        for (size_t n = 0; n < N; n++) {
            u128 acc = 0;
            for (size_t j = 0; j < nB; j++) acc += (u128)u[j * N + n] * punctured_mod(b->bsk, nB, j, msk);
            u64 bc = (u64)(acc % msk);
            u64 alpha = mulmod(addmod(bc, msk - w[nB * N + n], msk), binv_msk, msk);
            for (size_t i = 0; i < k; i++) {
                u64 q = c->q[i];
                u128 a2 = 0;
                for (size_t j = 0; j < nB; j++) a2 += (u128)u[j * N + n] * punctured_mod(b->bsk, nB, j, q);
                u64 v = (u64)(a2 % q);
                u64 bq_ = prod_mod(b->bsk, nB, q);
                u64 corr;
                if (alpha > msk >> 1)
                    corr = mulmod((msk - alpha) % q, bq_, q);          /* -alpha*B, alpha negative */
                else
                    corr = mulmod(alpha % q, bq_ ? q - bq_ : 0, q);   /* alpha * (-B mod q) */
                out[i * N + n] = addmod(v, corr, q);
            }
        }
    }
    free(aq); free(bq); free(ab); free(bb); free(w); free(yq); free(u);
    return 0;
}

/* relinearize_inplace for BFV (src/evaluate.cu:1342-1374) through
 * keyswitch_inplace (src/eval_key_switch.cu:95-182) with the BFV branches of
 * DRNSTool::modup (src/rns_bconv.cu:530-628: own limbs copied in coefficient
 * form, complement limbs FastBConv of the digit, then NTT of all QlP limbs)
 * and moddown_from_NTT (776-828: INTT of all QlP limbs, delta = FastBConv
 * P -> Ql, (cx - delta) * P^-1 in coefficient form), added to ct[0], ct[1]. */
int orc_bfv_relinearize(const orc_ctx *c, const u64 *ct3, const u64 *key, size_t dnum, u64 *out2) {
    size_t N = c->N, size_Ql = c->size_Q, alpha = c->size_P, size_QlP = size_Ql + c->size_P;
    size_t beta = orc_beta(c, size_Ql), QN = size_Ql * N;
    if (beta > dnum) return -1;
    const u64 *c2 = ct3 + 2 * QN;
    u64 *qlp = (u64 *)malloc(size_QlP * 8);
    size_t *rows = (size_t *)malloc(size_QlP * sizeof(size_t));
    for (size_t p = 0; p < size_QlP; p++) {
        rows[p] = qlp_row(c, size_Ql, p);
        qlp[p] = c->q[rows[p]];
    }
    u64 *t = (u64 *)malloc(beta * size_QlP * N * 8);
    for (size_t j = 0; j < beta; j++) {
        size_t st = alpha * j, sz = (j == beta - 1) ? size_Ql - alpha * (beta - 1) : alpha;
        u64 *tj = t + j * size_QlP * N;
        size_t no = size_QlP - sz;
        u64 *obase = (u64 *)malloc(no * 8), *conv = (u64 *)malloc(no * N * 8);
        size_t *opos = (size_t *)malloc(no * sizeof(size_t));
        size_t m = 0;
        for (size_t p = 0; p < size_QlP; p++)
            if (p < st || p >= st + sz) { obase[m] = qlp[p]; opos[m] = p; m++; }
        orc_fastbconv(N, qlp + st, sz, obase, no, c2 + st * N, conv);
        memcpy(tj + st * N, c2 + st * N, sz * N * 8);
        for (size_t o = 0; o < no; o++) memcpy(tj + opos[o] * N, conv + o * N, N * 8);
        ntt_rows(c, tj, rows, size_QlP, 0);
        free(obase); free(conv); free(opos);
    }
    u64 *cx = (u64 *)malloc(2 * size_QlP * N * 8);
    orc_inner_product(c, size_Ql, t, key, dnum, cx);
    u64 *delta = (u64 *)malloc(QN * 8);
    for (int h = 0; h < 2; h++) {
        u64 *cxi = cx + h * size_QlP * N;
        ntt_rows(c, cxi, rows, size_QlP, 1);
        orc_fastbconv(N, c->q + c->size_Q, c->size_P, c->q, size_Ql, cxi + QN, delta);
        for (size_t i = 0; i < size_Ql; i++) {
            u64 q = c->q[i], pinv = invmod(prod_mod(c->q + c->size_Q, c->size_P, q), q);
            for (size_t n = 0; n < N; n++) {
                u64 v = mulmod(submod(cxi[i * N + n], delta[i * N + n], q), pinv, q);
                out2[h * QN + i * N + n] = addmod(ct3[h * QN + i * N + n], v, q);
            }
        }
    }
    free(qlp); free(rows); free(t); free(cx); free(delta);
    return 0;
}

/* ------------------------------------------------------------------ */
/* BFV multiply with HPS (mul_tech hps)                                */
/* ------------------------------------------------------------------ */

/* little-endian multiword helpers for the HPS constants */
#define ORC_BIG_MAX 40
typedef struct { u64 w[ORC_BIG_MAX]; size_t n; } orc_big;

static void big_set(orc_big *a, u64 v) { memset(a, 0, sizeof(*a)); a->w[0] = v; a->n = 1; }
static void big_mul(orc_big *a, u64 m) {
    u128 carry = 0;
    for (size_t i = 0; i < a->n; i++) {
        u128 v = (u128)a->w[i] * m + carry;
        a->w[i] = (u64)v;
        carry = v >> 64;
    }
    if (carry) a->w[a->n++] = (u64)carry;
}
static u64 big_mod(const orc_big *a, u64 m) {
    u128 r = 0;
    for (size_t i = a->n; i-- > 0;) r = ((r << 64) | a->w[i]) % m;
    return (u64)r;
}
static void big_div(orc_big *a, u64 m) {   /* a = floor(a / m) */
    u128 r = 0;
    for (size_t i = a->n; i-- > 0;) {
        u128 cur = (r << 64) | a->w[i];
        a->w[i] = (u64)(cur / m);
        r = cur % m;
    }
    while (a->n > 1 && a->w[a->n - 1] == 0) a->n--;
}

/* HPS constants of the top data level (src/rns.cu:686-783 and the
 * BaseConverter tables, src/host/rns.cu:438-497): R = size_Q + 1 primes below
 * the smallest q (get_primes_below, src/host/numth.cu:235-262). The over-Q
 * variant (mul_tech hps_overq, src/rns.cu:792-890) takes R = size_Q primes and
 * the t/R scale-and-round constants of QR -> Q. */
struct orc_hps {
    size_t k, nr;
    u64 t;
    int overq;
    u64 *r;                        /* [nr] */
    u64 *rp, *rps, *irp, *irps, *ninv;
    double *frac;                  /* [k] tRSHatInvModsDivsFrac */
    u64 *divs_mod_r;               /* [nr][k + 1] tRSHatInvModsDivsModr */
    /* over Q */
    size_t ks;                     /* source moduli of the switch: q_0 .. q_{ks-1} */
    u64 *negrq;                    /* [ks] negPQHatInvModq of the source -> R converter (P = R) */
    u64 *qinv_r;                   /* [nr][ks] QInvModp: q_i^-1 mod r_j */
    double *frac_r;                /* [nr] tQlSlHatInvModsDivsFrac */
    u64 *divs_mod_q;               /* [k][nr + 1] tQlSlHatInvModsDivsModq */
};

static void hps_overq_consts(const orc_ctx *c, orc_hps *h) {
    size_t k = h->k, nr = h->nr, ns = k + nr, ks = h->ks;
    h->negrq = (u64 *)malloc(ks * 8);
    h->qinv_r = (u64 *)malloc(nr * ks * 8);
    h->frac_r = (double *)malloc(nr * sizeof(double));
    h->divs_mod_q = (u64 *)malloc(k * (nr + 1) * 8);
    for (size_t i = 0; i < ks; i++) {
        u64 q = c->q[i];
        u64 v = mulmod(prod_mod(h->r, nr, q), invmod(punctured_mod(c->q, ks, i, q), q), q);
        h->negrq[i] = q - v;
        for (size_t j = 0; j < nr; j++) h->qinv_r[j * ks + i] = invmod(q % h->r[j], h->r[j]);
    }
    /* theta_i = t Q [(S/s_i)^-1]_{s_i}, S = Q R (src/rns.cu:833-890) */
    u64 *s = (u64 *)malloc(ns * 8);
    memcpy(s, c->q, k * 8);
    memcpy(s + k, h->r, nr * 8);
    for (size_t i = 0; i < ns; i++) {
        u64 shinv = invmod(punctured_mod(s, ns, i, s[i]), s[i]);
        orc_big th;
        big_set(&th, h->t);
        for (size_t m = 0; m < k; m++) big_mul(&th, c->q[m]);
        big_mul(&th, shinv);
        if (i >= k) h->frac_r[i - k] = (double)big_mod(&th, s[i]) / (double)s[i];
        big_div(&th, s[i]);
        for (size_t m = 0; m < k; m++) {
            if (i >= k) h->divs_mod_q[m * (nr + 1) + (i - k)] = big_mod(&th, c->q[m]);
            else if (i == m) h->divs_mod_q[m * (nr + 1) + nr] = big_mod(&th, c->q[m]);
        }
    }
    free(s);
}

orc_hps *orc_hps_create(const orc_ctx *c, u64 t) { return orc_hps_create2(c, t, 0); }

/* HPS tool for Ql = q_0 .. q_{k-1}; R from the smallest prime of all of Q
 * (src/rns.cu:686-699, 792-801: the R of a lower level is a prefix of the top one) */
static orc_hps *hps_create(const orc_ctx *c, u64 t, int overq, size_t k, size_t ks) {
    orc_hps *h = (orc_hps *)calloc(1, sizeof(orc_hps));
    size_t nr = overq ? k : k + 1, N = c->N;
    h->k = k; h->nr = nr; h->t = t; h->overq = overq; h->ks = ks;
    u64 qmin = c->q[0];
    for (size_t i = 1; i < c->size_Q; i++) if (c->q[i] < qmin) qmin = c->q[i];
    h->r = (u64 *)malloc(nr * 8);
    int bits = bit_count(qmin);
    u64 lower = (u64)1 << (bits - 1), v = qmin - 2 * N;
    size_t m = 0;
    while (m < nr && v > lower) {
        if (orc_is_prime(v)) h->r[m++] = v;
        v -= 2 * N;
    }
    h->rp = (u64 *)malloc(nr * N * 8); h->rps = (u64 *)malloc(nr * N * 8);
    h->irp = (u64 *)malloc(nr * N * 8); h->irps = (u64 *)malloc(nr * N * 8);
    h->ninv = (u64 *)malloc(nr * 8);
    for (size_t j = 0; j < nr; j++)
        orc_ntt_tables(h->r[j], c->logN, h->rp + j * N, h->rps + j * N, h->irp + j * N, h->irps + j * N, h->ninv + j);
    /* S = Q R; theta_i = t R [(S/s_i)^-1]_{s_i} */
    size_t ns = k + nr;
    u64 *s = (u64 *)malloc(ns * 8);
    memcpy(s, c->q, k * 8);
    memcpy(s + k, h->r, nr * 8);
    h->frac = (double *)malloc(k * sizeof(double));
    h->divs_mod_r = (u64 *)malloc(nr * (k + 1) * 8);
    for (size_t i = 0; i < ns; i++) {
        u64 shinv = invmod(punctured_mod(s, ns, i, s[i]), s[i]);
        orc_big th;
        big_set(&th, t);
        for (size_t j = 0; j < nr; j++) big_mul(&th, h->r[j]);
        big_mul(&th, shinv);
        if (i < k) h->frac[i] = (double)big_mod(&th, s[i]) / (double)s[i];
        big_div(&th, s[i]);
        for (size_t j = 0; j < nr; j++) {
            if (i < k) h->divs_mod_r[j * (k + 1) + i] = big_mod(&th, h->r[j]);
            else if (i == k + j) h->divs_mod_r[j * (k + 1) + k] = big_mod(&th, h->r[j]);
        }
    }
    free(s);
    if (overq) hps_overq_consts(c, h);
    return h;
}

orc_hps *orc_hps_create2(const orc_ctx *c, u64 t, int overq) {
    return hps_create(c, t, overq, c->size_Q, c->size_Q);
}

void orc_hps_destroy(orc_hps *h) {
    if (!h) return;
    free(h->r); free(h->rp); free(h->rps); free(h->irp); free(h->irps); free(h->ninv);
    free(h->frac); free(h->divs_mod_r);
    free(h->negrq); free(h->qinv_r); free(h->frac_r); free(h->divs_mod_q);
    free(h);
}

size_t orc_hps_r_size(const orc_hps *h) { return h->nr; }
void orc_hps_r(const orc_hps *h, u64 *out) { memcpy(out, h->r, h->nr * 8); }

/* The FP64 sums of the HPS path, all in one form: acc = init, then for each
 * term in input order acc = fma((double) x_i, w_i, acc) -- the reference's
 * `acc += static_cast<double>(x_i) * w_i` (src/rns_bconv.cu:170-195,
 * src/rns.cu:1708-1714, 1757-1763) as nvcc contracts it (multiply-add fused,
 * u64 -> double rounded to nearest even). x_i = x[i * stride]. */
static double fma_chain(double init, const u64 *x, size_t stride, const double *w, size_t n) {
    double acc = init;
    for (size_t i = 0; i < n; i++) acc = fma((double)x[i * stride], w[i], acc);
    return acc;
}

void orc_fma_chain(const u64 *x, size_t nterms, size_t N, const double *w, double init, double *out) {
    for (size_t n = 0; n < N; n++) out[n] = fma_chain(init, x + n, N, w, nterms);
}

/* DBaseConverter::bConv_HPS (src/rns_bconv.cu:277-372): y_i = [x_i ihat_i^-1]_{i};
 * out_p = sum_i y_i [ihat_i]_p - v [I]_p with v = llround(sum_i y_i fl(1 / i_i))
 * (the FP64 sum from 0 by fma_chain; ibase_.qiInv() = 1.0 / (double) q_i,
 * src/host/rns.cu:319-323). sum (may be null): the FP64 sum per coefficient */
void orc_bconv_hps(size_t N, const u64 *ib, size_t ni, const u64 *ob, size_t no, const u64 *in, u64 *out,
                   double *sum) {
    u64 *y = (u64 *)malloc(ni * N * 8);
    double *inv = (double *)malloc(ni * sizeof(double));
    double *fs = (double *)malloc(N * sizeof(double));
    for (size_t i = 0; i < ni; i++) {
        u64 hinv = invmod(punctured_mod(ib, ni, i, ib[i]), ib[i]);
        inv[i] = 1.0 / (double)ib[i];
        for (size_t n = 0; n < N; n++) y[i * N + n] = mulmod(in[i * N + n], hinv, ib[i]);
    }
    for (size_t n = 0; n < N; n++) fs[n] = fma_chain(0.0, y + n, N, inv, ni);
    if (sum) memcpy(sum, fs, N * sizeof(double));
    for (size_t o = 0; o < no; o++) {
        u64 p = ob[o], imodp = prod_mod(ib, ni, p);
        for (size_t n = 0; n < N; n++) {
            u128 acc = 0;
            for (size_t i = 0; i < ni; i++) acc += (u128)y[i * N + n] * punctured_mod(ib, ni, i, p);
            long long v = llround(fs[n]);
            u64 sub = mulmod((u64)v % p, imodp, p);
            out[o * N + n] = submod((u64)(acc % p), sub, p);
        }
    }
    free(y); free(inv); free(fs);
}

static void bconv_hps(size_t N, const u64 *ib, size_t ni, const u64 *ob, size_t no, const u64 *in, u64 *out) {
    orc_bconv_hps(N, ib, ni, ob, no, in, out, NULL);
}

static void r_ntt(const orc_hps *h, size_t N, u64 *data, int inverse) {
    for (size_t j = 0; j < h->nr; j++) {
        if (inverse)
            ntt_inv_limb(data + j * N, N, h->r[j], h->irp + j * N, h->irps + j * N, h->ninv[j]);
        else
            ntt_fwd_limb(data + j * N, N, h->r[j], h->rp + j * N, h->rps + j * N);
    }
}

/* DBaseConverter::bConv_BEHZ_var1 Q -> R (src/rns_bconv.cu:231-245):
 * y_i = [x_i negPQHatInvModq_i]_{q_i}, out_j = [sum_i y_i q_i^-1]_{r_j}.
 * For X = CRT(x) in [0, Q): out = (R X + u) / Q with u = sum_i y_i Qhat_i in
 * [0, k Q), i.e. the modulus switch Q -> R of X rounded up by at most k. */
void orc_hps_modswitch(const orc_ctx *c, const orc_hps *h, const u64 *x, u64 *out) {
    size_t N = c->N, ks = h->ks, nr = h->nr;
    for (size_t n = 0; n < N; n++) {
        u64 y[64];
        for (size_t i = 0; i < ks; i++) y[i] = mulmod(x[i * N + n], h->negrq[i], c->q[i]);
        for (size_t j = 0; j < nr; j++) {
            u128 acc = 0;
            for (size_t i = 0; i < ks; i++) acc += (u128)y[i] * h->qinv_r[j * ks + i];
            out[j * N + n] = (u64)(acc % h->r[j]);
        }
    }
}

/* The scale-and-round step of one polynomial d [k + nr][N] (Q part, then R part).
 * mul_tech hps: scaleAndRound_HPS_QR_R (src/rns.cu:1700-1746): nu = 0.5 +
 * sum_i x_{q_i} frac_i (fma_chain), alpha = (u64) nu, out_j = [sum_i x_{q_i} D_ji +
 * x_{r_j} D_jk + alpha]_{r_j}, out [nr][N]. hps_overq / leveled:
 * scaleAndRound_HPS_QlRl_Ql (src/rns.cu:1749-1796): nu = 0.5 + sum_j x_{r_j}
 * frac_j, out_i = [sum_j x_{r_j} D_ij + x_{q_i} D_ik + alpha]_{q_i}, out [k][N].
 * nu (may be null): the FP64 sum per coefficient */
void orc_hps_scale_round(const orc_ctx *c, const orc_hps *h, const u64 *d, u64 *out, double *nu_out) {
    size_t N = c->N, k = h->k, nr = h->nr, QN = k * N;
    for (size_t n = 0; n < N; n++) {
        if (h->overq) {
            double nu = fma_chain(0.5, d + QN + n, N, h->frac_r, nr);
            if (nu_out) nu_out[n] = nu;
            u64 alpha = (u64)nu;
            for (size_t i = 0; i < k; i++) {
                u64 qi = c->q[i];
                const u64 *dm = h->divs_mod_q + i * (nr + 1);
                u128 acc = 0;
                for (size_t j = 0; j < nr; j++) acc += (u128)d[QN + j * N + n] * dm[j];
                acc += (u128)d[i * N + n] * dm[nr];
                out[i * N + n] = addmod((u64)(acc % qi), alpha % qi, qi);
            }
            continue;
        }
        double nu = fma_chain(0.5, d + n, N, h->frac, k);
        if (nu_out) nu_out[n] = nu;
        u64 alpha = (u64)nu;
        for (size_t j = 0; j < nr; j++) {
            u64 rj = h->r[j];
            const u64 *dm = h->divs_mod_r + j * (k + 1);
            u128 acc = 0;
            for (size_t i = 0; i < k; i++) acc += (u128)d[i * N + n] * dm[i];
            acc += (u128)d[QN + j * N + n] * dm[k];
            out[j * N + n] = addmod((u64)(acc % rj), alpha % rj, rj);
        }
    }
}

void orc_hps_frac(const orc_hps *h, double *frac_q, double *frac_r) {
    if (frac_q) memcpy(frac_q, h->frac, h->k * sizeof(double));
    if (frac_r && h->overq) memcpy(frac_r, h->frac_r, h->nr * sizeof(double));
}

int orc_bfv_multiply_hps(const orc_ctx *c, const orc_hps *h, const u64 *ct1, const u64 *ct2, u64 *out3) {
    return orc_bfv_multiply_hps2(c, h, ct1, ct2, 0, out3);
}

/* bfv_multiply_hps (src/evaluate.cu:648-801). mul_tech hps: extend both
 * operands Q -> QR (bConv_HPS), NTT, tensor, INTT, scaleAndRound_HPS_QR_R
 * (src/rns.cu:1700-1746: nu = 0.5 + sum x_i frac_i with fused multiply-adds,
 * alpha = (u64) nu), then bConv_HPS R -> Q. hps_overq: the second operand is
 * switched Q -> R (bConv_BEHZ_var1) and converted back R -> Q (bConv_HPS); a
 * squaring (square != 0) extends the one operand like the first (lines
 * 717-731); the result is scaleAndRound_HPS_QlRl_Ql (src/rns.cu:1749-1796:
 * nu = 0.5 + sum_j x_{r_j} frac_j, alpha = (u64) nu, out_i = [sum_j x_{r_j}
 * D_ij + x_{q_i} D_ik + alpha]_{q_i}). ct1, ct2 [2][k][N] -> out3 [3][k][N] */
int orc_bfv_multiply_hps2(const orc_ctx *c, const orc_hps *h, const u64 *ct1, const u64 *ct2, int square,
                          u64 *out3) {
    size_t N = c->N, k = h->k, nr = h->nr, QN = k * N, RN = nr * N, SN = QN + RN;
    u64 *a = (u64 *)malloc(3 * SN * 8), *b = (u64 *)malloc(2 * SN * 8);
    const u64 *src[2] = {ct1, square ? ct1 : ct2};
    u64 *dst[2] = {a, b};
    for (int s = 0; s < 2; s++)
        for (int p = 0; p < 2; p++) {
            u64 *d = dst[s] + p * SN;
            if (h->overq && s == 1 && !square) {
                orc_hps_modswitch(c, h, src[s] + p * QN, d + QN);
                bconv_hps(N, h->r, nr, c->q, k, d + QN, d);
            } else {
                memcpy(d, src[s] + p * QN, QN * 8);
                bconv_hps(N, c->q, k, h->r, nr, d, d + QN);
            }
            orc_ntt_forward(c, d, k, 0);
            r_ntt(h, N, d + QN, 0);
        }
    for (size_t l = 0; l < k + nr; l++) {
        u64 q = l < k ? c->q[l] : h->r[l - k];
        for (size_t n = 0; n < N; n++) {
            size_t o = l * N + n;
            u64 a0 = a[o], a1 = a[SN + o], b0 = b[o], b1 = b[SN + o];
            u64 r0 = mulmod(a0, b0, q), r2 = mulmod(a1, b1, q);
            u64 r1 = mulmod(addmod(a0, a1, q), addmod(b0, b1, q), q);
            r1 = submod(submod(r1, r0, q), r2, q);
            a[o] = r0; a[SN + o] = r1; a[2 * SN + o] = r2;
        }
    }
    u64 *tr = (u64 *)malloc(RN * 8);
    for (int p = 0; p < 3; p++) {
        u64 *d = a + p * SN;
        orc_ntt_inverse(c, d, k, 0);
        r_ntt(h, N, d + QN, 1);
        if (h->overq) {
            orc_hps_scale_round(c, h, d, out3 + p * QN, NULL);
            continue;
        }
        orc_hps_scale_round(c, h, d, tr, NULL);
        bconv_hps(N, h->r, nr, c->q, k, tr, out3 + p * QN);
    }
    free(a); free(b); free(tr);
    return 0;
}

/* ---- hps_overq_leveled (src/evaluate.cu:549-646, 648-801, 819-1027; src/rns.cu:920-1001, 1799-1846) ---- */

orc_hps *orc_hps_create_level(const orc_ctx *c, u64 t, size_t d) {
    size_t k0 = c->size_Q;
    if (d >= k0) return NULL;
    return hps_create(c, t, 1, k0 - d, d ? k0 : k0);
}

/* FindLevelsToDrop (src/evaluate.cu:549-646): sigma = 3.2f, assurance 36,
 * ternary key bound 1 (include/host/hestdparms.h:152-153); natural logs as
 * the reference evaluates them */
size_t orc_hps_levels_to_drop(const orc_ctx *c, u64 t, size_t depth, int key_switch, int asymmetric) {
    double sigma = (double)3.2f, alpha = 36.0, p = (double)t, Bkey = 1.0;
    uint32_t kP = (uint32_t)c->size_P, numPartQ = (uint32_t)orc_beta(c, c->size_Q), n = (uint32_t)c->N;
    u64 qmax = 0;
    for (size_t i = 0; i < c->size_Q; i++) if (c->q[i] > qmax) qmax = c->q[i];
    double dcrtBits = (double)bit_count(qmax);
    double Berr = sigma * sqrt(alpha);
    double delta = 2. * sqrt(n);
    double Vnorm = asymmetric ? (1. + delta * Bkey) / 2. : Berr * (1. + 2. * delta * Bkey);
    double noiseKS = kP * (numPartQ * delta * Berr + delta * Bkey + 1.0) / 2;
    double C1 = delta * delta * p * Bkey;
    double C2 = delta * delta * Bkey * Bkey / 2.0 + noiseKS;
    double logq = depth > 0 ? log(4 * p) + (double)(depth - 1) * log(C1) + log(C1 * Vnorm + (double)depth * C2)
                            : log(p * (4 * Vnorm));
    double loge = logq / log(2) - 2 - log2(p);
    double logExtra = key_switch ? log2(noiseKS) : log2(delta);
    int32_t levels = (int32_t)floor((loge - (double)(2 * depth) - 16 - logExtra) / dcrtBits);
    if (levels < 0) levels = 0;
    else if (levels > (int32_t)c->size_Q - 1) levels = (int32_t)c->size_Q - 1;
    return (size_t)levels;
}

/* scaleAndRound_HPS_Q_Ql (src/rns.cu:1799-1806, tables 920-1001): theta_i = Ql
 * [(Q/q_i)^-1]_{q_i}; nu = 0.5 + sum_j x_{Ql+j} frac_j (fused multiply-adds in
 * order), alpha = (u64) nu; out_i = [sum_j x_{Ql+j} D_ij + x_i D_id + alpha]_{q_i},
 * i.e. round(Ql X / Q) for X = CRT_Q(x). x [k0][N] -> out [Ql][N], d >= 1 */
void orc_hps_scale_q_ql(const orc_ctx *c, size_t d, const u64 *x, u64 *out) {
    orc_hps_scale_q_ql_nu(c, d, x, out, NULL);
}

void orc_hps_scale_q_ql_nu(const orc_ctx *c, size_t d, const u64 *x, u64 *out, double *nu_out) {
    size_t N = c->N, k0 = c->size_Q, Ql = k0 - d;
    double *frac = (double *)malloc(d * sizeof(double));
    u64 *D = (u64 *)malloc(Ql * (d + 1) * 8);
    for (size_t i = 0; i < k0; i++) {
        orc_big th;
        big_set(&th, 1);
        for (size_t m = 0; m < Ql; m++) big_mul(&th, c->q[m]);
        big_mul(&th, invmod(punctured_mod(c->q, k0, i, c->q[i]), c->q[i]));
        if (i >= Ql) frac[i - Ql] = (double)big_mod(&th, c->q[i]) / (double)c->q[i];
        big_div(&th, c->q[i]);
        for (size_t m = 0; m < Ql; m++) {
            if (i >= Ql) D[m * (d + 1) + (i - Ql)] = big_mod(&th, c->q[m]);
            else if (i == m) D[m * (d + 1) + d] = big_mod(&th, c->q[m]);
        }
    }
    for (size_t n = 0; n < N; n++) {
        double nu = fma_chain(0.5, x + Ql * N + n, N, frac, d);
        if (nu_out) nu_out[n] = nu;
        u64 alpha = (u64)nu;
        for (size_t i = 0; i < Ql; i++) {
            u64 qi = c->q[i];
            u128 acc = (u128)x[i * N + n] * D[i * (d + 1) + d];
            for (size_t j = 0; j < d; j++) acc += (u128)x[(Ql + j) * N + n] * D[i * (d + 1) + j];
            out[i * N + n] = addmod((u64)(acc % qi), alpha % qi, qi);
        }
    }
    free(frac); free(D);
}

/* ExpandCRTBasis_Ql_Q (src/rns.cu:1808-1834): out_i = [x_i (Q/Ql)]_{q_i} for
 * i < Ql, 0 for the dropped limbs. x [Ql][N] -> out [k0][N] (may alias) */
void orc_hps_expand(const orc_ctx *c, size_t d, const u64 *x, u64 *out) {
    size_t N = c->N, k0 = c->size_Q, Ql = k0 - d;
    for (size_t i = 0; i < Ql; i++) {
        u64 qi = c->q[i], f = prod_mod(c->q + Ql, d, qi);
        for (size_t n = 0; n < N; n++) out[i * N + n] = mulmod(x[i * N + n], f, qi);
    }
    for (size_t i = Ql; i < k0; i++) memset(out + i * N, 0, N * 8);
}

/* bfv_multiply_hps / the multiply of bfv_mul_relin_hps with d dropped levels
 * (src/evaluate.cu:648-801, 819-960), h = orc_hps_create_level(c, t, d):
 * ct1 scaled Q -> Ql (scaleAndRound_HPS_Q_Ql) then extended Ql -> Rl (bConv_HPS);
 * ct2 switched Q -> Rl (bConv_BEHZ_var1 from all of Q) and converted Rl -> Ql
 * (bConv_HPS); a squaring extends ct1 only; NTT over QlRl, tensor, INTT,
 * scaleAndRound_HPS_QlRl_Ql, then the first expand_polys results expanded Ql -> Q.
 * ct1, ct2 [2][k0][N] -> out3 [3][k0][N]; limbs >= Ql of an unexpanded result are 0 */
int orc_bfv_multiply_hps_level(const orc_ctx *c, const orc_hps *h, size_t d, const u64 *ct1, const u64 *ct2,
                               int square, size_t expand_polys, u64 *out3) {
    size_t N = c->N, k0 = c->size_Q, k = h->k, nr = h->nr, QN0 = k0 * N, QN = k * N, RN = nr * N, SN = QN + RN;
    if (k + d != k0 || !h->overq) return -1;
    u64 *a = (u64 *)malloc(3 * SN * 8), *b = (u64 *)malloc(2 * SN * 8);
    const u64 *src[2] = {ct1, square ? ct1 : ct2};
    u64 *dst[2] = {a, b};
    for (int s2 = 0; s2 < 2; s2++)
        for (int p = 0; p < 2; p++) {
            u64 *dd = dst[s2] + p * SN;
            const u64 *x = src[s2] + p * QN0;
            if (s2 == 1 && !square) {
                orc_hps_modswitch(c, h, x, dd + QN);
                bconv_hps(N, h->r, nr, c->q, k, dd + QN, dd);
            } else {
                if (d) orc_hps_scale_q_ql(c, d, x, dd);
                else memcpy(dd, x, QN * 8);
                bconv_hps(N, c->q, k, h->r, nr, dd, dd + QN);
            }
            orc_ntt_forward(c, dd, k, 0);
            r_ntt(h, N, dd + QN, 0);
        }
    for (size_t l = 0; l < k + nr; l++) {
        u64 q = l < k ? c->q[l] : h->r[l - k];
        for (size_t n = 0; n < N; n++) {
            size_t o = l * N + n;
            u64 a0 = a[o], a1 = a[SN + o], b0 = b[o], b1 = b[SN + o];
            u64 r0 = mulmod(a0, b0, q), r2 = mulmod(a1, b1, q);
            u64 r1 = mulmod(addmod(a0, a1, q), addmod(b0, b1, q), q);
            r1 = submod(submod(r1, r0, q), r2, q);
            a[o] = r0; a[SN + o] = r1; a[2 * SN + o] = r2;
        }
    }
    memset(out3, 0, 3 * QN0 * 8);
    for (int p = 0; p < 3; p++) {
        u64 *dd = a + p * SN, *o = out3 + p * QN0;
        orc_ntt_inverse(c, dd, k, 0);
        r_ntt(h, N, dd + QN, 1);
        orc_hps_scale_round(c, h, dd, o, NULL);
        if (d && (size_t)p < expand_polys) orc_hps_expand(c, d, o, o);
    }
    free(a); free(b);
    return 0;
}

/* hoisting_inplace for CKKS (src/evaluate.cu:1670-1865): the sum of the
 * rotations of ct by elts[0..n-1] with one modup of c1:
 *   acc_c0 = sum_i pi_i(c0),  acc_cx = sum_i IP(pi_i(modup(c1)), key_i)
 *   out = (acc_c0 + moddown(acc_cx0), moddown(acc_cx1))
 * keys [n][dnum][2][QP][N]; pi_i permutes every limb of every digit. */
int orc_hoisting(const orc_ctx *c, size_t size_Ql, const u64 *ct, const uint32_t *elts, size_t n,
                 const u64 *keys, size_t dnum, u64 *out2) {
    size_t N = c->N, LN = size_Ql * N, QlP = size_Ql + c->size_P, LPN = QlP * N;
    size_t beta = orc_beta(c, size_Ql);
    if (beta > dnum || n == 0) return -1;
    size_t key_words = dnum * 2 * c->size_QP * N;
    u64 *t = (u64 *)malloc(beta * LPN * 8), *tp = (u64 *)malloc(beta * LPN * 8);
    u64 *acc0 = (u64 *)calloc(LN, 8), *acx = (u64 *)calloc(2 * LPN, 8), *cx = (u64 *)malloc(2 * LPN * 8);
    uint32_t *perm = (uint32_t *)malloc(N * 4);
    orc_modup(c, size_Ql, ct + LN, t);
    for (size_t e = 0; e < n; e++) {
        orc_galois_perm_table(c->logN, elts[e], perm);
        for (size_t i = 0; i < size_Ql; i++) {
            u64 q = c->q[i];
            for (size_t k = 0; k < N; k++) acc0[i * N + k] = addmod(acc0[i * N + k], ct[i * N + perm[k]], q);
        }
        for (size_t r = 0; r < beta * QlP; r++)
            for (size_t k = 0; k < N; k++) tp[r * N + k] = t[r * N + perm[k]];
        orc_inner_product(c, size_Ql, tp, keys + e * key_words, dnum, cx);
        for (size_t h = 0; h < 2; h++)
            for (size_t p = 0; p < QlP; p++) {
                u64 q = c->q[qlp_row(c, size_Ql, p)];
                for (size_t k = 0; k < N; k++)
                    acx[h * LPN + p * N + k] = addmod(acx[h * LPN + p * N + k], cx[h * LPN + p * N + k], q);
            }
    }
    orc_moddown(c, size_Ql, acx, out2);
    orc_moddown(c, size_Ql, acx + LPN, out2 + LN);
    for (size_t i = 0; i < size_Ql; i++)
        for (size_t k = 0; k < N; k++) out2[i * N + k] = addmod(out2[i * N + k], acc0[i * N + k], c->q[i]);
    free(t); free(tp); free(acc0); free(acx); free(cx); free(perm);
    return 0;
}

/* hoisting_inplace for BFV (src/evaluate.cu:1670-1865, coefficient form): c0 is
 * permuted in the coefficient domain (apply_galois: X^j -> X^{j g mod 2N}, negated
 * when j g mod 2N >= N, include/galois.cuh), c1 is key-switched as in the CKKS
 * form on its NTT and brought back: the BFV modup / moddown_from_NTT of the
 * reference are the NTT-domain ones conjugated by the (exact) transform. */
int orc_hoisting_bfv(const orc_ctx *c, size_t size_Ql, const u64 *ct, const uint32_t *elts, size_t n,
                     const u64 *keys, size_t dnum, u64 *out2) {
    size_t N = c->N, LN = size_Ql * N;
    u64 *x = (u64 *)calloc(2 * LN, 8), *acc0 = (u64 *)calloc(LN, 8);
    memcpy(x + LN, ct + LN, LN * 8);
    orc_ntt_forward(c, x + LN, size_Ql, 0);
    int rc = orc_hoisting(c, size_Ql, x, elts, n, keys, dnum, out2);
    if (rc == 0) {
        orc_ntt_inverse(c, out2, size_Ql, 0);
        orc_ntt_inverse(c, out2 + LN, size_Ql, 0);
        for (size_t e = 0; e < n; e++) {
            u64 g = elts[e];
            for (size_t i = 0; i < size_Ql; i++) {
                u64 q = c->q[i];
                for (size_t j = 0; j < N; j++) {
                    u64 d = (u64)((u128)j * g % (2 * N));
                    u64 v = ct[i * N + j];
                    if (d >= N) { d -= N; v = v ? q - v : 0; }
                    acc0[i * N + d] = addmod(acc0[i * N + d], v, q);
                }
            }
        }
        for (size_t i = 0; i < LN; i++) out2[i] = addmod(out2[i], acc0[i], c->q[i / N]);
    }
    free(x); free(acc0);
    return rc;
}

/* ---- seeded uniform sampler (src/prng.cu:17-210) ----
 * Block: the Salsa20 hash (Bernstein's specification: quarterround, rowround,
 * columnround, doubleround x 10, then the input added word-wise) of the 16
 * words seed[0..7], nonce low, nonce high, seed[8..13] (little-endian words of
 * the 64-byte seed). Sampler: block b (< L N / 8) yields coefficients 8b..8b+7
 * of limb b / (N/8); a word above (2^64-1) - ((2^64-1) mod q) - 1 is replaced,
 * from that position on, by a fresh block for nonce b + r L N (r counting the
 * blocks drawn for b); accepted words are reduced mod q. */
static uint32_t salsa_rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

void orc_salsa_quarterround(uint32_t y[4]) {
    y[1] ^= salsa_rotl32(y[0] + y[3], 7);
    y[2] ^= salsa_rotl32(y[1] + y[0], 9);
    y[3] ^= salsa_rotl32(y[2] + y[1], 13);
    y[0] ^= salsa_rotl32(y[3] + y[2], 18);
}

static void salsa_qr_at(uint32_t *x, int a, int b, int c, int d) {
    uint32_t y[4] = {x[a], x[b], x[c], x[d]};
    orc_salsa_quarterround(y);
    x[a] = y[0]; x[b] = y[1]; x[c] = y[2]; x[d] = y[3];
}

void orc_salsa_hash(const uint8_t in[64], uint8_t out[64]) {
    uint32_t w[16], x[16];
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)in[4 * i] | ((uint32_t)in[4 * i + 1] << 8) | ((uint32_t)in[4 * i + 2] << 16) |
               ((uint32_t)in[4 * i + 3] << 24);
    memcpy(x, w, sizeof(x));
    for (int r = 0; r < 10; r++) {
        /* columnround: columns (0,4,8,12), (5,9,13,1), (10,14,2,6), (15,3,7,11) */
        salsa_qr_at(x, 0, 4, 8, 12);
        salsa_qr_at(x, 5, 9, 13, 1);
        salsa_qr_at(x, 10, 14, 2, 6);
        salsa_qr_at(x, 15, 3, 7, 11);
        /* rowround: rows (0,1,2,3), (5,6,7,4), (10,11,8,9), (15,12,13,14) */
        salsa_qr_at(x, 0, 1, 2, 3);
        salsa_qr_at(x, 5, 6, 7, 4);
        salsa_qr_at(x, 10, 11, 8, 9);
        salsa_qr_at(x, 15, 12, 13, 14);
    }
    for (int i = 0; i < 16; i++) {
        uint32_t v = x[i] + w[i];
        out[4 * i] = (uint8_t)v; out[4 * i + 1] = (uint8_t)(v >> 8);
        out[4 * i + 2] = (uint8_t)(v >> 16); out[4 * i + 3] = (uint8_t)(v >> 24);
    }
}

static void seeded_block64(const uint8_t *seed, u64 nonce, u64 out[8]) {
    uint8_t in[64], h[64];
    memcpy(in, seed, 32);
    for (int i = 0; i < 8; i++) in[32 + i] = (uint8_t)(nonce >> (8 * i));
    memcpy(in + 40, seed + 32, 24);
    orc_salsa_hash(in, h);
    for (int k = 0; k < 8; k++) {
        u64 v = 0;
        for (int i = 7; i >= 0; i--) v = (v << 8) | h[8 * k + i];
        out[k] = v;
    }
}

void orc_sample_uniform_seeded(const orc_ctx *c, const uint8_t seed[64], size_t L, uint64_t *out) {
    size_t N = c->N, per = N / 8;
    for (size_t b = 0; b < per * L; b++) {
        u64 q = c->q[b / per];
        u64 max_accept = ~0ull - (~0ull % q) - 1;
        u64 r[8];
        seeded_block64(seed, b, r);
        u64 draws = 1;
        for (int i = 0; i < 8; i++) {
            while (r[i] > max_accept) {
                seeded_block64(seed, b + draws * (u64)N * L, r);
                draws++;
            }
            out[b * 8 + i] = r[i] % q;
        }
    }
}

/* ------------------------------------------------------------------ */
/* BFV decryption scale-and-round                                      */
/* ------------------------------------------------------------------ */

/* PhantomSecretKey::bfv_decrypt after the phase (src/secretkey.cu:571-635):
 * phase = canonical residues [size_Ql][N] of c0 + c1 s (+ c2 s^2), coefficient
 * form, at the level of the first size_Ql primes; out[N] = the plaintext. behz != 0: behz_decrypt_scale_and_round
 * (src/rns.cu:1008-1090; gamma = the largest 61-bit prime = 1 mod 2N, constants
 * src/rns.cu:330-390), with the final centring of perform_final_multiplication
 * (984-1006). Otherwise HPS (hps_decrypt_scale_and_round, src/rns.cu:1634-1690,
 * kernels 1519-1633; constants 583-673): FP64 sums with fused multiply-adds as
 * the device compiler contracts them, the form picked from qMSB, tMSB and the
 * bit count of size_Q exactly as the reference picks it. */
int orc_bfv_decrypt(const orc_ctx *c, size_t size_Ql, u64 t, int behz, const u64 *phase, u64 *out) {
    const size_t N = c->N, k = size_Ql;
    if (size_Ql < 1 || size_Ql > c->size_Q) return -1;
    const u64 *q = c->q;
    if (behz) {
        u64 g;
        if (orc_get_primes(N, 61, 1, &g) != 0) return -1;
        u64 neg_t = invmod(prod_mod(q, k, t), t), neg_g = invmod(prod_mod(q, k, g), g);
        neg_t = neg_t ? t - neg_t : 0;
        neg_g = neg_g ? g - neg_g : 0;
        const u64 ginv_t = invmod(g % t, t);
        for (size_t n = 0; n < N; n++) {
            u64 xt = 0, xg = 0;
            for (size_t i = 0; i < k; i++) {
                const u64 x = mulmod(phase[i * N + n], mulmod(t % q[i], g % q[i], q[i]), q[i]);
                const u64 y = mulmod(x, invmod(punctured_mod(q, k, i, q[i]), q[i]), q[i]);
                xt = addmod(xt, mulmod(y, punctured_mod(q, k, i, t), t), t);
                xg = addmod(xg, mulmod(y, punctured_mod(q, k, i, g), g), g);
            }
            xt = mulmod(xt, neg_t, t);
            xg = mulmod(xg, neg_g, g);
            const u64 v = xg > (g >> 1) ? addmod(xt, (g - xg) % t, t) : submod(xt, xg % t, t);
            out[n] = mulmod(v, ginv_t, t);
        }
        return 0;
    }
    u64 qmax = 0;   /* over all of Q; the size term over the level's primes (src/rns.cu:576-592) */
    for (size_t i = 0; i < c->size_Q; i++) qmax = q[i] > qmax ? q[i] : qmax;
    const int qmsb = bit_count(qmax), sizeqmsb = bit_count((u64)k), tmsb = bit_count(t);
    const int large = qmsb + sizeqmsb >= 52;
    const int h = qmsb >> 1;
    const int lazy = large ? (h + tmsb + sizeqmsb < 52) : (qmsb + tmsb + sizeqmsb < 52);
    u64 *om = (u64 *)malloc(k * 8), *omb = (u64 *)malloc(k * 8);
    double *fr = (double *)malloc(k * 8), *frb = (double *)malloc(k * 8);
    for (size_t i = 0; i < k; i++) {
        const u64 v = invmod(punctured_mod(q, k, i, q[i]), q[i]);
        const u128 big = (u128)t * v;
        om[i] = (u64)((big / q[i]) % t);
        fr[i] = (double)(u64)(big % q[i]) / (double)q[i];
        const u64 vb = (u64)(((u128)v << h) % q[i]);
        const u128 bigb = (u128)t * vb;
        omb[i] = (u64)((bigb / q[i]) % t);
        frb[i] = (double)(u64)(bigb % q[i]) / (double)q[i];
    }
    const double tinv = 1.0 / (double)t;
    for (size_t n = 0; n < N; n++) {
        double fs = 0.0;
        u64 is = 0;
        for (size_t i = 0; i < k; i++) {
            const u64 x = phase[i * N + n];
            if (!large) {
                fs = fma((double)x, fr[i], fs);
                is += lazy ? x * om[i] : mulmod(x, om[i], t);
            } else {
                const u64 hi = x >> h, lo = x & (((u64)1 << h) - 1);
                fs = fma((double)lo, fr[i], fs);
                fs = fma((double)hi, frb[i], fs);
                is += lazy ? lo * om[i] : mulmod(lo, om[i], t);
                is += lazy ? hi * omb[i] : mulmod(hi, omb[i], t);
            }
        }
        fs += (double)is;
        const u64 quot = (u64)(fs * tinv);
        fs -= (double)(t * quot);
        out[n] = (u64)llround(fs);
    }
    free(om);
    free(omb);
    free(fr);
    free(frb);
    return 0;
}
