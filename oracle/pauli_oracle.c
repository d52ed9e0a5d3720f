/*
 * pauli_oracle.c -- CPU restatement of the PauLIB PauliSum hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it to check or
 * time the HIP engine.  The engine itself never links, loads or calls it.
 *
 * Every function restates one reference function (paths relative to the
 * reference repository); the term order, the key order, the stable-sort
 * tie-breaking and the in-order segment sums are kept, because the parity
 * bar (bit-exact strings/phases, coefficients to 1e-12) depends on them.
 *
 * Data layout (same as the engine's device layout): plane-major SoA,
 *   x[k*cap + t], z[k*cap + t]  word k (qubits 64k..64k+63) of term t,
 *   f[t] flags (bit0 = i, bit1 = -1), v[t] coefficient.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   multiply / commutator   Cayley tables, scalar-vs-bulk KATs of the
 *                           reference tests, brute force against dense 2x2
 *                           Pauli matrices, oracle/_ref differential.
 *   sort_and_combine /      reference KATs (merge, sign separation, zero kept,
 *   canonicalize            idempotence, fold, zero drop, imaginary bit),
 *                           std::map-style brute force, oracle/_ref.
 *   Clifford H/S/CNOT       reference truth tables, dense-matrix conjugation.
 *   evolve                  reference closed forms (0, +-pi/2, pi, pi/3),
 *                           dense 2^n conjugation, oracle/_ref.
 *   propagate               statevector <0|U^dag O U|0> (no truncation),
 *                           oracle/_ref (with truncation).
 *   trim                    reference KATs with distinct magnitudes; the
 *                           tie order of std::nth_element is parity unpinned.
 *
 * Angles: evolve and apply_RX/RY/RZ take std::cos and std::sin of one angle,
 * which the reference's compilers merge into one call of the C library's
 * sincos (its results can differ from separate cos and sin calls in the
 * last bit); the oracle calls sincos itself (orc_cos_sin).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* cos and sin of an angle as the reference's build computes them */
static void orc_cos_sin(double th, double *c, double *s) { sincos(th, s, c); }

#define ORC_SIGN 0x02u
#define ORC_PHASE 0x01u

typedef struct orc_sum {
    int as;        /* words per plane (ArraySize) */
    int64_t nq;    /* logical qubit count (num_qubits_) */
    int64_t n;     /* terms */
    int64_t cap;   /* plane stride */
    uint64_t *x;
    uint64_t *z;
    uint8_t *f;
    double *v;
} orc_sum;

#define X(s, k, t) ((s)->x[(int64_t)(k) * (s)->cap + (t)])
#define Z(s, k, t) ((s)->z[(int64_t)(k) * (s)->cap + (t)])

static int popc64(uint64_t w) { return __builtin_popcountll(w); }
static int parity64(uint64_t w) { return __builtin_parityll(w); }

/* ---------------------------------------------------------------- storage */

static int reserve(orc_sum *s, int64_t cap) {
    if (cap <= s->cap) return 0;
    int64_t nc = s->cap ? s->cap : 16;
    while (nc < cap) nc *= 2;
    uint64_t *nx = (uint64_t *)calloc((size_t)(s->as * nc), sizeof(uint64_t));
    uint64_t *nz = (uint64_t *)calloc((size_t)(s->as * nc), sizeof(uint64_t));
    uint8_t *nf = (uint8_t *)calloc((size_t)nc, 1);
    double *nv = (double *)calloc((size_t)nc, sizeof(double));
    if (!nx || !nz || !nf || !nv) { free(nx); free(nz); free(nf); free(nv); return -1; }
    for (int k = 0; k < s->as; ++k) {
        if (s->n) {
            memcpy(nx + (int64_t)k * nc, s->x + (int64_t)k * s->cap, (size_t)s->n * 8);
            memcpy(nz + (int64_t)k * nc, s->z + (int64_t)k * s->cap, (size_t)s->n * 8);
        }
    }
    if (s->n) { memcpy(nf, s->f, (size_t)s->n); memcpy(nv, s->v, (size_t)s->n * 8); }
    free(s->x); free(s->z); free(s->f); free(s->v);
    s->x = nx; s->z = nz; s->f = nf; s->v = nv; s->cap = nc;
    return 0;
}

orc_sum *orc_sum_new(int as, int64_t nq) {
    if (as < 1 || as > 16) return NULL;
    orc_sum *s = (orc_sum *)calloc(1, sizeof(orc_sum));
    if (!s) return NULL;
    s->as = as; s->nq = nq;
    return s;
}

void orc_sum_free(orc_sum *s) {
    if (!s) return;
    free(s->x); free(s->z); free(s->f); free(s->v); free(s);
}

int orc_sum_as(const orc_sum *s) { return s->as; }
int64_t orc_sum_len(const orc_sum *s) { return s->n; }
int64_t orc_sum_nq(const orc_sum *s) { return s->nq; }
void orc_sum_set_nq(orc_sum *s, int64_t nq) { s->nq = nq; }

/* Replace the contents with n terms given as packed planes x[k*n + t]. */
int orc_sum_set(orc_sum *s, int64_t n, const uint64_t *x, const uint64_t *z,
                const uint8_t *f, const double *v) {
    s->n = 0;
    if (reserve(s, n)) return -1;
    for (int k = 0; k < s->as; ++k) {
        if (n) {
            memcpy(s->x + (int64_t)k * s->cap, x + (int64_t)k * n, (size_t)n * 8);
            memcpy(s->z + (int64_t)k * s->cap, z + (int64_t)k * n, (size_t)n * 8);
        }
    }
    if (n) { memcpy(s->f, f, (size_t)n); memcpy(s->v, v, (size_t)n * 8); }
    s->n = n;
    return 0;
}

/* Copy out as packed planes x[k*n + t]. */
void orc_sum_get(const orc_sum *s, uint64_t *x, uint64_t *z, uint8_t *f, double *v) {
    const int64_t n = s->n;
    for (int k = 0; k < s->as; ++k) {
        if (n) {
            memcpy(x + (int64_t)k * n, s->x + (int64_t)k * s->cap, (size_t)n * 8);
            memcpy(z + (int64_t)k * n, s->z + (int64_t)k * s->cap, (size_t)n * 8);
        }
    }
    if (n) { memcpy(f, s->f, (size_t)n); memcpy(v, s->v, (size_t)n * 8); }
}

static int push(orc_sum *s, const uint64_t *xw, const uint64_t *zw, uint8_t f, double v) {
    if (s->n == s->cap && reserve(s, s->n + 1)) return -1;
    const int64_t t = s->n;
    for (int k = 0; k < s->as; ++k) { X(s, k, t) = xw[k]; Z(s, k, t) = zw[k]; }
    s->f[t] = f; s->v[t] = v;
    s->n = t + 1;
    return 0;
}

/* ------------------------------------------------------- pair arithmetic */

/* PauliString::multiply, include/PauliStringCore.h:149-174: a <- a * b.
 * Bit-sliced mod-4 counter of the per-site phase; 2*popcount(c2) mod 4 is
 * 2*parity(c2), which is the form PauliSum::multiply uses
 * (include/PauliSumOperations.h:302). */
static uint8_t pair_product(int as, const uint64_t *ax, const uint64_t *az, uint8_t af,
                            const uint64_t *bx, const uint64_t *bz, uint8_t bf,
                            uint64_t *ox, uint64_t *oz) {
    uint64_t c1 = 0, c2 = 0;
    for (int k = 0; k < as; ++k) {
        const uint64_t xa = ax[k], za = az[k], xb = bx[k], zb = bz[k];
        ox[k] = xa ^ xb;
        oz[k] = za ^ zb;
        const uint64_t xbz = xa & zb;
        const uint64_t anti = (xb & za) ^ xbz;
        c2 ^= (c1 ^ xa ^ za ^ xb ^ zb ^ xbz) & anti;
        c1 ^= anti;
    }
    return (uint8_t)((af + bf + popc64(c1) + 2 * parity64(c2)) & 3);
}

/* PauliString::inner_product, include/PauliStringCore.h:125-136. */
static int commutes(int as, const uint64_t *ax, const uint64_t *az,
                    const uint64_t *bx, const uint64_t *bz) {
    uint64_t acc = 0;
    for (int k = 0; k < as; ++k) acc ^= (ax[k] & bz[k]) ^ (az[k] & bx[k]);
    return parity64(acc) == 0;
}

static void load_term(const orc_sum *s, int64_t t, uint64_t *xw, uint64_t *zw) {
    for (int k = 0; k < s->as; ++k) { xw[k] = X(s, k, t); zw[k] = Z(s, k, t); }
}

/* --------------------------------------------------------- bulk products */

/* PauliSum::multiply, include/PauliSumOperations.h:255-306.
 * out[i*n + j] = lhs[j] * rhs[i]; value = lhs.v[j] * rhs.v[i];
 * num_qubits_ stays the lhs value. */
int orc_multiply(orc_sum *lhs, const orc_sum *rhs) {
    if (lhs->as != rhs->as) return -2;
    const int as = lhs->as;
    const int64_t n = lhs->n, m = rhs->n;
    orc_sum out = {0};
    out.as = as; out.nq = lhs->nq;
    if (reserve(&out, n * m)) return -1;
    uint64_t ax[16], az[16], bx[16], bz[16], ox[16], oz[16];
    for (int64_t i = 0; i < m; ++i) {
        load_term(rhs, i, bx, bz);
        for (int64_t j = 0; j < n; ++j) {
            load_term(lhs, j, ax, az);
            const int64_t t = i * n + j;
            const uint8_t fl = pair_product(as, ax, az, lhs->f[j], bx, bz, rhs->f[i], ox, oz);
            for (int k = 0; k < as; ++k) { out.x[(int64_t)k * out.cap + t] = ox[k]; out.z[(int64_t)k * out.cap + t] = oz[k]; }
            out.f[t] = fl;
            out.v[t] = lhs->v[j] * rhs->v[i];
        }
    }
    out.n = n * m;
    free(lhs->x); free(lhs->z); free(lhs->f); free(lhs->v);
    lhs->x = out.x; lhs->z = out.z; lhs->f = out.f; lhs->v = out.v;
    lhs->cap = out.cap; lhs->n = out.n;
    return 0;
}

/* commutator_impl, include/PauliSumOperations.h:315-346, in the order a
 * single thread produces: i over a, j over b.  Anticommuting pairs only;
 * p = a_i * b_j, flags + 1 for i[A,B]; value (2 * a.v[i]) * b.v[j].
 * num_qubits_ is the max of both operands. */
orc_sum *orc_commutator(const orc_sum *a, const orc_sum *b, int by_i) {
    if (a->as != b->as) return NULL;
    const int as = a->as;
    orc_sum *out = orc_sum_new(as, a->nq > b->nq ? a->nq : b->nq);
    if (!out) return NULL;
    uint64_t ax[16], az[16], bx[16], bz[16], ox[16], oz[16];
    for (int64_t i = 0; i < a->n; ++i) {
        load_term(a, i, ax, az);
        for (int64_t j = 0; j < b->n; ++j) {
            load_term(b, j, bx, bz);
            if (commutes(as, ax, az, bx, bz)) continue;
            uint8_t fl = pair_product(as, ax, az, a->f[i], bx, bz, b->f[j], ox, oz);
            if (by_i) fl = (uint8_t)((fl + 1) & 3);
            if (push(out, ox, oz, fl, (2.0 * a->v[i]) * b->v[j])) { orc_sum_free(out); return NULL; }
        }
    }
    return out;
}

/* ------------------------------------------------------------- collation */

/* key_less, include/PauliSumOperations.h:376-386: x words from word 0, then
 * z words, then flags; unsigned compares. */
static int key_cmp(const orc_sum *s, int64_t a, int64_t b) {
    for (int k = 0; k < s->as; ++k) {
        const uint64_t p = X(s, k, a), q = X(s, k, b);
        if (p != q) return p < q ? -1 : 1;
    }
    for (int k = 0; k < s->as; ++k) {
        const uint64_t p = Z(s, k, a), q = Z(s, k, b);
        if (p != q) return p < q ? -1 : 1;
    }
    if (s->f[a] != s->f[b]) return s->f[a] < s->f[b] ? -1 : 1;
    return 0;
}

/* Stable bottom-up merge sort of idx by key_cmp (ties keep index order,
 * which is what std::stable_sort plus stable std::merge of the fixed chunk
 * grid produce, include/PauliSumOperations.h:388-414). */
static int stable_sort_idx(const orc_sum *s, int64_t *idx, int64_t n) {
    int64_t *tmp = (int64_t *)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
    if (!tmp) return -1;
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t p = lo, q = mid, o = lo;
            while (p < mid && q < hi) {
                if (key_cmp(s, idx[q], idx[p]) < 0) tmp[o++] = idx[q++];
                else tmp[o++] = idx[p++];
            }
            while (p < mid) tmp[o++] = idx[p++];
            while (q < hi) tmp[o++] = idx[q++];
        }
        memcpy(idx, tmp, (size_t)n * sizeof(int64_t));
    }
    free(tmp);
    return 0;
}

/* PauliSum::sort_and_combine, include/PauliSumOperations.h:368-461.
 * Stable sort, then one term per distinct (x, z, flags) whose value is
 * acc = 0; acc += v over the segment in sorted (= original index) order.
 * Exact zeros are kept. */
int orc_sort_and_combine(orc_sum *s) {
    const int64_t n = s->n;
    if (n == 0) return 0;
    int64_t *idx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!idx) return -1;
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    if (stable_sort_idx(s, idx, n)) { free(idx); return -1; }
    orc_sum out = {0};
    out.as = s->as; out.nq = s->nq;
    if (reserve(&out, n)) { free(idx); return -1; }
    int64_t m = 0;
    for (int64_t i = 0; i < n;) {
        int64_t e = i + 1;
        while (e < n && key_cmp(s, idx[e], idx[i]) == 0) ++e;
        const int64_t first = idx[i];
        double acc = 0.0;
        for (int64_t r = i; r < e; ++r) acc += s->v[idx[r]];
        for (int k = 0; k < s->as; ++k) {
            out.x[(int64_t)k * out.cap + m] = X(s, k, first);
            out.z[(int64_t)k * out.cap + m] = Z(s, k, first);
        }
        out.f[m] = s->f[first];
        out.v[m] = acc;
        ++m;
        i = e;
    }
    free(idx);
    out.n = m;
    free(s->x); free(s->z); free(s->f); free(s->v);
    s->x = out.x; s->z = out.z; s->f = out.f; s->v = out.v; s->cap = out.cap; s->n = m;
    return 0;
}

/* PauliSum::compact, include/PauliSumUtils.h:286-308: in-order stream
 * compaction by a keep mask. */
static void compact(orc_sum *s, const uint8_t *keep) {
    int64_t w = 0;
    for (int64_t i = 0; i < s->n; ++i) {
        if (!keep[i]) continue;
        if (w != i) {
            for (int k = 0; k < s->as; ++k) { X(s, k, w) = X(s, k, i); Z(s, k, w) = Z(s, k, i); }
            s->f[w] = s->f[i];
            s->v[w] = s->v[i];
        }
        ++w;
    }
    s->n = w;
}

/* PauliSum::canonicalize, include/PauliSumUtils.h:314-328: fold the sign
 * flag into the coefficient, sort_and_combine, drop exact zeros. */
int orc_canonicalize(orc_sum *s) {
    for (int64_t i = 0; i < s->n; ++i) {
        if (s->f[i] & ORC_SIGN) { s->f[i] ^= ORC_SIGN; s->v[i] = -s->v[i]; }
    }
    if (orc_sort_and_combine(s)) return -1;
    uint8_t *keep = (uint8_t *)malloc((size_t)(s->n ? s->n : 1));
    if (!keep) return -1;
    for (int64_t i = 0; i < s->n; ++i) keep[i] = (s->v[i] != 0.0);
    compact(s, keep);
    free(keep);
    return 0;
}

/* operator+= (add then sort_and_combine), include/PauliSumCore.h:209-237. */
int orc_add(orc_sum *lhs, const orc_sum *rhs) {
    if (lhs->as != rhs->as) return -2;
    uint64_t xw[16], zw[16];
    for (int64_t t = 0; t < rhs->n; ++t) {
        load_term(rhs, t, xw, zw);
        if (push(lhs, xw, zw, rhs->f[t], rhs->v[t])) return -1;
    }
    return orc_sort_and_combine(lhs);
}

/* ------------------------------------------------------------ observables */

/* PauliSum::zero_state_expectation, include/PauliSumUtils.h:162-177:
 * sum of sign * value over terms with no X/Y and no imaginary phase, in
 * term order. */
double orc_zero_state_expectation(const orc_sum *s) {
    double r = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        int has_xy = 0;
        for (int k = 0; k < s->as; ++k) if (X(s, k, i)) { has_xy = 1; break; }
        if (has_xy) continue;
        if (s->f[i] & ORC_PHASE) continue;
        const double sign = (s->f[i] & ORC_SIGN) ? -1.0 : 1.0;
        r += sign * s->v[i];
    }
    return r;
}

/* PauliSum::term_weight, include/PauliSumUtils.h:277-283. */
static int64_t term_weight(const orc_sum *s, int64_t t) {
    int64_t w = 0;
    for (int k = 0; k < s->as; ++k) w += popc64(X(s, k, t) | Z(s, k, t));
    return w;
}

/* PauliSum::truncate_weight, include/PauliSumUtils.h:386-393. */
int orc_truncate_weight(orc_sum *s, int64_t max_weight) {
    uint8_t *keep = (uint8_t *)malloc((size_t)(s->n ? s->n : 1));
    if (!keep) return -1;
    for (int64_t i = 0; i < s->n; ++i) keep[i] = term_weight(s, i) <= max_weight;
    compact(s, keep);
    free(keep);
    return 0;
}

/* PauliSum::cutoff, include/PauliSumUtils.h:377-384: keep |c| >= eps. */
int orc_cutoff(orc_sum *s, double eps) {
    uint8_t *keep = (uint8_t *)malloc((size_t)(s->n ? s->n : 1));
    if (!keep) return -1;
    for (int64_t i = 0; i < s->n; ++i) keep[i] = fabs(s->v[i]) >= eps;
    compact(s, keep);
    free(keep);
    return 0;
}

/* PauliSum::trim(max_terms), include/PauliSumUtils.h:330-343: keep the
 * max_terms largest |c| in term order.  std::nth_element leaves the choice
 * among equal magnitudes to the library; here ties keep the lower index
 * (parity unpinned for ties). */
static const orc_sum *g_trim_sum;
static int trim_cmp(const void *pa, const void *pb) {
    const int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    const double ma = fabs(g_trim_sum->v[a]), mb = fabs(g_trim_sum->v[b]);
    if (ma != mb) return ma > mb ? -1 : 1;
    return a < b ? -1 : (a > b);
}
int orc_trim(orc_sum *s, int64_t max_terms) {
    const int64_t n = s->n;
    if (n <= max_terms) return 0;
    int64_t *idx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    uint8_t *keep = (uint8_t *)calloc((size_t)n, 1);
    if (!idx || !keep) { free(idx); free(keep); return -1; }
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    g_trim_sum = s;
    qsort(idx, (size_t)n, sizeof(int64_t), trim_cmp);
    for (int64_t i = 0; i < max_terms; ++i) keep[idx[i]] = 1;
    compact(s, keep);
    free(idx); free(keep);
    return 0;
}

/* PauliSum::trim(max_terms, keep), include/PauliSumUtils.h:345-375: terms
 * whose (x, z) pattern occurs in keep are always kept; the others compete
 * for max(0, max_terms - protected) places by |c| (ties: lower index, as
 * orc_trim; parity unpinned for ties).  Kept terms stay in term order. */
static int same_pattern(const orc_sum *a, int64_t i, const orc_sum *b, int64_t j) {
    for (int k = 0; k < a->as; ++k)
        if (X(a, k, i) != X(b, k, j) || Z(a, k, i) != Z(b, k, j)) return 0;
    return 1;
}
int orc_trim_keep(orc_sum *s, int64_t max_terms, const orc_sum *keep) {
    if (keep && keep->as != s->as) return -2;
    const int64_t n = s->n;
    if (n <= max_terms) return 0;
    uint8_t *mask = (uint8_t *)calloc((size_t)(n ? n : 1), 1);
    int64_t *idx = (int64_t *)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
    if (!mask || !idx) { free(mask); free(idx); return -1; }
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        int hit = 0;
        for (int64_t j = 0; keep && j < keep->n && !hit; ++j) hit = same_pattern(s, i, keep, j);
        if (hit) mask[i] = 1;
        else idx[m++] = i;
    }
    const int64_t prot = n - m;
    const int64_t budget = max_terms > prot ? max_terms - prot : 0;
    g_trim_sum = s;
    qsort(idx, (size_t)m, sizeof(int64_t), trim_cmp);
    for (int64_t i = 0; i < budget && i < m; ++i) mask[idx[i]] = 1;
    compact(s, mask);
    free(mask); free(idx);
    return 0;
}

/* PauliSum::scale, include/PauliSumUtils.h:395-398. */
int orc_scale(orc_sum *s, double c) {
    for (int64_t i = 0; i < s->n; ++i) s->v[i] *= c;
    return 0;
}

/* operator-=, include/PauliSumCore.h:245-250: append rhs with -1 * value
 * (append_scaled, :221-232), then sort_and_combine. */
int orc_sub(orc_sum *lhs, const orc_sum *rhs) {
    if (lhs->as != rhs->as) return -2;
    const int64_t m = rhs->n;  /* rhs may be lhs: take its length first */
    uint64_t xw[16], zw[16];
    for (int64_t t = 0; t < m; ++t) {
        load_term(rhs, t, xw, zw);
        if (push(lhs, xw, zw, rhs->f[t], -1.0 * rhs->v[t])) return -1;
    }
    return orc_sort_and_combine(lhs);
}

/* PauliSum::lookup, include/PauliSumUtils.h:253-268: every term with the
 * pattern (x, z) in term order; phase relative to flags, +-i skipped. */
double orc_lookup(const orc_sum *s, const uint64_t *x, const uint64_t *z, uint8_t flags) {
    double r = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        int match = 1;
        for (int k = 0; k < s->as && match; ++k) match = X(s, k, i) == x[k] && Z(s, k, i) == z[k];
        if (!match) continue;
        const unsigned rel = ((unsigned)s->f[i] - (unsigned)flags) & 3u;
        if (rel & ORC_PHASE) continue;
        r += (rel & ORC_SIGN) ? -s->v[i] : s->v[i];
    }
    return r;
}

/* PauliSum::trace, include/PauliSumUtils.h:229-243. */
double orc_trace(const orc_sum *s) {
    double r = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        int ident = 1;
        for (int k = 0; k < s->as && ident; ++k) ident = (X(s, k, i) | Z(s, k, i)) == 0;
        if (!ident || (s->f[i] & ORC_PHASE)) continue;
        r += ((s->f[i] & ORC_SIGN) ? -1.0 : 1.0) * s->v[i];
    }
    return ldexp(r, (int)s->nq);
}

/* PauliSum::frobenius_norm2, include/PauliSumUtils.h:401-410. */
double orc_frobenius_norm2(const orc_sum *s) {
    double r = 0.0;
    for (int64_t i = 0; i < s->n; ++i) r += s->v[i] * s->v[i];
    return r;
}

/* ------------------------------------------------------------ Clifford */

/* PauliSum::apply_H, include/PauliSumClifford.h:43-62. */
int orc_apply_H(orc_sum *s, int64_t c) {
    const int w = (int)(c / 64);
    const uint64_t m = 1ull << (c % 64);
    if (w >= s->as) return -2;
    for (int64_t i = 0; i < s->n; ++i) {
        uint64_t xw = X(s, w, i), zw = Z(s, w, i);
        const unsigned xc = (xw & m) != 0, zc = (zw & m) != 0;
        s->f[i] ^= (uint8_t)((xc & zc) << 1);
        xw ^= zw & m; zw ^= xw & m; xw ^= zw & m;
        X(s, w, i) = xw; Z(s, w, i) = zw;
    }
    return 0;
}

/* PauliSum::apply_S, include/PauliSumClifford.h:64-81. */
int orc_apply_S(orc_sum *s, int64_t c) {
    const int w = (int)(c / 64);
    const uint64_t m = 1ull << (c % 64);
    if (w >= s->as) return -2;
    for (int64_t i = 0; i < s->n; ++i) {
        const uint64_t xw = X(s, w, i), zw = Z(s, w, i);
        const unsigned xc = (xw & m) != 0, zc = (zw & m) != 0;
        s->f[i] ^= (uint8_t)((xc & zc) << 1);
        Z(s, w, i) = zw ^ (xw & m);
    }
    return 0;
}

/* PauliSum::apply_CNOT, include/PauliSumClifford.h:212-239. */
int orc_apply_CNOT(orc_sum *s, int64_t c, int64_t t) {
    const int wc = (int)(c / 64), wt = (int)(t / 64);
    const uint64_t mc = 1ull << (c % 64), mt = 1ull << (t % 64);
    if (wc >= s->as || wt >= s->as) return -2;
    for (int64_t i = 0; i < s->n; ++i) {
        const uint64_t xc = (X(s, wc, i) & mc) != 0, zc = (Z(s, wc, i) & mc) != 0;
        const uint64_t xt = (X(s, wt, i) & mt) != 0, zt = (Z(s, wt, i) & mt) != 0;
        s->f[i] ^= (uint8_t)((xc & zt & (1 ^ xt ^ zc)) << 1);
        X(s, wt, i) ^= (0 - xc) & mt;
        Z(s, wc, i) ^= (0 - zt) & mc;
    }
    return 0;
}

int orc_apply_gate(orc_sum *s, int g, int64_t c, int64_t t);
#define ORC_NGATES 21

/* ------------------------------------------------------------- rotation */

static int is_close_one(double v) { return fabs(v - 1.0) <= 1e-14; } /* :19-21 */

/* PauliSum::multiply(i, ps), include/PauliSumOperations.h:185-209. */
static void term_times(orc_sum *s, int64_t i, const uint64_t *px, const uint64_t *pz, uint8_t pf) {
    uint64_t ax[16], az[16], ox[16], oz[16];
    load_term(s, i, ax, az);
    const uint8_t fl = pair_product(s->as, ax, az, s->f[i], px, pz, pf, ox, oz);
    for (int k = 0; k < s->as; ++k) { X(s, k, i) = ox[k]; Z(s, k, i) = oz[k]; }
    s->f[i] = fl;
}

/* One rotation of PauliSum::evolve, include/PauliSumOperations.h:106-183,
 * with cos/sin already taken (the caller passes std::cos/std::sin of the
 * full angle from the host libm, :113-114). ps = P with the i bit flipped.
 * General path: anticommuting term j is scaled by cos in place and the term
 * (Q_j * ps) with value (-sin) * v_j is appended in j order; then
 * operator+= (append + sort_and_combine). */
int orc_evolve_cs(orc_sum *s, const uint64_t *px, const uint64_t *pz, uint8_t pflags,
                  double cos_t, double sin_t) {
    const int as = s->as;
    const uint8_t psf = (uint8_t)(pflags ^ ORC_PHASE);
    const int64_t m = s->n;
    uint64_t ax[16], az[16];
    if (is_close_one(cos_t)) return 0;
    if (is_close_one(-cos_t)) {
        for (int64_t j = 0; j < m; ++j) {
            load_term(s, j, ax, az);
            if (commutes(as, ax, az, px, pz)) continue;
            s->f[j] ^= ORC_SIGN;
        }
        return 0;
    }
    if (is_close_one(sin_t)) {
        for (int64_t j = 0; j < m; ++j) {
            load_term(s, j, ax, az);
            if (commutes(as, ax, az, px, pz)) continue;
            s->f[j] ^= ORC_SIGN;
            term_times(s, j, px, pz, psf);
        }
        return 0;
    }
    if (is_close_one(-sin_t)) {
        for (int64_t j = 0; j < m; ++j) {
            load_term(s, j, ax, az);
            if (commutes(as, ax, az, px, pz)) continue;
            term_times(s, j, px, pz, psf);
        }
        return 0;
    }
    orc_sum *nw = orc_sum_new(as, s->nq);
    if (!nw) return -1;
    uint64_t ox[16], oz[16];
    for (int64_t j = 0; j < m; ++j) {
        load_term(s, j, ax, az);
        if (commutes(as, ax, az, px, pz)) continue;
        const double orig = s->v[j];
        s->v[j] *= cos_t;
        const uint8_t fl = pair_product(as, ax, az, s->f[j], px, pz, psf, ox, oz);
        if (push(nw, ox, oz, fl, -sin_t * orig)) { orc_sum_free(nw); return -1; }
    }
    const int rc = orc_add(s, nw);
    orc_sum_free(nw);
    return rc;
}

/* PauliSum::evolve over every rotation term in order, with std::cos and
 * std::sin of the term's coefficient (the angle). */
int orc_evolve(orc_sum *s, const orc_sum *rot) {
    if (s->as != rot->as) return -2;
    uint64_t px[16], pz[16];
    for (int64_t r = 0; r < rot->n; ++r) {
        load_term(rot, r, px, pz);
        double c, sn;
        orc_cos_sin(rot->v[r], &c, &sn);
        const int rc = orc_evolve_cs(s, px, pz, rot->f[r], c, sn);
        if (rc) return rc;
    }
    return 0;
}

/* apply_truncation, include/PauliPropagation.h:35-47.  max_weight < 0,
 * cutoff <= 0 and max_terms < 0 mean "not set". */
int orc_apply_truncation_keep(orc_sum *s, int64_t max_weight, double cutoff, int64_t max_terms,
                              const orc_sum *keep) {
    int rc = orc_canonicalize(s);
    if (!rc && max_weight >= 0) rc = orc_truncate_weight(s, max_weight);
    if (!rc && cutoff > 0.0) rc = orc_cutoff(s, cutoff);
    if (!rc && max_terms >= 0) rc = keep ? orc_trim_keep(s, max_terms, keep) : orc_trim(s, max_terms);
    return rc;
}
int orc_apply_truncation(orc_sum *s, int64_t max_weight, double cutoff, int64_t max_terms) {
    return orc_apply_truncation_keep(s, max_weight, cutoff, max_terms, NULL);
}

/* propagate, include/PauliPropagation.h:53-63: gates in reverse order, each
 * a one-term evolve followed by apply_truncation. */
int orc_propagate(orc_sum *s, const orc_sum *rot, int64_t max_weight, double cutoff,
                  int64_t max_terms) {
    if (s->as != rot->as) return -2;
    uint64_t px[16], pz[16];
    for (int64_t l = rot->n; l-- > 0;) {
        load_term(rot, l, px, pz);
        double c, sn;
        orc_cos_sin(rot->v[l], &c, &sn);
        int rc = orc_evolve_cs(s, px, pz, rot->f[l], c, sn);
        if (!rc) rc = orc_apply_truncation(s, max_weight, cutoff, max_terms);
        if (rc) return rc;
    }
    return 0;
}

/* Heisenberg driver for a mixed Clifford + rotation program, the shape of
 * the brickwork workloads: the reference ships no such driver, so this
 * composes its methods exactly as a caller would -- ops are given in
 * application order and processed last to first; a Clifford op calls the
 * bulk method (apply_H / apply_S / apply_CNOT); a rotation op is
 * propagate()'s body: one-term evolve then apply_truncation.
 *   kind[l]: 0 = H(q0), 1 = S(q0), 2 = CNOT(q0, q1), 3 = rotation by
 *   theta[l] about the Pauli word (px, pz)[l*as .. l*as+as-1].
 * cos/sin of the angles come from the host libm. */
int orc_run_program(orc_sum *s, int64_t nops, const int32_t *kind, const int64_t *q0,
                    const int64_t *q1, const uint64_t *px, const uint64_t *pz,
                    const double *theta, int64_t max_weight, double cutoff, int64_t max_terms) {
    const int as = s->as;
    for (int64_t l = nops; l-- > 0;) {
        int rc = 0;
        switch (kind[l]) {
            case 0: rc = orc_apply_H(s, q0[l]); break;
            case 1: rc = orc_apply_S(s, q0[l]); break;
            case 2: rc = orc_apply_CNOT(s, q0[l], q1[l]); break;
            case 3:
            {
                double c, sn;
                orc_cos_sin(theta[l], &c, &sn);
                rc = orc_evolve_cs(s, px + l * as, pz + l * as, 0, c, sn);
            }
                if (!rc) rc = orc_apply_truncation(s, max_weight, cutoff, max_terms);
                break;
            default:
                /* kinds 16 + gate id: any gate of orc_apply_gate */
                if (kind[l] < 16 || kind[l] >= 16 + ORC_NGATES) return -2;
                rc = orc_apply_gate(s, kind[l] - 16, q0[l], q1[l]);
                break;
        }
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------- extended Clifford set */

/* Gate ids shared with the engine's plh_apply_clifford and ref_bridge.cpp:
 *   0 I  1 X  2 Y  3 Z  4 H  5 S  6 H_YZ  7 H_XY  8 sqrt_X  9 sqrt_Y
 *  10 CNOT 11 CZ 12 CY 13 XC_X 14 XC_Z 15 XC_Y 16 YC_X 17 YC_Z 18 YC_Y
 *  19 SWAP 20 iSWAP
 * Each case restates the reference's per-term bit update; the qubit bits of
 * one term are read before any is written, so control and target may share
 * a word (as in the reference, which takes no __restrict__ for them). */
static int gate_is_two(int g) { return g >= 10; }

static void gate_term(orc_sum *s, int g, int64_t i, int wc, uint64_t mc, int wt, uint64_t mt) {
    const uint64_t xc = (X(s, wc, i) & mc) != 0, zc = (Z(s, wc, i) & mc) != 0;
    uint64_t xt = 0, zt = 0;
    if (gate_is_two(g)) { xt = (X(s, wt, i) & mt) != 0; zt = (Z(s, wt, i) & mt) != 0; }
    uint64_t sg = 0;                 /* sign flip (flags bit 1) */
    uint64_t dxc = 0, dzc = 0, dxt = 0, dzt = 0;  /* bits toggled */
    switch (g) {
        case 0: return;                                        /* apply_I :208-210 */
        case 1: sg = zc; break;                                /* apply_X :162-174 */
        case 2: sg = xc ^ zc; break;                           /* apply_Y :177-190 */
        case 3: sg = xc; break;                                /* apply_Z :193-205 */
        case 4: sg = xc & zc; dxc = xc ^ zc; dzc = xc ^ zc; break;   /* apply_H :44-62 (swap) */
        case 5: sg = xc & zc; dzc = xc; break;                 /* apply_S :65-81 */
        case 6: sg = xc & (1 ^ zc); dxc = zc; break;           /* apply_H_YZ :84-100 */
        case 7: sg = (1 ^ xc) & zc; dzc = xc; break;           /* apply_H_XY :103-119 */
        case 8: sg = (1 ^ xc) & zc; dxc = zc; break;           /* apply_sqrt_X :122-138 */
        case 9: sg = xc & (1 ^ zc); dxc = xc ^ zc; dzc = xc ^ zc; break; /* apply_sqrt_Y :141-159 */
        case 10: sg = xc & zt & (1 ^ xt ^ zc); dxt = xc; dzc = zt; break; /* apply_CNOT :213-239 */
        case 11: sg = xc & xt & (zc ^ zt); dzt = xc; dzc = xt; break;     /* apply_CZ :242-266 */
        case 12: /* apply_CY :269-294 */
            sg = (xc & zc & zt & (1 ^ xt)) | (xc & xt & (1 ^ zc) & (1 ^ zt));
            dzc = xt ^ zt; dxt = xc; dzt = xc; break;
        case 13: /* apply_XC_X :297-321 */
            sg = (xc & zc & zt & (1 ^ xt)) | (xt & zc & zt & (1 ^ xc));
            dxc = zt; dxt = zc; break;
        case 14: /* apply_XC_Z :324-348 */
            sg = (xc & xt & zc & zt) | (xt & zc & (1 ^ xc) & (1 ^ zt));
            dxc = xt; dzt = zc; break;
        case 15: /* apply_XC_Y :351-376 */
            sg = (xc & xt & zc & (1 ^ zt)) | (zc & zt & (1 ^ xc) & (1 ^ xt));
            dxc = xt ^ zt; dxt = zc; dzt = zc; break;
        case 16: /* apply_YC_X :379-404 */
            sg = (xc & xt & zt & (1 ^ zc)) | (zc & zt & (1 ^ xc) & (1 ^ xt));
            dxc = zt; dzc = zt; dxt = xc ^ zc; break;
        case 17: /* apply_YC_Z :407-432 */
            sg = (xt & zc & zt & (1 ^ xc)) | (xc & xt & (1 ^ zc) & (1 ^ zt));
            dxc = xt; dzc = xt; dzt = xc ^ zc; break;
        case 18: /* apply_YC_Y :435-462 */
            sg = (xc & zt & (1 ^ xt) & (1 ^ zc)) | (xt & zc & (1 ^ xc) & (1 ^ zt));
            dxc = xt ^ zt; dzc = xt ^ zt; dxt = xc ^ zc; dzt = xc ^ zc; break;
        case 19: /* apply_SWAP :465-490 */
            dxc = xc ^ xt; dxt = xc ^ xt; dzc = zc ^ zt; dzt = zc ^ zt; break;
        default: return;
    }
    s->f[i] ^= (uint8_t)(sg << 1);
    X(s, wc, i) ^= (0 - dxc) & mc;
    Z(s, wc, i) ^= (0 - dzc) & mc;
    if (gate_is_two(g)) {
        X(s, wt, i) ^= (0 - dxt) & mt;
        Z(s, wt, i) ^= (0 - dzt) & mt;
    }
}

/* PauliSum::apply_<gate>, include/PauliSumClifford.h:44-498.  iSWAP is the
 * reference's literal sequence SWAP, CZ, S(c), S(t) (:493-498).  Returns -2
 * for an unknown gate, a qubit outside the words, or c == t. */
int orc_apply_gate(orc_sum *s, int g, int64_t c, int64_t t) {
    if (g < 0 || g >= ORC_NGATES || c < 0 || c >= 64 * (int64_t)s->as) return -2;
    if (gate_is_two(g) && (t < 0 || t >= 64 * (int64_t)s->as || t == c)) return -2;
    if (g == 20) {
        int rc = orc_apply_gate(s, 19, c, t);
        if (!rc) rc = orc_apply_gate(s, 11, c, t);
        if (!rc) rc = orc_apply_gate(s, 5, c, 0);
        if (!rc) rc = orc_apply_gate(s, 5, t, 0);
        return rc;
    }
    const int wc = (int)(c / 64), wt = gate_is_two(g) ? (int)(t / 64) : 0;
    const uint64_t mc = 1ull << (c % 64), mt = gate_is_two(g) ? 1ull << (t % 64) : 0;
    for (int64_t i = 0; i < s->n; ++i) gate_term(s, g, i, wc, mc, wt, mt);
    return 0;
}

/* PauliSum::apply_RZ / apply_RX / apply_RY, include/PauliSumOperations.h:
 * 28-104 (axis 3 = Z, 1 = X, 2 = Y): half-angle cos/sin; each term that
 * anticommutes with the axis Pauli at c keeps value*cos_half, and a copy
 * with value*sin_half, left-multiplied by i*P_c (multiply_iZ / iX / iY,
 * include/PauliSumClifford.h:15-40), is appended in term order; then
 * operator+= (append + sort_and_combine). */
int orc_apply_rot_axis(orc_sum *s, int axis, int64_t c, double angle) {
    if (axis < 1 || axis > 3 || c < 0 || c >= 64 * (int64_t)s->as) return -2;
    double cos_half, sin_half;
    orc_cos_sin(angle / 2, &cos_half, &sin_half);
    const int w = (int)(c / 64);
    const uint64_t m = 1ull << (c % 64);
    orc_sum *nw = orc_sum_new(s->as, s->nq);
    if (!nw) return -1;
    uint64_t xw[16], zw[16];
    for (int64_t i = 0; i < s->n; ++i) {
        const uint64_t xc = (X(s, w, i) & m) != 0, zc = (Z(s, w, i) & m) != 0;
        const uint64_t anti = axis == 3 ? xc : axis == 1 ? zc : (xc ^ zc);
        if (!anti) continue;
        const double t_val = s->v[i] * sin_half;
        s->v[i] *= cos_half;
        load_term(s, i, xw, zw);
        uint8_t f = s->f[i];
        if (axis == 1) { f ^= (uint8_t)(xc << 1); xw[w] ^= m; }            /* multiply_iX */
        else if (axis == 3) { f ^= (uint8_t)((1 ^ zc) << 1); zw[w] ^= m; } /* multiply_iZ */
        else { f ^= (uint8_t)((1 ^ xc) << 1); xw[w] ^= m; zw[w] ^= m; }    /* multiply_iY */
        if (push(nw, xw, zw, f, t_val)) { orc_sum_free(nw); return -1; }
    }
    const int rc = orc_add(s, nw);
    orc_sum_free(nw);
    return rc;
}

/* PauliSum::elementwise_multiply, include/PauliSumOperations.h:220-250:
 * term i <- term_i * rhs_i (phase as PauliString::multiply), value
 * v_i * rhs.v_i.  The reference reads rhs[i] for every i < n; here a rhs of
 * another length or word count is refused (-2). */
int orc_elementwise_multiply(orc_sum *s, const orc_sum *rhs) {
    if (s->as != rhs->as || s->n != rhs->n) return -2;
    uint64_t ax[16], az[16], bx[16], bz[16], ox[16], oz[16];
    for (int64_t i = 0; i < s->n; ++i) {
        load_term(s, i, ax, az);
        load_term(rhs, i, bx, bz);
        s->f[i] = pair_product(s->as, ax, az, s->f[i], bx, bz, rhs->f[i], ox, oz);
        for (int k = 0; k < s->as; ++k) { X(s, k, i) = ox[k]; Z(s, k, i) = oz[k]; }
        s->v[i] *= rhs->v[i];
    }
    return 0;
}
