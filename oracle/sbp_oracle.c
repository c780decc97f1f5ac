/*
 * sbp_oracle.c -- CPU restatement of vtsynergy/SBP's hot path under the
 * replay contract described in sbp_oracle.h.  Test infrastructure only:
 * it is the parity checker and the timed CPU baseline.
 *
 * Structure follows the reference closely on purpose (a Delta of four maps
 * with zero-net entries kept, "entries" plus "unchanged cells" loops,
 * incremental commits), so that the GPU's histogram-based formulation is
 * checked against an independent derivation.
 */
#define _GNU_SOURCE
#include "sbp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as    */
/* easy as 1, 2, 3", SC'11).                                                 */
/* ------------------------------------------------------------------------- */
void orc_philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

uint32_t orc_uniform_below(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

double orc_unit_double(uint32_t a, uint32_t b) {
    uint64_t u = (((uint64_t)a << 32) | b) >> 11;
    return (double)u * 0x1p-53;
}

static int32_t random_block(int32_t current, int32_t B, uint32_t w) {
    /* common::propose_random_block (src/common.cpp:200-209): candidates = U[0, B-2]; with B < 2 there is no
     * other block (U[0, -1] is undefined in the reference): the current block, no move */
    if (B < 2) return current;
    int32_t proposed = (int32_t)orc_uniform_below(w, (uint32_t)(B - 1));
    if (proposed >= current) proposed++;
    return proposed;
}

/* glibc logf over a contiguous range of float bit patterns (for the device logf check). */
void orc_logf_range(uint32_t first_bits, uint32_t count, float *out) {
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t bits = first_bits + i;
        float x;
        memcpy(&x, &bits, 4);
        out[i] = logf(x);
    }
}

/* ------------------------------------------------------------------------- */
/* 128-bit fixed point accumulation (scale 2^64, truncation toward zero).    */
/* ------------------------------------------------------------------------- */
typedef __int128 fx128;

static fx128 fx_from(double d) {
    uint64_t bits;
    memcpy(&bits, &d, 8);
    int neg = (int)(bits >> 63);
    int e = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & 0xfffffffffffffull;
    if (e == 0) return 0; /* zero or subnormal: below the 2^-64 grid */
    m |= 1ull << 52;
    int shift = e - 1075 + 64; /* value * 2^64 = m * 2^shift */
    unsigned __int128 mag;
    if (shift >= 0) {
        if (shift > 74) abort(); /* |d| >= 2^63: outside the contract's range */
        mag = (unsigned __int128)m << shift;
    } else if (shift > -64) {
        mag = (unsigned __int128)(m >> (-shift));
    } else {
        mag = 0;
    }
    fx128 v = (fx128)mag;
    return neg ? -v : v;
}

static double fx_to(fx128 v) {
    int64_t hi = (int64_t)(v >> 64);
    uint64_t lo = (uint64_t)v;
    return (double)hi + (double)lo * 0x1p-64;
}

void orc_fx_from_double(double d, uint64_t out[2]) {
    fx128 v = fx_from(d);
    out[0] = (uint64_t)v;
    out[1] = (uint64_t)(v >> 64);
}
double orc_fx_to_double(int64_t hi, uint64_t lo) { return (double)hi + (double)lo * 0x1p-64; }

/* fastlog (include/fastlog.hpp:21-30): logf of the integer, 0 for 0. */
static inline double fastlog(int64_t x) {
    if (x == 0) return 0.0;
    return (double)logf((float)x);
}

/* common::cell_entropy (include/common.hpp:39-49). */
static inline double cell_entropy(int64_t value, int64_t d_in, int64_t d_out) {
    if (value == 0) return 0.0;
    double e = (double)value * (fastlog(value) - (fastlog(d_in) + fastlog(d_out)));
    if (isnan(e) || isinf(e)) abort();
    return e;
}

/* ------------------------------------------------------------------------- */
/* Sorted sparse vectors (one per row and one per column of the block matrix) */
/* ------------------------------------------------------------------------- */
typedef struct { int32_t *idx; int64_t *val; int32_t len, cap; } svec;

static int32_t sv_find(const svec *v, int32_t key) {
    int32_t lo = 0, hi = v->len;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (v->idx[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}
static int64_t sv_get(const svec *v, int32_t key) {
    int32_t p = sv_find(v, key);
    return (p < v->len && v->idx[p] == key) ? v->val[p] : 0;
}
/* Adds delta; erases the cell when it reaches zero (DictTransposeMatrix::sub / update_edge_counts). */
static void sv_add(svec *v, int32_t key, int64_t delta) {
    int32_t p = sv_find(v, key);
    if (p < v->len && v->idx[p] == key) {
        v->val[p] += delta;
        if (v->val[p] == 0) {
            memmove(v->idx + p, v->idx + p + 1, sizeof(int32_t) * (size_t)(v->len - p - 1));
            memmove(v->val + p, v->val + p + 1, sizeof(int64_t) * (size_t)(v->len - p - 1));
            v->len--;
        }
        return;
    }
    if (delta == 0) return;
    if (v->len == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 4;
        v->idx = (int32_t *)realloc(v->idx, sizeof(int32_t) * (size_t)v->cap);
        v->val = (int64_t *)realloc(v->val, sizeof(int64_t) * (size_t)v->cap);
    }
    memmove(v->idx + p + 1, v->idx + p, sizeof(int32_t) * (size_t)(v->len - p));
    memmove(v->val + p + 1, v->val + p, sizeof(int64_t) * (size_t)(v->len - p));
    v->idx[p] = key;
    v->val[p] = delta;
    v->len++;
}
static void sv_copy(svec *dst, const svec *src) {
    dst->len = src->len;
    dst->cap = src->len;
    dst->idx = src->len ? (int32_t *)malloc(sizeof(int32_t) * (size_t)src->len) : NULL;
    dst->val = src->len ? (int64_t *)malloc(sizeof(int64_t) * (size_t)src->len) : NULL;
    if (src->len) {
        memcpy(dst->idx, src->idx, sizeof(int32_t) * (size_t)src->len);
        memcpy(dst->val, src->val, sizeof(int64_t) * (size_t)src->len);
    }
}

/* ------------------------------------------------------------------------- */
/* Blockmodel (include/blockmodel/blockmodel.hpp:46-219)                      */
/* ------------------------------------------------------------------------- */
struct orc_bm {
    int32_t B, N;
    int32_t *assign;
    int32_t *size;
    int64_t *d_out, *d_in, *d_tot;
    svec *row, *col;
    double overall_entropy;
    int64_t num_to_merge;
    int empty;
    int32_t np;  /* entropy model: 0 parametric (entropy::mdl), 1 nonparametric (entropy::nonparametric) */
    int32_t dc;  /* nonparametric with degree correction (args.degreecorrected) */
    svec *hin, *hout;  /* dc: per block, vertex in-/out-degree -> number of its vertices (degree histograms) */
    const orc_graph *g;  /* dc: the graph the histograms are keyed by */
};

static void bm_add(orc_bm *bm, int32_t r, int32_t c, int64_t d) {
    sv_add(&bm->row[r], c, d);
    sv_add(&bm->col[c], r, d);
}
static inline int64_t bm_get(const orc_bm *bm, int32_t r, int32_t c) {
    const svec *a = &bm->row[r], *b = &bm->col[c];
    return a->len <= b->len ? sv_get(a, c) : sv_get(b, r);
}

static orc_bm *bm_alloc(int32_t B, int32_t N) {
    orc_bm *bm = (orc_bm *)calloc(1, sizeof(orc_bm));
    bm->B = B;
    bm->N = N;
    bm->assign = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    bm->size = (int32_t *)calloc((size_t)B, sizeof(int32_t));
    bm->d_out = (int64_t *)calloc((size_t)B, sizeof(int64_t));
    bm->d_in = (int64_t *)calloc((size_t)B, sizeof(int64_t));
    bm->d_tot = (int64_t *)calloc((size_t)B, sizeof(int64_t));
    bm->row = (svec *)calloc((size_t)B, sizeof(svec));
    bm->col = (svec *)calloc((size_t)B, sizeof(svec));
    bm->overall_entropy = DBL_MAX;
    bm->num_to_merge = (int64_t)((double)B * 0.5);  /* BLOCK_REDUCTION_RATE, blockmodel.hpp:70-76 */
    return bm;
}

void orc_bm_free(orc_bm *bm) {
    if (!bm) return;
    for (int32_t b = 0; b < bm->B; ++b) {
        free(bm->row[b].idx); free(bm->row[b].val);
        free(bm->col[b].idx); free(bm->col[b].val);
    }
    free(bm->row); free(bm->col);
    if (bm->hin) {
        for (int32_t b = 0; b < bm->B; ++b) {
            free(bm->hin[b].idx); free(bm->hin[b].val);
            free(bm->hout[b].idx); free(bm->hout[b].val);
        }
        free(bm->hin); free(bm->hout);
    }
    free(bm->assign); free(bm->size); free(bm->d_out); free(bm->d_in); free(bm->d_tot);
    free(bm);
}

/* Blockmodel::initialize_edge_counts (src/blockmodel/blockmodel.cpp:193-276). */
static void bm_fill(orc_bm *bm, const orc_graph *g) {
    for (int32_t v = 0; v < g->N; ++v) {
        int32_t b = bm->assign[v];
        bm->size[b]++;
        for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) {
            int32_t nb = bm->assign[g->out_adj[e]];
            bm_add(bm, b, nb, 1);
            bm->d_out[b] += 1;
            bm->d_tot[b] += 1;
        }
        for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
            int32_t nb = bm->assign[g->in_adj[e]];
            bm->d_in[b] += 1;
            if (b != nb) bm->d_tot[b] += 1;
        }
    }
}

orc_bm *orc_bm_new(const orc_graph *g, int32_t B, const int32_t *assignment) {
    orc_bm *bm = bm_alloc(B, g->N);
    memcpy(bm->assign, assignment, sizeof(int32_t) * (size_t)g->N);
    bm_fill(bm, g);
    return bm;
}

/* Blockmodel::copy (src/blockmodel/blockmodel.cpp:139-153): num_blocks_to_merge resets to 0. */
orc_bm *orc_bm_copy(const orc_bm *src) {
    orc_bm *bm = (orc_bm *)calloc(1, sizeof(orc_bm));
    *bm = *src;
    bm->assign = (int32_t *)malloc(sizeof(int32_t) * (size_t)src->N);
    memcpy(bm->assign, src->assign, sizeof(int32_t) * (size_t)src->N);
    bm->size = (int32_t *)malloc(sizeof(int32_t) * (size_t)src->B);
    memcpy(bm->size, src->size, sizeof(int32_t) * (size_t)src->B);
    bm->d_out = (int64_t *)malloc(sizeof(int64_t) * (size_t)src->B);
    memcpy(bm->d_out, src->d_out, sizeof(int64_t) * (size_t)src->B);
    bm->d_in = (int64_t *)malloc(sizeof(int64_t) * (size_t)src->B);
    memcpy(bm->d_in, src->d_in, sizeof(int64_t) * (size_t)src->B);
    bm->d_tot = (int64_t *)malloc(sizeof(int64_t) * (size_t)src->B);
    memcpy(bm->d_tot, src->d_tot, sizeof(int64_t) * (size_t)src->B);
    bm->row = (svec *)calloc((size_t)src->B, sizeof(svec));
    bm->col = (svec *)calloc((size_t)src->B, sizeof(svec));
    for (int32_t b = 0; b < src->B; ++b) {
        sv_copy(&bm->row[b], &src->row[b]);
        sv_copy(&bm->col[b], &src->col[b]);
    }
    if (src->hin) {
        bm->hin = (svec *)calloc((size_t)src->B, sizeof(svec));
        bm->hout = (svec *)calloc((size_t)src->B, sizeof(svec));
        for (int32_t b = 0; b < src->B; ++b) {
            sv_copy(&bm->hin[b], &src->hin[b]);
            sv_copy(&bm->hout[b], &src->hout[b]);
        }
    }
    bm->num_to_merge = 0;
    return bm;
}

int32_t orc_bm_num_blocks(const orc_bm *bm) { return bm->B; }

static int64_t vtx_out(const orc_graph *g, int32_t v) { return g->out_off[v + 1] - g->out_off[v]; }
static int64_t vtx_in(const orc_graph *g, int32_t v) { return g->in_off[v + 1] - g->in_off[v]; }
/* Degree histograms of initialize_edge_counts (src/blockmodel/blockmodel.cpp:234-235): keyed by the vertex's
 * out_neighbors / in_neighbors sizes. */
static void bm_fill_hist(orc_bm *bm, const orc_graph *g) {
    if (bm->hin) {
        for (int32_t b = 0; b < bm->B; ++b) {
            free(bm->hin[b].idx); free(bm->hin[b].val);
            free(bm->hout[b].idx); free(bm->hout[b].val);
        }
        free(bm->hin); free(bm->hout);
    }
    bm->hin = (svec *)calloc((size_t)bm->B, sizeof(svec));
    bm->hout = (svec *)calloc((size_t)bm->B, sizeof(svec));
    bm->g = g;
    for (int32_t v = 0; v < g->N; ++v) {
        sv_add(&bm->hout[bm->assign[v]], (int32_t)vtx_out(g, v), 1);
        sv_add(&bm->hin[bm->assign[v]], (int32_t)vtx_in(g, v), 1);
    }
}
/* model: 0 parametric, 1 nonparametric, 2 nonparametric degree-corrected (histograms built from g) */
void orc_bm_set_model(orc_bm *bm, const orc_graph *g, int32_t model) {
    bm->np = model >= 1;
    bm->dc = model == 2;
    if (bm->dc && (!bm->hin || bm->g != g)) bm_fill_hist(bm, g);
}
static void bm_model_from(orc_bm *bm, const orc_graph *g, const orc_params *p) {
    orc_bm_set_model(bm, g, p->nonparametric ? (p->degreecorrected ? 2 : 1) : 0);
}

int64_t orc_bm_nnz(const orc_bm *bm) {
    int64_t n = 0;
    for (int32_t b = 0; b < bm->B; ++b) n += bm->row[b].len;
    return n;
}

void orc_bm_export(const orc_bm *bm, int32_t *row_ptr, int32_t *cols, int32_t *cnts, int64_t *d_out,
                   int64_t *d_in, int64_t *d_tot, int32_t *sizes, int32_t *assignment) {
    int64_t k = 0;
    for (int32_t b = 0; b < bm->B; ++b) {
        if (row_ptr) row_ptr[b] = (int32_t)k;
        for (int32_t i = 0; i < bm->row[b].len; ++i, ++k) {
            if (cols) cols[k] = bm->row[b].idx[i];
            if (cnts) cnts[k] = (int32_t)bm->row[b].val[i];
        }
    }
    if (row_ptr) row_ptr[bm->B] = (int32_t)k;
    if (d_out) memcpy(d_out, bm->d_out, sizeof(int64_t) * (size_t)bm->B);
    if (d_in) memcpy(d_in, bm->d_in, sizeof(int64_t) * (size_t)bm->B);
    if (d_tot) memcpy(d_tot, bm->d_tot, sizeof(int64_t) * (size_t)bm->B);
    if (sizes) memcpy(sizes, bm->size, sizeof(int32_t) * (size_t)bm->B);
    if (assignment) memcpy(assignment, bm->assign, sizeof(int32_t) * (size_t)bm->N);
}

/* Blockmodel::validate (src/blockmodel/blockmodel.cpp:527-632): incremental state == rebuild. */
int orc_bm_validate(const orc_bm *bm, const orc_graph *g) {
    orc_bm *fresh = orc_bm_new(g, bm->B, bm->assign);
    int ok = 1;
    for (int32_t b = 0; b < bm->B && ok; ++b) {
        if (fresh->d_out[b] != bm->d_out[b] || fresh->d_in[b] != bm->d_in[b] || fresh->d_tot[b] != bm->d_tot[b] ||
            fresh->size[b] != bm->size[b] || fresh->row[b].len != bm->row[b].len || fresh->col[b].len != bm->col[b].len)
            ok = 0;
        for (int32_t i = 0; ok && i < bm->row[b].len; ++i)
            if (fresh->row[b].idx[i] != bm->row[b].idx[i] || fresh->row[b].val[i] != bm->row[b].val[i]) ok = 0;
        for (int32_t i = 0; ok && i < bm->col[b].len; ++i)
            if (fresh->col[b].idx[i] != bm->col[b].idx[i] || fresh->col[b].val[i] != bm->col[b].val[i]) ok = 0;
    }
    orc_bm_free(fresh);
    return ok;
}

/* ------------------------------------------------------------------------- */
/* MDL (src/entropy.cpp:597-604; src/blockmodel/blockmodel.cpp:288-306)       */
/* ------------------------------------------------------------------------- */
double orc_log_posterior(const orc_bm *bm) {
    fx128 acc = 0;
    for (int32_t r = 0; r < bm->B; ++r) {
        const svec *row = &bm->row[r];
        for (int32_t i = 0; i < row->len; ++i) {
            int32_t c = row->idx[i];
            double value = (double)row->val[i];
            double d_out = (double)bm->d_out[r], d_in = (double)bm->d_in[c];
            if (d_in == 0.0 || d_out == 0.0) abort(); /* reference exits (-1000) */
            /* values * nat_log(values / (d_out * d_in)); nat_log is logf */
            double ratio = value / (d_out * d_in);
            double term = value * (double)logf((float)ratio);
            acc += fx_from(term);
        }
    }
    return fx_to(acc);
}

static double model_term(int64_t E, int32_t N, int32_t B) {
    double x = pow((double)B, 2) / (double)E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return ((double)E * h) + ((double)N * log((double)B));
}

double orc_np_mdl(const orc_bm *bm, const orc_graph *g);

double orc_mdl(const orc_bm *bm, const orc_graph *g) {
    if (bm->np) return orc_np_mdl(bm, g);  /* entropy::mdl, src/entropy.cpp:597-599 */
    return model_term(g->E, g->N, bm->B) - orc_log_posterior(bm);
}

/* ------------------------------------------------------------------------- */
/* Nonparametric description length (entropy::nonparametric,                 */
/* src/entropy.cpp:617-1225, include/entropy.hpp:101-194): directed graphs,   */
/* not degree-corrected.  fastlgamma(x) is glibc lgamma(double(x)), +inf at   */
/* 0 (include/fastlgamma.hpp:19-27, src/fastlgamma.cpp).  The cell and block  */
/* terms of sparse_entropy and the Delta entries of a move are summed in      */
/* 128-bit fixed point; every other sum keeps the reference's order.         */
/* ------------------------------------------------------------------------- */
double orc_lgamma(int64_t x) {
    if (x <= 0) return INFINITY;
    int sign;
    return lgamma_r((double)x, &sign);
}
/* fastlbinom (include/entropy.hpp:145-149) */
double orc_lbinom(int64_t n, int64_t k) {
    if (n == 0 || k == 0 || k > n) return 0;
    return ((orc_lgamma(n + 1) - orc_lgamma(k + 1)) - orc_lgamma(n - k + 1));
}
/* eterm_exact (include/entropy.hpp:103-112), directed */
static double np_eterm(int64_t w) { return -orc_lgamma(w + 1); }
/* fastlgamma taking size_t (include/fastlgamma.hpp:19-27): a negative int or long argument converts to
 * 2^64 + x, whose lgamma the reference then takes; 0 gives +inf. */
static double flg_size(int64_t x) {
    if (x < 0) { int sign; return lgamma_r((double)(uint64_t)x, &sign); }
    return orc_lgamma(x);
}
/* vterm_exact (include/entropy.hpp:114-135), directed */
static double np_vterm_m(int dc, int64_t d_out, int64_t d_in, int64_t w) {
    if (dc) return flg_size(d_out + 1) + flg_size(d_in + 1);
    if (w == 0) return 0.0;
    return (double)(d_out + d_in) * fastlog(w);
}
static double np_vterm(int64_t d_out, int64_t d_in, int64_t w) { return np_vterm_m(0, d_out, d_in, w); }

/* spence (src/spence.cpp:94-155): the Cephes dilogarithm, with the published rational coefficients. */
static const double SPENCE_A[8] = {
    4.65128586073990045278E-5, 7.31589045238094711071E-3, 1.33847639578309018650E-1, 8.79691311754530315341E-1,
    2.71149851196553469920E0,  4.25697156008121755724E0,  3.29771340985225106936E0,  1.00000000000000000126E0,
};
static const double SPENCE_B[8] = {
    6.90990488912553276999E-4, 2.54043763932544379113E-2, 2.82974860602568089943E-1, 1.41172597751831069617E0,
    3.63800533345137075418E0,  5.03278880143316990390E0,  3.54771340985225096217E0,  9.99999999999999998740E-1,
};
static double spence_poly(double x, const double *c) {
    double ans = c[0];
    for (int i = 1; i < 8; ++i) ans = ans * x + c[i];
    return ans;
}
double orc_spence(double x) {
    if (x < 0.0) return NAN;
    if (x == 1.0) return 0.0;
    if (x == 0.0) return M_PI * M_PI / 6.0;
    int flag = 0;
    double w;
    if (x > 2.0) { x = 1.0 / x; flag |= 2; }
    if (x > 1.5) { w = (1.0 / x) - 1.0; flag |= 2; }
    else if (x < 0.5) { w = -x; flag |= 1; }
    else w = x - 1.0;
    double y = -w * spence_poly(w, SPENCE_A) / spence_poly(w, SPENCE_B);
    if (flag & 1) y = (M_PI * M_PI) / 6.0 - log(x) * log1p(-x) - y;
    if (flag & 2) { double z = log(x); y = -0.5 * z * z - y; }
    return y;
}
/* get_v (src/entropy.cpp:676-686): fixed point of v = u sqrt(spence(exp(-v))) */
static double dc_get_v(double u, double epsilon) {
    double v = u, delta = 1;
    while (delta > epsilon) {
        double n_v = u * sqrt(orc_spence(exp(-v)));
        delta = fabs(n_v - v);
        v = n_v;
    }
    return v;
}
/* log_q (src/entropy.cpp:688-717), size_t arguments: log of the number of partitions of n into at most k
 * parts, approximated. */
double orc_log_q(uint64_t n, uint64_t k) {
    if (n <= 0 || k < 1) return 0;
    if (k > n) k = n;
    if ((double)k < pow((double)n, 1 / 4.)) return orc_lbinom((int64_t)(n - 1), (int64_t)(k - 1)) - flg_size((int64_t)(k + 1));
    double u = (double)k / sqrt((double)n);
    double v = dc_get_v(u, 1e-8);
    double lf = log(v) - log1p(-exp(-v) * (1 + u * u / 2)) / 2 - log(2) * 3 / 2. - log(u) - log(M_PI);
    double g = 2 * v / u - u * log1p(-exp(-v));
    return lf - log((double)n) + sqrt((double)n) * g;
}
/* get_delta_deg_dl_dist_change (src/entropy.cpp:969-1060), directed: the degree description-length change of
 * `block` when vweight vertices of in-/out-degree (vkin, vkout) leave (diff -1) or join (+1) it.  The int
 * and size_t conversions of the reference are kept. */
static double dc_change(const orc_bm *bm, int32_t block, int64_t vkin, int64_t vkout, int64_t vweight, int diff) {
    const int64_t total_r = bm->size[block], ep_r = bm->d_out[block], em_r = bm->d_in[block];
    const int tkin = (int)vkin, tkout = (int)vkout, n = (int)vweight;
    const int nd_in = (int)sv_get(&bm->hin[block], (int32_t)vkin), nd_out = (int)sv_get(&bm->hout[block], (int32_t)vkout);
#define DC_SK(delta) ((0.0 - flg_size((int64_t)(nd_in + (delta) + 1))) - flg_size((int64_t)(nd_out + (delta) + 1)))
#define DC_SE(delta, kin, kout) ((0.0 + orc_log_q((uint64_t)(em_r + (kin)), (uint64_t)(total_r + (delta)))) + \
                                 orc_log_q((uint64_t)(ep_r + (kout)), (uint64_t)(total_r + (delta))))
#define DC_SR(delta) (2 * flg_size(total_r + (delta) + 1))
    double S_b = 0, S_a = 0;
    S_b += DC_SK(0);
    S_a += DC_SK((int)(diff * vweight));
    S_b += DC_SE(0, 0, 0);
    S_a += DC_SE(diff * n, diff * tkin, diff * tkout);
    S_b += DC_SR(0);
    S_a += DC_SR(diff * n);
#undef DC_SK
#undef DC_SE
#undef DC_SR
    return S_a - S_b;
}
/* get_delta_deg_dl (src/entropy.cpp:1063-1096 vertex, :1166-1181 merge) */
static double dc_delta_deg_dl(const orc_bm *bm, int32_t r, int32_t s, int64_t vkin, int64_t vkout, int64_t weight) {
    if (r == s) return 0.0;
    double dS = 0;
    dS += dc_change(bm, r, vkin, vkout, weight, -1);
    dS += dc_change(bm, s, vkin, vkout, weight, +1);
    return dS;
}
/* get_deg_dl_dist (src/entropy.cpp:719-742), directed: block by block, in-histogram then out-histogram
 * entries in ascending degree. */
double orc_np_deg_dl(const orc_bm *bm) {
    if (!bm->dc) return 0.0;
    double S = 0;
    for (int32_t b = 0; b < bm->B; ++b) {
        S += orc_log_q((uint64_t)bm->d_out[b], (uint64_t)bm->size[b]);
        S += orc_log_q((uint64_t)bm->d_in[b], (uint64_t)bm->size[b]);
        uint64_t total = 0;
        for (int32_t i = 0; i < bm->hin[b].len; ++i) S -= flg_size(bm->hin[b].val[i] + 1);
        for (int32_t i = 0; i < bm->hout[b].len; ++i) {
            S -= flg_size(bm->hout[b].val[i] + 1);
            total += (uint64_t)bm->hout[b].val[i];
        }
        S += 2 * flg_size((int64_t)(total + 1));
    }
    return S;
}
/* get_edges_dl (src/entropy.cpp:733-738) */
double orc_np_edges_dl(int64_t B, int64_t E) { return orc_lbinom(B * B + E - 1, E); }
static int64_t np_nonempty(const orc_bm *bm) {
    int64_t n = 0;
    for (int32_t b = 0; b < bm->B; ++b) n += bm->size[b] > 0;
    return n;
}
/* get_partition_dl (src/entropy.cpp:657-669) */
double orc_np_partition_dl(const orc_bm *bm) {
    int64_t N = bm->N;
    double S = 0;
    S += orc_lbinom(N - 1, np_nonempty(bm) - 1);
    S += orc_lgamma(N + 1);
    for (int32_t b = 0; b < bm->B; ++b) S -= orc_lgamma(bm->size[b] + 1);
    S += fastlog(N);
    return S;
}
/* sparse_entropy (src/entropy.cpp:623-655) */
double orc_np_sparse_entropy(const orc_bm *bm) {
    fx128 acc = 0;
    for (int32_t r = 0; r < bm->B; ++r)
        for (int32_t i = 0; i < bm->row[r].len; ++i) acc += fx_from(np_eterm(bm->row[r].val[i]));
    for (int32_t b = 0; b < bm->B; ++b) acc += fx_from(np_vterm_m(bm->dc, bm->d_out[b], bm->d_in[b], bm->size[b]));
    if (bm->dc)  /* get_deg_entropy (src/entropy.cpp:619-623) for every vertex */
        for (int32_t v = 0; v < bm->N; ++v)
            acc += fx_from(-flg_size(vtx_in(bm->g, v) + 1) - flg_size(vtx_out(bm->g, v) + 1));
    return fx_to(acc);
}
/* nonparametric::mdl (src/entropy.cpp:740-767) */
double orc_np_mdl(const orc_bm *bm, const orc_graph *g) {
    double S = orc_np_sparse_entropy(bm), S_dl = 0;
    S_dl += orc_np_partition_dl(bm);
    S_dl += orc_np_deg_dl(bm);
    S_dl += orc_np_edges_dl(np_nonempty(bm), g->E);
    return S + S_dl * 1.0;  /* BETA_DL = 1 (include/entropy.hpp:14) */
}

/* entropy::null_mdl_v1 (src/entropy.cpp:571-585) */
double orc_null_mdl_v1(int64_t E) {
    double log_posterior_p = (double)E * log(1.0 / (double)E);
    double x = 1.0 / (double)E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return ((double)E * h) - log_posterior_p;
}

/* ------------------------------------------------------------------------- */
/* Delta (include/blockmodel/sparse/delta.hpp:22-149): four maps; cells that  */
/* net to zero stay in their map.                                            */
/* ------------------------------------------------------------------------- */
typedef struct { int32_t *key; int64_t *val; uint8_t *used; int32_t cap, count; } imap;

static void im_init(imap *m, int32_t expected) {
    int32_t cap = 16;
    while (cap < 2 * expected + 2) cap <<= 1;
    m->cap = cap;
    m->count = 0;
    m->key = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    m->val = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
    m->used = (uint8_t *)calloc((size_t)cap, 1);
}
static void im_free(imap *m) { free(m->key); free(m->val); free(m->used); }
static inline uint32_t im_hash(int32_t k) { return (uint32_t)k * 0x9E3779B1u; }
static void im_add(imap *m, int32_t k, int64_t d) {
    if (2 * (m->count + 1) > m->cap) {
        imap n;
        im_init(&n, m->cap);
        for (int32_t i = 0; i < m->cap; ++i)
            if (m->used[i]) im_add(&n, m->key[i], m->val[i]);
        im_free(m);
        *m = n;
    }
    uint32_t mask = (uint32_t)m->cap - 1, h = im_hash(k) & mask;
    while (m->used[h]) {
        if (m->key[h] == k) { m->val[h] += d; return; }
        h = (h + 1) & mask;
    }
    m->used[h] = 1;
    m->key[h] = k;
    m->val[h] = d;
    m->count++;
}
static int64_t im_get(const imap *m, int32_t k) {
    uint32_t mask = (uint32_t)m->cap - 1, h = im_hash(k) & mask;
    while (m->used[h]) {
        if (m->key[h] == k) return m->val[h];
        h = (h + 1) & mask;
    }
    return 0;
}

typedef struct {
    imap cur_row, prop_row, cur_col, prop_col;
    int32_t r, s;
    int64_t self_edge_weight;
} delta_t;

static void delta_init(delta_t *d, int32_t r, int32_t s, int32_t expected) {
    d->r = r; d->s = s; d->self_edge_weight = 0;
    im_init(&d->cur_row, expected); im_init(&d->prop_row, expected);
    im_init(&d->cur_col, expected); im_init(&d->prop_col, expected);
}
static void delta_free(delta_t *d) {
    im_free(&d->cur_row); im_free(&d->prop_row); im_free(&d->cur_col); im_free(&d->prop_col);
}
static void delta_add(delta_t *d, int32_t row, int32_t col, int64_t v) {
    if (row == d->r) im_add(&d->cur_row, col, v);
    else if (row == d->s) im_add(&d->prop_row, col, v);
    else if (col == d->r) im_add(&d->cur_col, row, v);
    else if (col == d->s) im_add(&d->prop_col, row, v);
    else abort(); /* std::logic_error in the reference */
}
static int64_t delta_get(const delta_t *d, int32_t row, int32_t col) {
    if (row == d->r) return im_get(&d->cur_row, col);
    if (row == d->s) return im_get(&d->prop_row, col);
    if (col == d->r) return im_get(&d->cur_col, row);
    if (col == d->s) return im_get(&d->prop_col, row);
    abort();
}

/* ------------------------------------------------------------------------- */
/* Finetune: vertex proposals (src/common.cpp:138-198, src/finetune.cpp)      */
/* ------------------------------------------------------------------------- */
typedef struct { int32_t s; int64_t k_out, k_in, k; int64_t vkin, vkout; } proposal_t;  /* vk*: graph degrees */

/* Vertex-mode propose_new_block with the canonical multinomial. */
static proposal_t propose_vertex(const orc_bm *bm, const orc_graph *g, int32_t v, const uint32_t w[4]) {
    proposal_t p;
    int32_t r = bm->assign[v];
    int64_t ko = g->out_off[v + 1] - g->out_off[v];
    int64_t ki_all = g->in_off[v + 1] - g->in_off[v], self_in = 0;
    for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) self_in += (g->in_adj[e] == v);
    p.k_out = ko;
    p.k_in = ki_all - self_in;
    p.vkout = ko;
    p.vkin = ki_all;
    p.k = p.k_out + p.k_in;
    if (p.k == 0) {
        p.s = random_block(r, bm->B, w[1]);
        return p;
    }
    uint32_t idx = orc_uniform_below(w[0], (uint32_t)p.k);
    int32_t u;
    if ((int64_t)idx < ko) {
        u = g->out_adj[g->out_off[v] + idx];
    } else {
        int64_t want = (int64_t)idx - ko, seen = 0;
        u = -1;
        for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
            if (g->in_adj[e] == v) continue;
            if (seen == want) { u = g->in_adj[e]; break; }
            seen++;
        }
    }
    int32_t t = bm->assign[u];
    /* neighbors_weights(t) (dict_transpose_matrix.cpp:176-188): row(t) then col(t) without row t */
    const svec *row = &bm->row[t], *col = &bm->col[t];
    int64_t total = 0;
    for (int32_t i = 0; i < row->len; ++i) total += row->val[i];
    for (int32_t i = 0; i < col->len; ++i) if (col->idx[i] != t) total += col->val[i];
    uint32_t target = orc_uniform_below(w[1], (uint32_t)total);
    int64_t prefix = 0;
    p.s = -1;
    for (int32_t i = 0; i < row->len && p.s < 0; ++i) {
        prefix += row->val[i];
        if ((int64_t)target < prefix) p.s = row->idx[i];
    }
    for (int32_t i = 0; i < col->len && p.s < 0; ++i) {
        if (col->idx[i] == t) continue;
        prefix += col->val[i];
        if ((int64_t)target < prefix) p.s = col->idx[i];
    }
    return p;
}

/* finetune::blockmodel_delta (src/finetune.cpp:201-235) with in-edges excluding self. */
static void vertex_delta(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t r, int32_t s, delta_t *d) {
    int64_t k = (g->out_off[v + 1] - g->out_off[v]) + (g->in_off[v + 1] - g->in_off[v]);
    delta_init(d, r, s, (int32_t)(k < 1000000 ? k : 1000000));
    for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) {
        int32_t u = g->out_adj[e], ob = bm->assign[u];
        if (u == v) {
            delta_add(d, s, s, 1);
            d->self_edge_weight = 1;
        } else {
            delta_add(d, s, ob, 1);
        }
        delta_add(d, r, ob, -1);
    }
    for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
        int32_t u = g->in_adj[e];
        if (u == v) continue; /* edge_weights(in, ignore_self=true) */
        int32_t ib = bm->assign[u];
        delta_add(d, ib, s, 1);
        delta_add(d, ib, r, -1);
    }
}

/* Sum of |cell entropy| terms a per-cell evaluation touches (the scale of its rounding error). */
static __thread double g_abs_terms;

typedef struct { const orc_bm *bm; int32_t r, s; int64_t shift_in, shift_out; } degfn;
static inline int64_t new_d_in(const degfn *f, int32_t x) {
    int64_t v = f->bm->d_in[x];
    if (x == f->r) v -= f->shift_in; else if (x == f->s) v += f->shift_in;
    return v;
}
static inline int64_t new_d_out(const degfn *f, int32_t x) {
    int64_t v = f->bm->d_out[x];
    if (x == f->r) v -= f->shift_out; else if (x == f->s) v += f->shift_out;
    return v;
}

static void imap_entries_delta(const orc_bm *bm, const delta_t *d, const imap *m, int which, const degfn *f,
                               int skip_new_in_r, fx128 *acc) {
    for (int32_t i = 0; i < m->cap; ++i) {
        if (!m->used[i]) continue;
        int32_t row, col;
        if (which == 0) { row = d->r; col = m->key[i]; }
        else if (which == 1) { row = d->s; col = m->key[i]; }
        else if (which == 2) { row = m->key[i]; col = d->r; }
        else { row = m->key[i]; col = d->s; }
        int64_t change = m->val[i];
        int64_t value = bm_get(bm, row, col);
        double t_old = cell_entropy(value, bm->d_in[col], bm->d_out[row]);
        *acc += fx_from(t_old);
        g_abs_terms += fabs(t_old);
        if (skip_new_in_r && (row == d->r || col == d->r)) continue;
        double t_new = cell_entropy(value + change, new_d_in(f, col), new_d_out(f, row));
        *acc -= fx_from(t_new);
        g_abs_terms += fabs(t_new);
    }
}

/* Replay-contract delta-MDL of a vertex move: the Delta entries term by term, the four cells at
 * rows/columns {r, s} term by term, and the unchanged cells of rows r, s and columns r, s in bulk.  An
 * unchanged cell (row, c) with c outside {r, s} changes only through d_out[row], so its old-minus-new
 * entropy is M * (fastlog(d_out'[row]) - fastlog(d_out[row])); summing M over those cells gives the row
 * total minus the Delta and special cells.  Equal to entropy::delta_mdl (src/entropy.cpp:298-403) up to
 * rounding: the per-cell form is kept below as delta_mdl_from and the two are compared in the tests. */
static double delta_mdl_factored(const orc_bm *bm, const delta_t *d, const proposal_t *p) {
    degfn f = { bm, d->r, d->s, p->k_in + d->self_edge_weight, p->k_out };
    const int32_t r = d->r, s = d->s;
    fx128 acc = 0;
    imap_entries_delta(bm, d, &d->cur_row, 0, &f, 0, &acc);
    imap_entries_delta(bm, d, &d->prop_row, 1, &f, 0, &acc);
    imap_entries_delta(bm, d, &d->cur_col, 2, &f, 0, &acc);
    imap_entries_delta(bm, d, &d->prop_col, 3, &f, 0, &acc);
    const int32_t sp[4][2] = { { r, r }, { r, s }, { s, r }, { s, s } };
    for (int k = 0; k < 4; ++k) {
        int32_t row = sp[k][0], col = sp[k][1];
        int64_t m = bm_get(bm, row, col);
        if (m == 0 || delta_get(d, row, col) != 0) continue;
        acc += fx_from(cell_entropy(m, bm->d_in[col], bm->d_out[row]));
        acc -= fx_from(cell_entropy(m, new_d_in(&f, col), new_d_out(&f, row)));
    }
    const int64_t mrr = bm_get(bm, r, r), mrs = bm_get(bm, r, s), msr = bm_get(bm, s, r), mss = bm_get(bm, s, s);
    int64_t S_rrow = bm->d_out[r] - mrr - mrs, S_srow = bm->d_out[s] - mss - msr;
    int64_t S_rcol = bm->d_in[r] - mrr - msr, S_scol = bm->d_in[s] - mss - mrs;
    for (int32_t i = 0; i < d->cur_row.cap; ++i)
        if (d->cur_row.used[i] && d->cur_row.key[i] != r && d->cur_row.key[i] != s) S_rrow -= bm_get(bm, r, d->cur_row.key[i]);
    for (int32_t i = 0; i < d->prop_row.cap; ++i)
        if (d->prop_row.used[i] && d->prop_row.key[i] != r && d->prop_row.key[i] != s) S_srow -= bm_get(bm, s, d->prop_row.key[i]);
    for (int32_t i = 0; i < d->cur_col.cap; ++i)
        if (d->cur_col.used[i]) S_rcol -= bm_get(bm, d->cur_col.key[i], r);
    for (int32_t i = 0; i < d->prop_col.cap; ++i)
        if (d->prop_col.used[i]) S_scol -= bm_get(bm, d->prop_col.key[i], s);
    acc += fx_from((double)S_rrow * (fastlog(new_d_out(&f, r)) - fastlog(bm->d_out[r])));
    acc += fx_from((double)S_srow * (fastlog(new_d_out(&f, s)) - fastlog(bm->d_out[s])));
    acc += fx_from((double)S_rcol * (fastlog(new_d_in(&f, r)) - fastlog(bm->d_in[r])));
    acc += fx_from((double)S_scol * (fastlog(new_d_in(&f, s)) - fastlog(bm->d_in[s])));
    return fx_to(acc);
}

/* entropy::delta_mdl(const Blockmodel&, const Delta&, const ProposalAndEdgeCounts&) (src/entropy.cpp:298-403) */
static double delta_mdl_from(const orc_bm *bm, const delta_t *d, const proposal_t *p) {
    degfn f = { bm, d->r, d->s, p->k_in + d->self_edge_weight, p->k_out };
    fx128 acc = 0;
    imap_entries_delta(bm, d, &d->cur_row, 0, &f, 0, &acc);
    imap_entries_delta(bm, d, &d->prop_row, 1, &f, 0, &acc);
    imap_entries_delta(bm, d, &d->cur_col, 2, &f, 0, &acc);
    imap_entries_delta(bm, d, &d->prop_col, 3, &f, 0, &acc);
    int32_t rs[2] = { d->r, d->s };
    for (int k = 0; k < 2; ++k) {
        int32_t row = rs[k];
        const svec *sv = &bm->row[row];
        for (int32_t i = 0; i < sv->len; ++i) {
            int32_t col = sv->idx[i];
            int64_t value = sv->val[i];
            if (delta_get(d, row, col) != 0) continue;
            double a = cell_entropy(value, bm->d_in[col], bm->d_out[row]);
            double b = cell_entropy(value, new_d_in(&f, col), new_d_out(&f, row));
            acc += fx_from(a);
            acc -= fx_from(b);
            g_abs_terms += fabs(a) + fabs(b);
        }
    }
    for (int k = 0; k < 2; ++k) {
        int32_t col = rs[k];
        const svec *sv = &bm->col[col];
        for (int32_t i = 0; i < sv->len; ++i) {
            int32_t row = sv->idx[i];
            int64_t value = sv->val[i];
            if (delta_get(d, row, col) != 0 || row == d->r || row == d->s) continue;
            double a = cell_entropy(value, bm->d_in[col], bm->d_out[row]);
            double b = cell_entropy(value, new_d_in(&f, col), new_d_out(&f, row));
            acc += fx_from(a);
            acc -= fx_from(b);
            g_abs_terms += fabs(a) + fabs(b);
        }
    }
    return fx_to(acc);
}

/* entropy::hastings_correction (src/entropy.cpp:497-561). */
static double hastings_from(const orc_bm *bm, const orc_graph *g, int32_t v, const delta_t *d, const proposal_t *p) {
    if (p->k == 0) return 1.0;
    int32_t r = d->r, s = d->s;
    /* block_counts in ascending block order */
    int64_t k = (g->out_off[v + 1] - g->out_off[v]) + (g->in_off[v + 1] - g->in_off[v]);
    int32_t *blocks = (int32_t *)malloc(sizeof(int32_t) * (size_t)(k + 1));
    int64_t nb = 0;
    for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) blocks[nb++] = bm->assign[g->out_adj[e]];
    for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e)
        if (g->in_adj[e] != v) blocks[nb++] = bm->assign[g->in_adj[e]];
    /* insertion-free counting by sorting */
    for (int64_t i = 1; i < nb; ++i) {
        int32_t x = blocks[i];
        int64_t j = i - 1;
        while (j >= 0 && blocks[j] > x) { blocks[j + 1] = blocks[j]; j--; }
        blocks[j + 1] = x;
    }
    fx128 fwd = 0, bwd = 0;
    int64_t B = bm->B;
    for (int64_t i = 0; i < nb;) {
        int32_t x = blocks[i];
        int64_t cnt = 0;
        while (i < nb && blocks[i] == x) { cnt++; i++; }
        double counts = (double)cnt;
        double proposal_weights = (double)(bm_get(bm, s, x) + bm_get(bm, x, s)) + 1.0;
        double block_degrees = (double)(bm->d_tot[x] + B);
        double block_weights = (double)(bm_get(bm, r, x) + delta_get(d, r, x) + bm_get(bm, x, r) + delta_get(d, x, r)) + 1.0;
        int64_t new_block_degree = bm->d_tot[x];
        if (x == r) {
            int64_t self = bm_get(bm, r, r) + delta_get(d, r, r);
            new_block_degree = (bm->d_out[r] - p->k_out) + (bm->d_in[r] - p->k_in) - self;
        } else if (x == s) {
            int64_t self = bm_get(bm, s, s) + delta_get(d, s, s);
            new_block_degree = (bm->d_out[s] + p->k_out) + (bm->d_in[s] + p->k_in) - self;
        }
        double proposal_degrees = (double)(new_block_degree + B);
        fwd += fx_from(counts * proposal_weights / block_degrees);
        bwd += fx_from(counts * block_weights / proposal_degrees);
    }
    free(blocks);
    return fx_to(bwd) / fx_to(fwd);
}

/* finetune::accept (src/finetune.cpp:22-32). */
static int accept_move(double dS, double H, int greedy, const uint32_t w[4]) {
    if (greedy) return dS < 0.0;
    double u = orc_unit_double(w[2], w[3]);
    double a = exp(-3.0 * dS) * H;
    a = (a >= 1.0) ? 1.0 : a;
    return u <= a;
}

static double np_vertex_dS(const orc_bm *bm, const delta_t *d, const proposal_t *p);

double orc_vertex_delta_mdl(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t s) {
    uint32_t w[4] = { 0, 0, 0, 0 };
    proposal_t p = propose_vertex(bm, g, v, w);
    p.s = s;
    delta_t d;
    vertex_delta(bm, g, v, bm->assign[v], s, &d);
    double dS = bm->np ? np_vertex_dS(bm, &d, &p) : delta_mdl_factored(bm, &d, &p);
    delta_free(&d);
    return dS;
}

/* entropy::delta_mdl cell by cell, as the reference evaluates it; *abs_terms = sum of |cell terms|. */
double orc_vertex_delta_mdl_cells(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t s, double *abs_terms) {
    uint32_t w[4] = { 0, 0, 0, 0 };
    proposal_t p = propose_vertex(bm, g, v, w);
    p.s = s;
    delta_t d;
    vertex_delta(bm, g, v, bm->assign[v], s, &d);
    g_abs_terms = 0.0;
    double dS = delta_mdl_from(bm, &d, &p);
    if (abs_terms) *abs_terms = g_abs_terms;
    delta_free(&d);
    return dS;
}

double orc_vertex_hastings(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t s) {
    uint32_t w[4] = { 0, 0, 0, 0 };
    proposal_t p = propose_vertex(bm, g, v, w);
    p.s = s;
    delta_t d;
    vertex_delta(bm, g, v, bm->assign[v], s, &d);
    double H = hastings_from(bm, g, v, &d, &p);
    delta_free(&d);
    return H;
}

/* propose_gibbs_move_v3 / propose_move + eval (src/finetune.cpp:447-462,842-901), against a fixed bm. */
/* entries_dS (src/entropy.cpp:771-798): eterm(M + change) - eterm(M) over the Delta entries. */
static void np_entries(const orc_bm *bm, const delta_t *d, fx128 *acc) {
    const imap *maps[4] = { &d->cur_row, &d->prop_row, &d->cur_col, &d->prop_col };
    for (int which = 0; which < 4; ++which) {
        const imap *mp = maps[which];
        for (int32_t i = 0; i < mp->cap; ++i) {
            if (!mp->used[i]) continue;
            int32_t row, col;
            if (which == 0) { row = d->r; col = mp->key[i]; }
            else if (which == 1) { row = d->s; col = mp->key[i]; }
            else if (which == 2) { row = mp->key[i]; col = d->r; }
            else { row = mp->key[i]; col = d->s; }
            int64_t value = bm_get(bm, row, col);
            *acc += fx_from(np_eterm(value + mp->val[i]) - np_eterm(value));
        }
    }
}
/* get_delta_partition_dl + get_delta_deg_dl (0) + get_delta_edges_dl (src/entropy.cpp:865-1132) for
 * moving `weight` vertices from r (size_r) to s (size_s). */
static double np_delta_dl(int64_t N, int64_t E, int64_t nonempty, int64_t size_r, int64_t size_s, int64_t weight,
                          double ddeg) {
    double S_b = 0, S_a = 0;
    S_b += -orc_lgamma(size_r + 1);
    S_a += -orc_lgamma(size_r - weight + 1);
    S_b += -orc_lgamma(size_s + 1);
    S_a += -orc_lgamma(size_s + weight + 1);
    int dB = 0;
    if (size_r == weight) dB--;
    if (size_s == 0) dB++;
    if (dB != 0) {
        S_b += orc_lbinom(N - 1, nonempty - 1);
        S_a += orc_lbinom(N - 1, nonempty + dB - 1);
    }
    double dpart = 0;
    dpart += S_a - S_b;
    double eb = 0, ea = 0;
    if (dB != 0) {
        eb += orc_np_edges_dl(nonempty, E);
        ea += orc_np_edges_dl(nonempty + dB, E);
    }
    double dedges = ea - eb;
    double dS_dl = 0;
    dS_dl += dpart;
    dS_dl += ddeg;  /* get_delta_deg_dl: 0 without degree correction */
    dS_dl += dedges;
    return dS_dl;
}
/* virtual_move_sparse (src/entropy.cpp:802-862) + the description-length deltas: nonparametric::delta_mdl
 * (src/entropy.cpp:1135-1164) with weight 1, block_merge_delta_mdl (:1183-1206) with weight = size(r). */
static double np_move_dS(const orc_bm *bm, const delta_t *d, int64_t k_out, int64_t k_in, int64_t weight,
                         int64_t vkin, int64_t vkout) {
    const int32_t r = d->r, s = d->s;
    if (r == s) return 0.0;
    int64_t E = 0;  /* the graph's edge count: every edge is one unit of some block's out-degree */
    for (int32_t b = 0; b < bm->B; ++b) E += bm->d_out[b];
    fx128 acc = 0;
    np_entries(bm, d, &acc);
    double dS = fx_to(acc);
    dS += np_vterm_m(bm->dc, bm->d_out[r] - k_out, bm->d_in[r] - k_in, bm->size[r] - weight);
    dS -= np_vterm_m(bm->dc, bm->d_out[r], bm->d_in[r], bm->size[r]);
    dS += np_vterm_m(bm->dc, bm->d_out[s] + k_out, bm->d_in[s] + k_in, bm->size[s] + weight);
    dS -= np_vterm_m(bm->dc, bm->d_out[s], bm->d_in[s], bm->size[s]);
    const double ddeg = bm->dc ? dc_delta_deg_dl(bm, r, s, vkin, vkout, weight) : 0.0;
    double dS_dl = np_delta_dl(bm->N, E, np_nonempty(bm), bm->size[r], bm->size[s], weight, ddeg);
    return dS + 1.0 * dS_dl;
}
static double np_vertex_dS(const orc_bm *bm, const delta_t *d, const proposal_t *p) {
    return np_move_dS(bm, d, p->k_out, p->k_in, 1, p->vkin, p->vkout);
}
/* the merge's degree terms take block r's degrees as the "vertex" degrees (src/entropy.cpp:1170-1173) */
static double np_merge_dS(const orc_bm *bm, const delta_t *d) {
    if (d->r == d->s || bm->size[d->r] == 0 || bm->size[d->s] == 0) return 0.0;
    return np_move_dS(bm, d, bm->d_out[d->r], bm->d_in[d->r], bm->size[d->r], bm->d_in[d->r], bm->d_out[d->r]);
}

static int evaluate_vertex(const orc_bm *bm, const orc_graph *g, uint64_t seed, uint32_t epoch, int32_t v,
                           int greedy, int32_t *s_out, double *dS_out, delta_t *keep, proposal_t *p_out) {
    uint32_t w[4];
    orc_philox(seed, (uint32_t)v, epoch, ORC_TAG_FT, 0, w);
    int32_t r = bm->assign[v];
    proposal_t p = propose_vertex(bm, g, v, w);
    *s_out = p.s;
    *dS_out = 0.0;
    if (p_out) *p_out = p;
    if (p.s == r) return 0;
    delta_t d;
    vertex_delta(bm, g, v, r, p.s, &d);
    /* hastings_correction is 1 for greedy or nonparametric runs (src/entropy.cpp:499) */
    double H = (greedy || bm->np) ? 1.0 : hastings_from(bm, g, v, &d, &p);
    double dS = bm->np ? np_vertex_dS(bm, &d, &p) : delta_mdl_factored(bm, &d, &p);
    *dS_out = dS;
    int acc = accept_move(dS, H, greedy, w);
    if (keep && acc) *keep = d; else delta_free(&d);
    return acc;
}

int orc_ft_propose(const orc_bm *bm, const orc_graph *g, uint64_t seed, uint32_t epoch, int32_t v, int32_t greedy,
                   int32_t *s_out, double *dS_out, int32_t *accept_out) {
    *accept_out = evaluate_vertex(bm, g, seed, epoch, v, greedy, s_out, dS_out, NULL, NULL);
    return 0;
}

/* Blockmodel::move_vertex(const VertexMove_v3&) (src/blockmodel/blockmodel.cpp:437-494). */
static int commit_v3(orc_bm *bm, const orc_graph *g, int32_t v, int32_t s) {
    int32_t r = bm->assign[v];
    if (bm->size[r] == 1) return 0;
    for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) {
        int32_t u = g->out_adj[e], ob = bm->assign[u];
        bm_add(bm, r, ob, -1);
        bm->d_out[r]--;
        bm->d_out[s]++;
        if (u == v) {
            bm_add(bm, s, s, 1);
            bm->d_in[r]--;
            bm->d_in[s]++;
        } else {
            bm_add(bm, s, ob, 1);
        }
    }
    for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
        int32_t u = g->in_adj[e];
        if (u == v) continue;
        int32_t ib = bm->assign[u];
        bm_add(bm, ib, r, -1);
        bm->d_in[r]--;
        bm_add(bm, ib, s, 1);
        bm->d_in[s]++;
    }
    bm->d_tot[r] = bm->d_in[r] + bm->d_out[r] - bm_get(bm, r, r);
    bm->d_tot[s] = bm->d_in[s] + bm->d_out[s] - bm_get(bm, s, s);
    bm->assign[v] = s;
    bm->size[r]--;
    bm->size[s]++;
    if (bm->dc) {  /* histograms keyed by the vertex's graph degrees (blockmodel.cpp:489-492) */
        sv_add(&bm->hout[r], (int32_t)vtx_out(g, v), -1);
        sv_add(&bm->hin[r], (int32_t)vtx_in(g, v), -1);
        sv_add(&bm->hout[s], (int32_t)vtx_out(g, v), 1);
        sv_add(&bm->hin[s], (int32_t)vtx_in(g, v), 1);
    }
    return 1;
}

/* Blockmodel::move_vertex(Vertex, const Delta&, ProposalAndEdgeCounts&) (src/blockmodel/blockmodel.cpp:408-435). */
static void commit_delta(orc_bm *bm, int32_t v, const delta_t *d, const proposal_t *p) {
    int32_t r = d->r, s = d->s;
    bm->assign[v] = s;
    const imap *maps[4] = { &d->cur_row, &d->prop_row, &d->cur_col, &d->prop_col };
    for (int which = 0; which < 4; ++which) {
        const imap *m = maps[which];
        for (int32_t i = 0; i < m->cap; ++i) {
            if (!m->used[i] || m->val[i] == 0) continue;
            int32_t row, col;
            if (which == 0) { row = r; col = m->key[i]; }
            else if (which == 1) { row = s; col = m->key[i]; }
            else if (which == 2) { row = m->key[i]; col = r; }
            else { row = m->key[i]; col = s; }
            bm_add(bm, row, col, m->val[i]);
        }
    }
    bm->d_out[r] -= p->k_out;
    bm->d_out[s] += p->k_out;
    bm->d_in[r] -= (p->k_in + d->self_edge_weight);
    bm->d_in[s] += (p->k_in + d->self_edge_weight);
    bm->d_tot[r] = bm->d_out[r] + bm->d_in[r] - bm_get(bm, r, r);
    bm->d_tot[s] = bm->d_out[s] + bm->d_in[s] - bm_get(bm, s, s);
    bm->size[r]--;
    bm->size[s]++;
    if (bm->dc) {  /* keyed by the proposal's edge counts, in-edges without self-loops (blockmodel.cpp:431-434) */
        sv_add(&bm->hout[r], (int32_t)p->k_out, -1);
        sv_add(&bm->hin[r], (int32_t)p->k_in, -1);
        sv_add(&bm->hout[s], (int32_t)p->k_out, 1);
        sv_add(&bm->hin[s], (int32_t)p->k_in, 1);
    }
}

/* Keyed shuffle: ascending Philox keys, ties by vertex id. */
typedef struct { uint32_t key; int32_t v; } keyed_t;
static int keyed_cmp(const void *a, const void *b) {
    const keyed_t *x = (const keyed_t *)a, *y = (const keyed_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->v > y->v) - (x->v < y->v);
}
static void keyed_shuffle(int32_t *items, int64_t n, uint64_t seed, uint32_t epoch, uint32_t tag) {
    keyed_t *k = (keyed_t *)malloc(sizeof(keyed_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        uint32_t w[4];
        orc_philox(seed, (uint32_t)items[i], epoch, tag, 0, w);
        k[i].key = w[0];
        k[i].v = items[i];
    }
    qsort(k, (size_t)n, sizeof(keyed_t), keyed_cmp);
    for (int64_t i = 0; i < n; ++i) items[i] = k[i].v;
    free(k);
}

/* Proposal phase + sequential commit in vertex-id order over one batch of `order[start, end)`. */
static void gibbs_batch(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch, const int32_t *order,
                        int64_t start, int64_t end, int32_t *moves_s, int32_t *proposed_s, orc_stats *st) {
    int32_t N = g->N;
    for (int32_t v = 0; v < N; ++v) moves_s[v] = -1;
    int64_t nprop = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nprop) num_threads(p->threads > 0 ? p->threads : 1)
    for (int64_t index = start; index < end; ++index) {
        int32_t v = order ? order[index] : (int32_t)index;
        int32_t s;
        double dS;
        int acc = evaluate_vertex(bm, g, p->seed, epoch, v, p->greedy, &s, &dS, NULL, NULL);
        nprop++;
        if (acc) moves_s[v] = s;
    }
    if (proposed_s)
        for (int32_t v = 0; v < N; ++v) if (moves_s[v] >= 0) proposed_s[v] = moves_s[v];
    int64_t moved = 0, refused = 0;
    for (int32_t v = 0; v < N; ++v) {
        if (moves_s[v] < 0) continue;
        if (commit_v3(bm, g, v, moves_s[v])) moved++; else refused++;
    }
    if (st) { st->proposals += nprop; st->moves += moved; st->refused += refused; }
}

/* One sweep of finetune::asynchronous_gibbs (src/finetune.cpp:50-86). */
int orc_async_gibbs_sweep(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch, int32_t *proposed_s,
                          orc_stats *st) {
    bm_model_from(bm, g, p);
    int32_t N = g->N;
    int32_t *order = NULL;
    if (!p->ordered) {
        order = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
        for (int32_t i = 0; i < N; ++i) order[i] = i;
        keyed_shuffle(order, N, p->seed, epoch, ORC_TAG_SHUFFLE);
    }
    if (proposed_s) for (int32_t v = 0; v < N; ++v) proposed_s[v] = -1;
    int32_t *moves_s = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int64_t batch_size = (int64_t)ceil((double)N / (double)p->batches);
    for (int64_t batch = 0; batch < N / batch_size; ++batch) {
        int64_t start = batch * batch_size;
        int64_t end = (batch + 1) * batch_size < N ? (batch + 1) * batch_size : N;
        gibbs_batch(bm, g, p, epoch, order, start, end, moves_s, proposed_s, st);
    }
    free(moves_s);
    free(order);
    if (st) st->sweeps++;
    return 0;
}

/* One sweep of sequential Metropolis-Hastings (src/finetune.cpp:793-803,822-869). */
int orc_mh_sweep(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch, int32_t ordered,
                 double *sum_dS, orc_stats *st) {
    bm_model_from(bm, g, p);
    int32_t N = g->N;
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int32_t i = 0; i < N; ++i) order[i] = i;
    if (!ordered) keyed_shuffle(order, N, p->seed, epoch, ORC_TAG_SHUFFLE);
    double total = 0;  /* accepted delta-MDL, added in visit order (src/finetune.cpp:795-800) */
    for (int32_t index = 0; index < N; ++index) {
        int32_t v = order[index];
        int32_t r = bm->assign[v];
        if (bm->size[r] == 1) continue;
        int32_t s;
        double dS;
        delta_t d;
        proposal_t prop;
        int acc = evaluate_vertex(bm, g, p->seed, epoch, v, p->greedy, &s, &dS, &d, &prop);
        if (st) st->proposals++;
        if (acc) {
            commit_delta(bm, v, &d, &prop);
            delta_free(&d);
            total += dS;
            if (st) st->moves++;
        }
    }
    free(order);
    if (sum_dS) *sum_dS = total;
    if (st) st->sweeps++;
    return 0;
}

/* finetune::early_stop, 4-argument form (src/finetune.cpp:265-283). */
static int early_stop4(int64_t iteration, int golden_ratio_not_reached, double initial_entropy, const double *de,
                       int64_t n) {
    int64_t last = n - 1;
    if (de[last] == 0.0) return 1;
    if (iteration < 3) return 0;
    double average = de[last] + de[last - 1] + de[last - 2];
    average /= -3.0;
    double threshold = golden_ratio_not_reached ? 5e-4 * initial_entropy : 1e-4 * initial_entropy;
    return average < threshold;
}

/* finetune::early_stop, 3-argument form (src/finetune.cpp:285-294). */
static int early_stop3(int64_t iteration, double initial_entropy, const double *de, int64_t n) {
    if (iteration < 3) return 0;
    int64_t last = n - 1;
    double average = de[last] + de[last - 1] + de[last - 2];
    average /= -3.0;
    return average < 1e-4 * initial_entropy;
}

/* Stable descending sort by degree (Graph::sort_vertices, src/graph.cpp:235-268). */
typedef struct { int64_t deg; int32_t v; } degv_t;
static int degv_cmp(const void *a, const void *b) {
    const degv_t *x = (const degv_t *)a, *y = (const degv_t *)b;
    if (x->deg != y->deg) return x->deg > y->deg ? -1 : 1;
    return (x->v > y->v) - (x->v < y->v);
}
static int64_t vertex_degree(const orc_graph *g, int32_t v) {
    int64_t self = 0;
    for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) if (g->out_adj[e] == v) { self = 1; break; }
    return (g->out_off[v + 1] - g->out_off[v]) + (g->in_off[v + 1] - g->in_off[v]) - self;
}
static void split_by_degree(const orc_graph *g, double mh_percent, int32_t *hdv, int64_t *nh, int32_t *ldv,
                            int64_t *nl) {
    degv_t *d = (degv_t *)malloc(sizeof(degv_t) * (size_t)g->N);
    for (int32_t v = 0; v < g->N; ++v) { d[v].deg = vertex_degree(g, v); d[v].v = v; }
    qsort(d, (size_t)g->N, sizeof(degv_t), degv_cmp);
    *nh = 0; *nl = 0;
    for (int32_t i = 0; i < g->N; ++i) {
        if ((double)i < mh_percent * (double)g->N) hdv[(*nh)++] = d[i].v; else ldv[(*nl)++] = d[i].v;
    }
    free(d);
}

/* finetune::mcmc dispatch (src/finetune.cpp:754-771) and the sweeps' outer loops. */
int orc_mcmc(orc_bm *bm, const orc_graph *g, const orc_params *p, int32_t golden_ratio_not_reached, uint32_t *epoch,
             orc_stats *st) {
    bm_model_from(bm, g, p);
    int32_t N = g->N;
    if (bm->B == 1) return 0;
    int algo = p->algorithm;
    int max_it = p->max_iterations > 0 ? p->max_iterations : 100;
    double *de = (double *)malloc(sizeof(double) * (size_t)max_it);
    int64_t nde = 0;
    if (algo == 0 || algo == 2) {
        bm->overall_entropy = orc_mdl(bm, g);
        double initial = bm->overall_entropy, last = initial;
        int32_t *hdv = NULL, *ldv = NULL;
        int64_t nh = 0, nl = 0;
        if (algo == 2) {
            hdv = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
            ldv = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
            split_by_degree(g, p->mh_percent, hdv, &nh, ldv, &nl);
        }
        int32_t *moves_s = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
        for (int64_t it = 0; it < max_it; ++it) {
            uint32_t ep = (*epoch)++;
            if (algo == 0) {
                orc_async_gibbs_sweep(bm, g, p, ep, NULL, st);
            } else {
                /* hybrid_mcmc (src/finetune.cpp:563-646) */
                int32_t *h = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh ? nh : 1));
                int32_t *l = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nl ? nl : 1));
                memcpy(h, hdv, sizeof(int32_t) * (size_t)nh);
                memcpy(l, ldv, sizeof(int32_t) * (size_t)nl);
                if (!p->ordered) {
                    keyed_shuffle(h, nh, p->seed, ep, ORC_TAG_SHUFFLE);
                    keyed_shuffle(l, nl, p->seed, ep, ORC_TAG_SHUFFLE);
                }
                for (int64_t i = 0; i < nh; ++i) {
                    int32_t v = h[i], s;
                    if (bm->size[bm->assign[v]] == 1) continue;
                    double dS;
                    delta_t d;
                    proposal_t prop;
                    int acc = evaluate_vertex(bm, g, p->seed, ep, v, p->greedy, &s, &dS, &d, &prop);
                    if (st) st->proposals++;
                    if (acc) {
                        commit_delta(bm, v, &d, &prop);
                        delta_free(&d);
                        if (st) st->moves++;
                    }
                }
                int64_t batch_size = (int64_t)ceil((double)nl / (double)p->batches);
                for (int64_t batch = 0; batch_size > 0 && batch < nl / batch_size; ++batch) {
                    int64_t start = batch * batch_size;
                    int64_t end = (batch + 1) * batch_size < nl ? (batch + 1) * batch_size : nl;
                    gibbs_batch(bm, g, p, ep, l, start, end, moves_s, NULL, st);
                }
                free(h);
                free(l);
                if (st) st->sweeps++;
            }
            double entropy = orc_mdl(bm, g);
            de[nde++] = entropy - last;
            last = entropy;
            if (early_stop4(it, golden_ratio_not_reached, initial, de, nde)) break;
        }
        free(moves_s);
        free(hdv);
        free(ldv);
    } else {
        /* metropolis_hastings (src/finetune.cpp:773-820) */
        bm->overall_entropy = orc_mdl(bm, g);
        for (int64_t it = 0; it < max_it; ++it) {
            uint32_t ep = (*epoch)++;
            double sum = 0.0;
            orc_mh_sweep(bm, g, p, ep, p->ordered, &sum, st);
            de[nde++] = sum;
            if (early_stop4(it, golden_ratio_not_reached, bm->overall_entropy, de, nde)) break;
        }
    }
    bm->overall_entropy = orc_mdl(bm, g);
    free(de);
    return 0;
}

/* finetune::finetune_assignment (src/finetune.cpp:903-932): ordered sequential MH over the full graph. */
int orc_finetune_assignment(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t *epoch, orc_stats *st) {
    bm_model_from(bm, g, p);
    int max_it = p->max_iterations > 0 ? p->max_iterations : 100;
    double *de = (double *)malloc(sizeof(double) * (size_t)max_it);
    int64_t nde = 0;
    bm->overall_entropy = orc_mdl(bm, g);
    for (int64_t it = 0; it < max_it; ++it) {
        uint32_t ep = (*epoch)++;
        double sum = 0.0;
        orc_mh_sweep(bm, g, p, ep, 1, &sum, st);
        de[nde++] = sum;
        if (early_stop3(it, bm->overall_entropy, de, nde)) break;
    }
    bm->overall_entropy = orc_mdl(bm, g);
    free(de);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Block merge (src/block_merge.cpp)                                          */
/* ------------------------------------------------------------------------- */

/* block_merge::blockmodel_delta (src/block_merge.cpp:14-38). */
static void merge_delta(const orc_bm *bm, int32_t r, int32_t s, delta_t *d) {
    delta_init(d, r, s, bm->row[r].len + bm->col[r].len);
    const svec *row = &bm->row[r], *col = &bm->col[r];
    for (int32_t i = 0; i < row->len; ++i) {
        int32_t c = row->idx[i];
        int64_t v = row->val[i];
        if (c == r || c == s) delta_add(d, s, s, v); else delta_add(d, s, c, v);
        delta_add(d, r, c, -v);
    }
    for (int32_t i = 0; i < col->len; ++i) {
        int32_t rw = col->idx[i];
        if (rw == r) continue;
        int64_t v = col->val[i];
        if (rw == s) delta_add(d, s, s, v); else delta_add(d, rw, s, v);
        delta_add(d, rw, r, -v);
    }
}

/* entropy::block_merge_delta_mdl(long, ProposalAndEdgeCounts, const Blockmodel&, const Delta&) (src/entropy.cpp:129-182). */
static double merge_dS_from(const orc_bm *bm, const delta_t *d, int64_t k_out, int64_t k_in) {
    degfn f = { bm, d->r, d->s, k_in, k_out };
    fx128 acc = 0;
    imap_entries_delta(bm, d, &d->cur_row, 0, &f, 1, &acc);
    imap_entries_delta(bm, d, &d->prop_row, 1, &f, 1, &acc);
    imap_entries_delta(bm, d, &d->cur_col, 2, &f, 1, &acc);
    imap_entries_delta(bm, d, &d->prop_col, 3, &f, 1, &acc);
    int32_t s = d->s, r = d->r;
    const svec *row = &bm->row[s];
    for (int32_t i = 0; i < row->len; ++i) {
        int32_t col = row->idx[i];
        int64_t value = row->val[i];
        if (delta_get(d, s, col) != 0) continue;
        double a = cell_entropy(value, bm->d_in[col], bm->d_out[s]);
        double b = cell_entropy(value, new_d_in(&f, col), new_d_out(&f, s));
        acc += fx_from(a);
        acc -= fx_from(b);
        g_abs_terms += fabs(a) + fabs(b);
    }
    const svec *col = &bm->col[s];
    for (int32_t i = 0; i < col->len; ++i) {
        int32_t rw = col->idx[i];
        int64_t value = col->val[i];
        if (delta_get(d, rw, s) != 0 || rw == r || rw == s) continue;
        double a = cell_entropy(value, bm->d_in[s], bm->d_out[rw]);
        double b = cell_entropy(value, new_d_in(&f, s), new_d_out(&f, rw));
        acc += fx_from(a);
        acc -= fx_from(b);
        g_abs_terms += fabs(a) + fabs(b);
    }
    return fx_to(acc);
}

/* Replay-contract delta-MDL of merging r into s: Delta entries term by term, cell (s, s) when it is
 * unchanged, and the unchanged cells of row s / column s in bulk (their d_in / d_out do not change, so
 * each contributes M * (fastlog(d') - fastlog(d)) of the s degree).  Equal to
 * entropy::block_merge_delta_mdl (src/entropy.cpp:129-182) up to rounding; the per-cell form is
 * merge_dS_from and the tests compare the two. */
static double merge_dS_factored(const orc_bm *bm, const delta_t *d, int64_t k_out, int64_t k_in) {
    degfn f = { bm, d->r, d->s, k_in, k_out };
    const int32_t r = d->r, s = d->s;
    fx128 acc = 0;
    imap_entries_delta(bm, d, &d->cur_row, 0, &f, 1, &acc);
    imap_entries_delta(bm, d, &d->prop_row, 1, &f, 1, &acc);
    imap_entries_delta(bm, d, &d->cur_col, 2, &f, 1, &acc);
    imap_entries_delta(bm, d, &d->prop_col, 3, &f, 1, &acc);
    const int64_t mss = bm_get(bm, s, s), msr = bm_get(bm, s, r), mrs = bm_get(bm, r, s);
    if (mss != 0 && delta_get(d, s, s) == 0) {
        acc += fx_from(cell_entropy(mss, bm->d_in[s], bm->d_out[s]));
        acc -= fx_from(cell_entropy(mss, new_d_in(&f, s), new_d_out(&f, s)));
    }
    int64_t S_srow = bm->d_out[s] - mss - msr, S_scol = bm->d_in[s] - mss - mrs;
    for (int32_t i = 0; i < d->prop_row.cap; ++i)
        if (d->prop_row.used[i] && d->prop_row.key[i] != r && d->prop_row.key[i] != s) S_srow -= bm_get(bm, s, d->prop_row.key[i]);
    for (int32_t i = 0; i < d->prop_col.cap; ++i)
        if (d->prop_col.used[i]) S_scol -= bm_get(bm, d->prop_col.key[i], s);
    acc += fx_from((double)S_srow * (fastlog(new_d_out(&f, s)) - fastlog(bm->d_out[s])));
    acc += fx_from((double)S_scol * (fastlog(new_d_in(&f, s)) - fastlog(bm->d_in[s])));
    return fx_to(acc);
}

double orc_merge_delta_mdl(const orc_bm *bm, int32_t r, int32_t s) {
    delta_t d;
    merge_delta(bm, r, s, &d);
    double dS = bm->np ? np_merge_dS(bm, &d) : merge_dS_factored(bm, &d, bm->d_out[r], bm->d_in[r]);
    delta_free(&d);
    return dS;
}

/* entropy::block_merge_delta_mdl cell by cell; *abs_terms = sum of |cell terms|. */
double orc_merge_delta_mdl_cells(const orc_bm *bm, int32_t r, int32_t s, double *abs_terms) {
    delta_t d;
    merge_delta(bm, r, s, &d);
    g_abs_terms = 0.0;
    double dS = merge_dS_from(bm, &d, bm->d_out[r], bm->d_in[r]);
    if (abs_terms) *abs_terms = g_abs_terms;
    delta_free(&d);
    return dS;
}

/* propose_merge_sparse (src/block_merge.cpp:251-285) with merge-mode propose_new_block. */
static int32_t propose_merge(const orc_bm *bm, int32_t r, const uint32_t w[4]) {
    const svec *row = &bm->row[r], *col = &bm->col[r];
    int64_t k = bm->d_out[r] + bm->d_in[r];
    int32_t B = bm->B;
    if (k == 0) return random_block(r, B, w[1]);
    /* choose_neighbor over concat(row(r), col(r)) weights */
    uint32_t target = orc_uniform_below(w[0], (uint32_t)k);
    int64_t prefix = 0;
    int32_t t = -1;
    for (int32_t i = 0; i < row->len && t < 0; ++i) {
        prefix += row->val[i];
        if ((int64_t)target < prefix) t = row->idx[i];
    }
    for (int32_t i = 0; i < col->len && t < 0; ++i) {
        prefix += col->val[i];
        if ((int64_t)target < prefix) t = col->idx[i];
    }
    /* neighbors_weights(t) with edges[r] = 0 */
    const svec *tr = &bm->row[t], *tc = &bm->col[t];
    int64_t total = 0;
    for (int32_t i = 0; i < tr->len; ++i) if (tr->idx[i] != r) total += tr->val[i];
    for (int32_t i = 0; i < tc->len; ++i) if (tc->idx[i] != t && tc->idx[i] != r) total += tc->val[i];
    if (total == 0) return random_block(r, B, w[1]);
    uint32_t tg = orc_uniform_below(w[1], (uint32_t)total);
    prefix = 0;
    for (int32_t i = 0; i < tr->len; ++i) {
        if (tr->idx[i] == r) continue;
        prefix += tr->val[i];
        if ((int64_t)tg < prefix) return tr->idx[i];
    }
    for (int32_t i = 0; i < tc->len; ++i) {
        if (tc->idx[i] == t || tc->idx[i] == r) continue;
        prefix += tc->val[i];
        if ((int64_t)tg < prefix) return tc->idx[i];
    }
    abort();
}

/* block_merge::merge_blocks proposal loop (src/block_merge.cpp:189-217), NUM_AGG_PROPOSALS_PER_BLOCK = 10. */
int orc_merge_propose(const orc_bm *bm, uint64_t seed, uint32_t epoch, int32_t *best_s, double *best_dS) {
    int32_t B = bm->B;
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t r = 0; r < B; ++r) {
        int32_t past[10];
        int32_t npast = 0;
        best_s[r] = -1;
        best_dS[r] = DBL_MAX;
        for (uint32_t i = 0; i < 10; ++i) {
            uint32_t w[4];
            orc_philox(seed, (uint32_t)r, epoch, ORC_TAG_MERGE, i, w);
            int32_t s = propose_merge(bm, r, w);
            int seen = 0;
            for (int32_t j = 0; j < npast; ++j) if (past[j] == s) seen = 1;
            double dS = DBL_MAX;
            if (!seen) {
                dS = orc_merge_delta_mdl(bm, r, s);
                past[npast++] = s;
            }
            if (dS < best_dS[r]) {
                best_dS[r] = dS;
                best_s[r] = s;
            }
        }
    }
    return 0;
}

/* Min-heap on (dS, block); ties by ascending block id. */
typedef struct { double dS; int32_t r, s; } merge_t;
static int merge_less(const merge_t *a, const merge_t *b) {
    if (a->dS != b->dS) return a->dS < b->dS;
    return a->r < b->r;
}
static void heap_push(merge_t *h, int64_t *n, merge_t m) {
    int64_t i = (*n)++;
    h[i] = m;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!merge_less(&h[i], &h[p])) break;
        merge_t t = h[i]; h[i] = h[p]; h[p] = t;
        i = p;
    }
}
static merge_t heap_pop(merge_t *h, int64_t *n) {
    merge_t top = h[0];
    h[0] = h[--(*n)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, rr = l + 1, m = i;
        if (l < *n && merge_less(&h[l], &h[m])) m = l;
        if (rr < *n && merge_less(&h[rr], &h[m])) m = rr;
        if (m == i) break;
        merge_t t = h[i]; h[i] = h[m]; h[m] = t;
        i = m;
    }
    return top;
}

/* Stable ascending argsort of doubles (utils::partial_sort_indices, include/utils.hpp:143-153). */
static const double *g_sort_keys;
static int idx_cmp(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    double kx = g_sort_keys[x], ky = g_sort_keys[y];
    if (kx < ky) return -1;
    if (kx > ky) return 1;
    return (x > y) - (x < y);
}

/* Relabel with block_map, compact labels in ascending order (Blockmodel::build_mapping), rebuild. */
static orc_bm *finish_merge(orc_bm *bm, const orc_graph *g, const int32_t *block_map, int64_t merged) {
    int32_t B = bm->B;
    int32_t *assign = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->N);
    for (int32_t v = 0; v < g->N; ++v) assign[v] = block_map[bm->assign[v]];
    int32_t *used = (int32_t *)calloc((size_t)B, sizeof(int32_t));
    for (int32_t v = 0; v < g->N; ++v) used[assign[v]] = 1;
    int32_t next = 0;
    for (int32_t b = 0; b < B; ++b) if (used[b]) used[b] = next++; else used[b] = -1;
    for (int32_t v = 0; v < g->N; ++v) assign[v] = used[assign[v]];
    int32_t newB = B - (int32_t)merged;
    if (next > newB) newB = next;
    orc_bm *out = orc_bm_new(g, newB, assign);
    orc_bm_set_model(out, g, bm->dc ? 2 : bm->np);
    out->overall_entropy = bm->overall_entropy;
    free(used);
    free(assign);
    return out;
}

orc_bm *orc_merge_apply(orc_bm *bm, const orc_graph *g, const int32_t *best_s, const double *best_dS,
                        int64_t num_to_merge, int32_t approximate) {
    int32_t B = bm->B;
    int32_t *block_map = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
    for (int32_t b = 0; b < B; ++b) block_map[b] = b;
    int64_t merged = 0;
    if (approximate) {
        /* Blockmodel::carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122) */
        int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
        for (int32_t b = 0; b < B; ++b) order[b] = b;
        g_sort_keys = best_dS;
        qsort(order, (size_t)B, sizeof(int32_t), idx_cmp);
        int64_t counter = 0;
        while (merged < num_to_merge && counter < B) {
            int32_t from = order[counter++];
            int32_t b = best_s[from];
            do { b = block_map[b]; } while (block_map[b] != b);
            if (b != from) {
                block_map[from] = b;
                merged++;
            }
        }
        for (int32_t x = 0; x < B; ++x) {
            int32_t b = x;
            do { b = block_map[b]; } while (block_map[b] != b);
            order[x] = b;
        }
        memcpy(block_map, order, sizeof(int32_t) * (size_t)B);
        free(order);
    } else {
        /* carry_out_best_merges_advanced (src/block_merge.cpp:40-117) with union-find in place of the
         * O(B) relabel: block_map values stay roots, so find() returns exactly the relabelled entry. */
        merge_t *heap = (merge_t *)malloc(sizeof(merge_t) * (size_t)(B + 1));
        int64_t n = 0;
        for (int32_t b = 0; b < B; ++b) {
            merge_t m = { best_dS[b], b, best_s[b] };
            heap_push(heap, &n, m);
        }
        int32_t *parent = block_map;
        while (merged < num_to_merge && n > 0) {
            merge_t m = heap_pop(heap, &n);
            int32_t from = m.r;
            int32_t to = m.s;
            while (parent[to] != to) { parent[to] = parent[parent[to]]; to = parent[to]; }
            if (from == to) continue;
            int64_t k_out = bm->d_out[from], k_in = bm->d_in[from];
            delta_t d;
            merge_delta(bm, from, to, &d);
            double actual = bm->np ? np_merge_dS(bm, &d) : merge_dS_factored(bm, &d, k_out, k_in);
            if (n > 0 && actual > heap[0].dS) {
                m.dS = actual;
                heap_push(heap, &n, m);
                delta_free(&d);
                continue;
            }
            parent[from] = to;
            /* Blockmodel::merge_block (src/blockmodel/blockmodel.cpp:337-363) */
            int64_t self_to = bm_get(bm, to, to) + delta_get(&d, to, to);
            const imap *maps[4] = { &d.cur_row, &d.prop_row, &d.cur_col, &d.prop_col };
            for (int which = 0; which < 4; ++which) {
                const imap *mp = maps[which];
                for (int32_t i = 0; i < mp->cap; ++i) {
                    if (!mp->used[i] || mp->val[i] == 0) continue;
                    int32_t row, col;
                    if (which == 0) { row = from; col = mp->key[i]; }
                    else if (which == 1) { row = to; col = mp->key[i]; }
                    else if (which == 2) { row = mp->key[i]; col = from; }
                    else { row = mp->key[i]; col = to; }
                    bm_add(bm, row, col, mp->val[i]);
                }
            }
            delta_free(&d);
            bm->d_out[from] = 0;
            bm->d_out[to] += k_out;
            bm->d_in[from] = 0;
            bm->d_in[to] += k_in;
            bm->d_tot[from] = 0;
            bm->d_tot[to] = bm->d_out[to] + bm->d_in[to] - self_to;
            bm->size[to] += bm->size[from];
            bm->size[from] = 0;
            if (bm->dc) {  /* Blockmodel::merge_block adds the histograms of `from` to `to` (blockmodel.cpp:357-362) */
                for (int32_t i = 0; i < bm->hin[from].len; ++i) sv_add(&bm->hin[to], bm->hin[from].idx[i], bm->hin[from].val[i]);
                for (int32_t i = 0; i < bm->hout[from].len; ++i) sv_add(&bm->hout[to], bm->hout[from].idx[i], bm->hout[from].val[i]);
            }
            merged++;
        }
        for (int32_t b = 0; b < B; ++b) {
            int32_t x = b;
            while (parent[x] != x) x = parent[x];
            int32_t y = b;
            while (parent[y] != x) { int32_t nx = parent[y]; parent[y] = x; y = nx; }
        }
        free(heap);
    }
    orc_bm *out = finish_merge(bm, g, block_map, merged);
    free(block_map);
    return out;
}

/* block_merge::merge_blocks (src/block_merge.cpp:178-228). Consumes bm. */
orc_bm *orc_merge_blocks(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch, orc_stats *st) {
    bm_model_from(bm, g, p);
    int32_t B = bm->B;
    int32_t *best_s = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
    double *best_dS = (double *)malloc(sizeof(double) * (size_t)B);
    orc_merge_propose(bm, p->seed, epoch, best_s, best_dS);
    if (st) { st->merge_proposals += 10 * (int64_t)B; st->merge_phases++; }
    orc_bm *out = orc_merge_apply(bm, g, best_s, best_dS, bm->num_to_merge, p->approximate);
    free(best_s);
    free(best_dS);
    orc_bm_free(bm);
    return out;
}

/* ------------------------------------------------------------------------- */
/* Driver: BlockmodelTriplet + sbp::stochastic_block_partition               */
/* (src/blockmodel/blockmodel_triplet.cpp:4-105, src/sbp.cpp:16-86)          */
/* ------------------------------------------------------------------------- */
typedef struct { orc_bm *bm[3]; int optimal; } triplet_t;

static int tri_empty(const triplet_t *t, int i) { return t->bm[i] == NULL; }
static void tri_set(triplet_t *t, int i, orc_bm *b) { if (t->bm[i]) orc_bm_free(t->bm[i]); t->bm[i] = b; }

static void tri_update(triplet_t *t, orc_bm *bm) {
    int index;
    if (tri_empty(t, 1)) {
        index = 1;
    } else {
        if (bm->overall_entropy <= t->bm[1]->overall_entropy) {
            int old_index = (t->bm[1]->B > bm->B) ? 0 : 2;
            tri_set(t, old_index, orc_bm_copy(t->bm[1]));
            index = 1;
        } else {
            index = (t->bm[1]->B > bm->B) ? 2 : 0;
        }
    }
    tri_set(t, index, orc_bm_copy(bm));
}

static int tri_is_done(triplet_t *t) {
    if ((!tri_empty(t, 0) && t->bm[0]->B - t->bm[2]->B == 2) ||
        (tri_empty(t, 0) && t->bm[1]->B - t->bm[2]->B == 1))
        t->optimal = 1;
    return t->optimal;
}

static orc_bm *tri_next(triplet_t *t, orc_bm *old) {
    old->num_to_merge = 0;
    tri_update(t, old);
    if (tri_empty(t, 2)) {
        orc_bm *b = orc_bm_copy(t->bm[1]);
        b->num_to_merge = (int64_t)((double)b->B * 0.5);
        if (b->num_to_merge == 0) t->optimal = 1;
        return b;
    }
    if (tri_is_done(t)) return orc_bm_copy(t->bm[1]);
    int index;
    if (tri_empty(t, 0) && t->bm[1]->B > t->bm[2]->B) {
        index = 1;
    } else {
        int64_t upper = tri_empty(t, 0) ? 0 - (int64_t)t->bm[1]->B : (int64_t)t->bm[0]->B - t->bm[1]->B;
        int64_t lower = (int64_t)t->bm[1]->B - t->bm[2]->B;
        index = (upper >= lower) ? 0 : 1;
    }
    int64_t next = t->bm[index + 1]->B;
    next += (int64_t)((double)(t->bm[index]->B - t->bm[index + 1]->B) * 0.618);
    orc_bm *b = orc_bm_copy(t->bm[index]);
    b->num_to_merge = b->B - next;
    return b;
}

int orc_sbp(const orc_graph *g, const orc_params *p, int32_t *assignment_out, int32_t *num_blocks_out,
            orc_stats *st) {
    int32_t N = g->N;
    int32_t *ident = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int32_t v = 0; v < N; ++v) ident[v] = v;
    orc_bm *bm = orc_bm_new(g, N, ident);
    free(ident);
    triplet_t t = { { NULL, NULL, NULL }, 0 };
    uint32_t epoch = 0;
    int64_t iteration = 0;
    while (!t.optimal) {
        bm = orc_merge_blocks(bm, g, p, epoch++, st);
        /* finetune::mcmc: async_gibbs only while iteration < asynciterations, else Metropolis-Hastings */
        orc_params q = *p;
        if (p->algorithm == 0 && iteration >= p->asynciterations) q.algorithm = 1;
        orc_mcmc(bm, g, &q, tri_empty(&t, 2), &epoch, st);
        iteration++;
        orc_bm *next = tri_next(&t, bm);
        orc_bm_free(bm);
        bm = next;
    }
    if (assignment_out) memcpy(assignment_out, bm->assign, sizeof(int32_t) * (size_t)N);
    if (num_blocks_out) *num_blocks_out = bm->B;
    if (st) st->mdl = bm->overall_entropy;
    orc_bm_free(bm);
    for (int i = 0; i < 3; ++i) if (t.bm[i]) orc_bm_free(t.bm[i]);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Sampling (src/sample.cpp)                                                  */
/* ------------------------------------------------------------------------- */

/* Sampling orders.  Each returns the sampled vertices in sample-id order.
 *   alg 0 = max_degree (sample.cpp:219-233): stable descending degree (utils::argsort), ties by vertex id;
 *   alg 1 = random (sample.cpp:235-247): ascending Philox key (TAG_SAMPLE, epoch 0), ties by vertex id;
 *   alg 2 = degree_product (sample.cpp:28-49): out-edges in CSR order, stable-sorted by descending
 *           degree(src) * degree(dst) (Graph::sorted_edge_list, graph.cpp:302-316); endpoints are taken in
 *           that order (source before destination) until the count reaches the target, checked after each
 *           edge;
 *   alg 3 = degree_weighted (new): weighted sampling without replacement with weights = degree, by
 *           Efraimidis-Spirakis keys: key(v) = -logf(u_v) / degree(v) in float, u_v = (2 (w0 >> 9) + 1) 2^-24,
 *           ascending key (degree 0: +inf), ties by vertex id.
 *   alg 4 = expansion_snowball (sample.cpp:94-155) and alg 5 = degree_product_snowball (sample.cpp:51-92):
 *           see orc_snowball below.
 * The target is (long)(fraction * N) (sample.cpp:228).  Writes mapping[N] (full id -> sample id or -1) and
 * vertices[n] (sample id -> full id). */
typedef struct { uint64_t key; int64_t idx; } key64_t;
static int key64_cmp(const void *a, const void *b) {
    const key64_t *x = (const key64_t *)a, *y = (const key64_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}
float orc_degree_weight_key(uint32_t w0, int64_t degree) {
    if (degree <= 0) return INFINITY;
    float u = (float)(((w0 >> 9) << 1) | 1u) * 0x1p-24f;
    return -logf(u) / (float)degree;
}
/* Snowball samplers, literally: common::random_integer(lo, hi) is draw j of
 * Philox(seed; j, 0, TAG_SNOWBALL, 0), lo + (w0 (hi - lo + 1)) >> 32, j counting the calls.
 * expansion_snowball (sample.cpp:94-155; es_add_vertex :94-110, es_update_contribution :112-124): the
 * neighbour set is a flag array whose k-th set entry in ascending id stands for std::advance over the
 * std::set; argmax (include/utils.hpp:241-253) scans all vertices for the first maximum, sampled ones
 * included; a vertex chosen again is appended again and its mapping moves to the new slot.
 * degree_product_snowball (sample.cpp:51-92): std::priority_queue<(src, dst, product)> ordered by product,
 * restated with libstdc++'s heap operations (push_heap sift-up; pop_heap moves the last element to a hole
 * that descends along the larger child, the right one unless the left is strictly larger, then sifts up),
 * which fix the order in which equal products pop.  When a restart leaves the frontier empty the reference
 * reads top() of an empty queue (undefined); here the loop continues.  Returns the vertices in sample
 * order; mapping[v] = v's last slot or -1. */
static int64_t snow_draw(uint64_t seed, uint32_t *next, int64_t lo, int64_t hi) {
    uint32_t w[4];
    orc_philox(seed, (*next)++, 0, ORC_TAG_SNOWBALL, 0, w);
    return lo + (int64_t)(((uint64_t)w[0] * (uint64_t)(hi - lo + 1)) >> 32);
}
typedef struct { int64_t src, dst, prod; } snow_edge;
static void snow_sift_up(snow_edge *h, int64_t hole, int64_t top, snow_edge v) {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && h[parent].prod < v.prod) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = v;
}
static void snow_push(snow_edge *h, int64_t *len, snow_edge v) { snow_sift_up(h, (*len)++, 0, v); }
static snow_edge snow_pop(snow_edge *h, int64_t *len) {
    snow_edge top = h[0];
    int64_t n = *len - 1;  /* elements left after the pop */
    snow_edge v = h[n];
    h[n] = h[0];
    int64_t hole = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (h[child].prod < h[child - 1].prod) child--;
        h[hole] = h[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    snow_sift_up(h, hole, 0, v);
    *len = n;
    return top;
}
static int64_t orc_snowball(const orc_graph *g, int32_t alg, int64_t target, uint64_t seed, int32_t *mapping,
                            int32_t *order) {
    int32_t N = g->N;
    uint32_t next = 0;
    int64_t n = 0;
    for (int32_t v = 0; v < N; ++v) mapping[v] = -1;
    if (alg == 4) {
        int64_t *contrib = (int64_t *)calloc((size_t)N, sizeof(int64_t));
        char *flag = (char *)calloc((size_t)N, 1);
        int64_t sum = 0, nset = 0;
#define ES_UPDATE(u)                                                                   \
        do {                                                                           \
            for (int64_t e_ = g->out_off[u]; e_ < g->out_off[(u) + 1]; ++e_) {         \
                int32_t nb_ = g->out_adj[e_];                                          \
                if (flag[nb_] || mapping[nb_] >= 0) continue;                          \
                contrib[u]++, sum++;                                                   \
            }                                                                          \
            for (int64_t e_ = g->in_off[u]; e_ < g->in_off[(u) + 1]; ++e_) {           \
                int32_t nb_ = g->in_adj[e_];                                           \
                if (contrib[nb_] > 0) contrib[nb_]--, sum--;                           \
            }                                                                          \
        } while (0)
#define ES_ADD(x)                                                                      \
        do {                                                                           \
            int32_t v_ = (x);                                                          \
            order[n] = v_;                                                             \
            mapping[v_] = (int32_t)n++;                                                \
            for (int64_t e = g->out_off[v_]; e < g->out_off[v_ + 1]; ++e) {            \
                int32_t nb = g->out_adj[e];                                            \
                if (flag[nb] || mapping[nb] >= 0) continue;                            \
                flag[nb] = 1, nset++;                                                  \
                if (contrib[nb] > 0) continue;                                         \
                ES_UPDATE(nb);                                                         \
            }                                                                          \
            if (flag[v_]) flag[v_] = 0, nset--;                                        \
        } while (0)
        ES_ADD((int32_t)snow_draw(seed, &next, 0, N - 1));
        while (n < target) {
            if (nset == 0) {
                int32_t v;
                do v = (int32_t)snow_draw(seed, &next, 0, N - 1); while (mapping[v] >= 0);
                ES_ADD(v);
            } else if (sum == 0) {
                int64_t k = snow_draw(seed, &next, 0, nset - 1), seen = -1;
                int32_t v = 0;
                for (int32_t u = 0; u < N; ++u)
                    if (flag[u] && ++seen == k) { v = u; break; }
                ES_ADD(v);
            } else {
                int32_t best = 0;
                for (int32_t u = 1; u < N; ++u)
                    if (contrib[u] > contrib[best]) best = u;
                ES_ADD(best);
            }
        }
#undef ES_ADD
#undef ES_UPDATE
        free(contrib);
        free(flag);
    } else {
        int64_t *deg = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
        for (int32_t v = 0; v < N; ++v) deg[v] = vertex_degree(g, v);
        snow_edge *h = (snow_edge *)malloc(sizeof(snow_edge) * (size_t)(2 * g->E + 1));
        int64_t len = 0;
#define DPS_ADD(x)                                                                     \
        do {                                                                           \
            int32_t v_ = (x);                                                          \
            mapping[v_] = (int32_t)n;                                                  \
            order[n++] = v_;                                                           \
            for (int64_t e = g->out_off[v_]; e < g->out_off[v_ + 1]; ++e) {            \
                int32_t nb = g->out_adj[e];                                            \
                if (mapping[nb] != -1) continue;                                       \
                snow_edge x_ = { v_, nb, deg[v_] * deg[nb] };                          \
                snow_push(h, &len, x_);                                                \
            }                                                                          \
            for (int64_t e = g->in_off[v_]; e < g->in_off[v_ + 1]; ++e) {              \
                int32_t nb = g->in_adj[e];                                             \
                if (mapping[nb] != -1) continue;                                       \
                snow_edge x_ = { nb, v_, deg[nb] * deg[v_] };                          \
                snow_push(h, &len, x_);                                                \
            }                                                                          \
        } while (0)
        DPS_ADD((int32_t)snow_draw(seed, &next, 0, N - 1));
        while (n < target) {
            if (len == 0) {
                int32_t v;
                do v = (int32_t)snow_draw(seed, &next, 0, N - 1); while (mapping[v] >= 0);
                DPS_ADD(v);
                if (len == 0) continue;
            }
            snow_edge top = snow_pop(h, &len);
            if (mapping[top.src] == -1) DPS_ADD((int32_t)top.src);
            if (mapping[top.dst] == -1) DPS_ADD((int32_t)top.dst);
        }
#undef DPS_ADD
        free(h);
        free(deg);
    }
    return n;
}

int64_t orc_sample(const orc_graph *g, int32_t alg, double fraction, uint64_t seed, int32_t *mapping,
                   int32_t *vertices) {
    int32_t N = g->N;
    int64_t target = (int64_t)(fraction * (double)N);
    if (alg == 4 || alg == 5) return orc_snowball(g, alg, target, seed, mapping, vertices);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int64_t n = 0;
    for (int32_t v = 0; v < N; ++v) mapping[v] = -1;
    if (alg == 0) {
        degv_t *d = (degv_t *)malloc(sizeof(degv_t) * (size_t)N);
        for (int32_t v = 0; v < N; ++v) { d[v].deg = vertex_degree(g, v); d[v].v = v; }
        qsort(d, (size_t)N, sizeof(degv_t), degv_cmp);
        for (int32_t i = 0; i < N; ++i) order[i] = d[i].v;
        free(d);
        n = target;
    } else if (alg == 1) {
        for (int32_t i = 0; i < N; ++i) order[i] = i;
        keyed_shuffle(order, N, seed, 0, ORC_TAG_SAMPLE);
        n = target;
    } else if (alg == 2) {
        int64_t E = g->E;
        int64_t *deg = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
        for (int32_t v = 0; v < N; ++v) deg[v] = vertex_degree(g, v);
        key64_t *k = (key64_t *)malloc(sizeof(key64_t) * (size_t)(E ? E : 1));
        int32_t *src = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E ? E : 1));
        for (int32_t v = 0; v < N; ++v)
            for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) src[e] = v;
        for (int64_t e = 0; e < E; ++e) {
            k[e].key = ~(uint64_t)(deg[src[e]] * deg[g->out_adj[e]]);  /* descending product */
            k[e].idx = e;                                             /* stable */
        }
        qsort(k, (size_t)E, sizeof(key64_t), key64_cmp);
        for (int64_t i = 0; i < E; ++i) {
            int32_t ends[2] = { src[k[i].idx], g->out_adj[k[i].idx] };
            for (int j = 0; j < 2; ++j)
                if (mapping[ends[j]] == -1) { mapping[ends[j]] = (int32_t)n; order[n++] = ends[j]; }
            if (n >= target) break;
        }
        free(k); free(src); free(deg);
    } else {
        key64_t *k = (key64_t *)malloc(sizeof(key64_t) * (size_t)N);
        for (int32_t v = 0; v < N; ++v) {
            uint32_t w[4];
            orc_philox(seed, (uint32_t)v, 0, ORC_TAG_SAMPLE, 0, w);
            float key = orc_degree_weight_key(w[0], vertex_degree(g, v));
            uint32_t bits;
            memcpy(&bits, &key, 4);
            k[v].key = bits;
            k[v].idx = v;
        }
        qsort(k, (size_t)N, sizeof(key64_t), key64_cmp);
        for (int32_t i = 0; i < N; ++i) order[i] = (int32_t)k[i].idx;
        free(k);
        n = target;
    }
    if (alg != 2)
        for (int64_t i = 0; i < n; ++i) mapping[order[i]] = (int32_t)i;
    for (int64_t i = 0; i < n; ++i) vertices[i] = order[i];
    free(order);
    return n;
}

/* sample::from_vertices (src/sample.cpp:202-217): the induced subgraph in sample ids.  Edges are added
 * for full vertices in ascending id, each in its out-list order (Graph::add_edge appends to the source's
 * out-list and the destination's in-list), so a sample vertex's out-list keeps its full out-list order and
 * its in-list is ordered by (full source id, position in the source's out-list).  Buffers hold n + 1
 * offsets and up to E entries.  Returns the sample's edge count. */
int64_t orc_subgraph(const orc_graph *g, const int32_t *mapping, const int32_t *vertices, int32_t n,
                     int64_t *out_off, int32_t *out_adj, int64_t *in_off, int32_t *in_adj) {
    int32_t N = g->N;
    for (int32_t i = 0; i <= n; ++i) { out_off[i] = 0; in_off[i] = 0; }
    for (int32_t i = 0; i < n; ++i) {
        int32_t f = vertices[i];
        for (int64_t e = g->out_off[f]; e < g->out_off[f + 1]; ++e) {
            int32_t d = mapping[g->out_adj[e]];
            if (d < 0) continue;
            out_off[i + 1]++;
            in_off[d + 1]++;
        }
    }
    for (int32_t i = 0; i < n; ++i) { out_off[i + 1] += out_off[i]; in_off[i + 1] += in_off[i]; }
    int64_t *ipos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int32_t i = 0; i <= n; ++i) ipos[i] = in_off[i];
    for (int32_t i = 0; i < n; ++i) {
        int64_t o = out_off[i];
        int32_t f = vertices[i];
        for (int64_t e = g->out_off[f]; e < g->out_off[f + 1]; ++e) {
            int32_t d = mapping[g->out_adj[e]];
            if (d >= 0) out_adj[o++] = d;
        }
    }
    for (int32_t f = 0; f < N; ++f) {  /* in-lists in ascending full source id */
        int32_t sf = mapping[f];
        if (sf < 0) continue;
        for (int64_t e = g->out_off[f]; e < g->out_off[f + 1]; ++e) {
            int32_t d = mapping[g->out_adj[e]];
            if (d >= 0) in_adj[ipos[d]++] = sf;
        }
    }
    free(ipos);
    return out_off[n];
}

/* sample::extend (src/sample.cpp:157-200): sequential plurality vote; ties to the smallest block. */
/* sample::detach (src/sample.cpp:12-26): the vertices of degree > 1 (Graph::degrees, src/graph.cpp:24-31)
 * in id order.  Returns their count; mapping[v] = index or -1. */
int64_t orc_detach(const orc_graph *g, int32_t *mapping, int32_t *vertices) {
    int64_t n = 0;
    for (int32_t v = 0; v < g->N; ++v) {
        mapping[v] = -1;
        if (vertex_degree(g, v) > 1) {
            mapping[v] = (int32_t)n;
            vertices[n++] = v;
        }
    }
    return n;
}

/* sample::reattach (src/sample.cpp:367-405).  Each vertex still unassigned at its turn (id order) draws
 * random_integer(0, B - 1), here draw j of Philox(seed; j, 0, TAG_REATTACH, 0); then every out-neighbour,
 * then every in-neighbour, in list order: an assigned neighbour's block is copied to the vertex, an
 * unassigned one takes the drawn block together with the vertex and ends that list; a vertex left without a
 * block takes the drawn one. */
int orc_reattach(const orc_graph *g, const int32_t *mapping, const int32_t *detached_assignment,
                 int32_t detached_blocks, uint64_t seed, int32_t *assignment) {
    const int32_t N = g->N;
    uint32_t next = 0;
    for (int32_t v = 0; v < N; ++v) assignment[v] = mapping[v] >= 0 ? detached_assignment[mapping[v]] : -1;
    for (int32_t v = 0; v < N; ++v) {
        if (assignment[v] != -1) continue;
        uint32_t w[4];
        orc_philox(seed, next++, 0, ORC_TAG_REATTACH, 0, w);
        const int32_t random_community = (int32_t)(((uint64_t)w[0] * (uint64_t)detached_blocks) >> 32);
        for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) {
            int32_t u = g->out_adj[e];
            if (assignment[u] == -1) {
                assignment[v] = random_community;
                assignment[u] = random_community;
                break;
            }
            assignment[v] = assignment[u];
        }
        for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
            int32_t u = g->in_adj[e];
            if (assignment[u] == -1) {
                assignment[v] = random_community;
                assignment[u] = random_community;
                break;
            }
            assignment[v] = assignment[u];
        }
        if (assignment[v] < 0) assignment[v] = random_community;
    }
    return 0;
}

int orc_extend(const orc_graph *g, const int32_t *mapping, const int32_t *sample_assignment, int32_t sample_blocks,
               uint64_t seed, int32_t *assignment) {
    int32_t N = g->N;
    for (int32_t v = 0; v < N; ++v) assignment[v] = mapping[v] >= 0 ? sample_assignment[mapping[v]] : -1;
    int64_t *counts = (int64_t *)calloc((size_t)sample_blocks, sizeof(int64_t));
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)(sample_blocks + 1));
    for (int32_t v = 0; v < N; ++v) {
        if (assignment[v] != -1) continue;
        int32_t nt = 0;
        for (int64_t e = g->out_off[v]; e < g->out_off[v + 1]; ++e) {
            int32_t c = assignment[g->out_adj[e]];
            if (c == -1) continue;
            if (counts[c]++ == 0) touched[nt++] = c;
        }
        for (int64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
            int32_t u = g->in_adj[e];
            int32_t c = assignment[u];
            if (c == -1 || u == v) continue;
            if (counts[c]++ == 0) touched[nt++] = c;
        }
        if (nt == 0) {
            uint32_t w[4];
            orc_philox(seed, (uint32_t)v, 0, ORC_TAG_EXTEND, 0, w);
            assignment[v] = (int32_t)orc_uniform_below(w[0], (uint32_t)sample_blocks);
            continue;
        }
        int64_t best = 0;
        int32_t best_c = -1;
        for (int32_t i = 0; i < nt; ++i) {
            int32_t c = touched[i];
            if (counts[c] > best || (counts[c] == best && c < best_c)) { best = counts[c]; best_c = c; }
            counts[c] = 0;
        }
        assignment[v] = best_c;
    }
    free(counts);
    free(touched);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Divisive step: divisive::split_communities (src/top_down_sbp.cpp:488-549) */
/* with propose_split / propose_random_split (:191-240), extract_subgraph     */
/* (:98-132) and apply_best_splits (:30-66), literally: each block's induced  */
/* subgraph is built, each proposal builds the two-block blockmodel of it and */
/* evaluates nonparametric::mdl on it.  Draws follow the replay contract:     */
/* proposal p of block b has size lo + uniform_below(w, hi - lo + 1), w word  */
/* p % 4 of Philox(seed; b, epoch, ORC_TAG_SPLIT_SIZE, p / 4); subgraph       */
/* vertex i (full id v) is on side 0 when uniform_below(w, n) < size, w word  */
/* p % 4 of Philox(seed; v, epoch, ORC_TAG_SPLIT_SIDE, p / 4).                */
/* ------------------------------------------------------------------------- */
#define ORC_SPLIT_PROPOSALS 10  /* NUM_AGG_PROPOSALS_PER_BLOCK (include/top_down_sbp.hpp:28) */

static void split_sides(uint64_t seed, uint32_t epoch, int32_t b, int32_t p, const int32_t *vertices, int32_t n,
                        int32_t *side) {
    uint32_t w[4];
    orc_philox(seed, (uint32_t)b, epoch, ORC_TAG_SPLIT_SIZE, (uint32_t)(p / 4), w);
    /* common::random_integer(n * 0.2, n * 0.8): long parameters, so the bounds truncate */
    long lo = (long)((double)n * 0.2), hi = (long)((double)n * 0.8);
    long k = lo + (long)orc_uniform_below(w[p % 4], (uint32_t)(hi - lo + 1));
    for (int32_t i = 0; i < n; ++i) {
        orc_philox(seed, (uint32_t)vertices[i], epoch, ORC_TAG_SPLIT_SIDE, (uint32_t)(p / 4), w);
        /* discrete_distribution({k, n - k}): side 0 with probability k / n */
        side[i] = (long)orc_uniform_below(w[p % 4], (uint32_t)n) < k ? 0 : 1;
    }
}

/* A stable sort of block ids by descending value (utils::argsort, include/utils.hpp:165-172). */
static void argsort_desc(const double *val, int32_t *idx, int32_t *tmp, int32_t n) {
    for (int32_t w = 1; w < n; w *= 2) {
        for (int32_t lo = 0; lo < n; lo += 2 * w) {
            int32_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int32_t a = lo, c = mid, o = lo;
            while (a < mid && c < hi) tmp[o++] = (val[idx[c]] > val[idx[a]]) ? idx[c++] : idx[a++];
            while (a < mid) tmp[o++] = idx[a++];
            while (c < hi) tmp[o++] = idx[c++];
        }
        memcpy(idx, tmp, sizeof(int32_t) * (size_t)n);
    }
}

int32_t orc_split_communities(const orc_graph *g, int32_t *assign, int32_t B, int32_t target, uint64_t seed,
                              uint32_t epoch, int32_t nonparametric, double *delta_out) {
    int32_t N = g->N;
    int32_t *size = (int32_t *)calloc((size_t)B + 1, sizeof(int32_t));
    for (int32_t v = 0; v < N; ++v) size[assign[v]]++;
    int32_t *mapping = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N + 1));
    int32_t *vertices = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N + 1));
    int32_t *side = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N + 1));
    int32_t *zeros = (int32_t *)calloc((size_t)N + 1, sizeof(int32_t));
    int64_t cap = g->E > 0 ? g->E : 1;
    int64_t *out_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 2));
    int64_t *in_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 2));
    int32_t *out_adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t *in_adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    double *delta = (double *)malloc(sizeof(double) * (size_t)(B + 1));
    int32_t *choice = (int32_t *)malloc(sizeof(int32_t) * (size_t)(B + 1));
    for (int32_t v = 0; v < N; ++v) mapping[v] = -1;
    for (int32_t b = 0; b < B; ++b) {
        delta[b] = DBL_MAX;
        choice[b] = -1;
        /* extract_subgraph: the block's vertices ascending, translator[vertex] = index */
        int32_t n = 0;
        for (int32_t v = 0; v < N; ++v)
            if (assign[v] == b) { mapping[v] = n; vertices[n++] = v; }
        /* Do not attempt to split small clusters (:506) */
        if ((double)size[b] < 0.001 * (double)N) {
            for (int32_t i = 0; i < n; ++i) mapping[vertices[i]] = -1;
            continue;
        }
        orc_graph sub = {n, 0, out_off, out_adj, in_off, in_adj};
        if (n < 2) {  /* a subgraph of fewer than 2 vertices keeps no edges (:114-116) */
            for (int32_t i = 0; i <= n; ++i) { out_off[i] = 0; in_off[i] = 0; }
        } else {
            sub.E = orc_subgraph(g, mapping, vertices, n, out_off, out_adj, in_off, in_adj);
        }
        double old_entropy;
        if (nonparametric) {  /* null_mdl_v1: the one-block nonparametric model of the subgraph */
            orc_bm *null_model = orc_bm_new(&sub, 1, zeros);
            old_entropy = orc_np_mdl(null_model, &sub);
            orc_bm_free(null_model);
        } else {
            old_entropy = orc_null_mdl_v1(sub.E);
        }
        for (int32_t p = 0; p < ORC_SPLIT_PROPOSALS; ++p) {
            orc_bm *split;
            if (n < 2) split = orc_bm_new(&sub, 1, zeros);
            else {
                split_sides(seed, epoch, b, p, vertices, n, side);
                split = orc_bm_new(&sub, 2, side);
            }
            double new_entropy = orc_np_mdl(split, &sub);
            orc_bm_free(split);
            double d = new_entropy - old_entropy;
            if (d < delta[b]) { delta[b] = d; choice[b] = p; }
        }
        for (int32_t i = 0; i < n; ++i) mapping[vertices[i]] = -1;
    }
    if (delta_out) memcpy(delta_out, delta, sizeof(double) * (size_t)B);
    /* apply_best_splits */
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(B + 1));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(B + 1));
    for (int32_t b = 0; b < B; ++b) order[b] = b;
    argsort_desc(delta, order, tmp, B);
    int32_t num_blocks = B;
    for (int32_t j = B - 1; j > -1; --j) {
        int32_t b = order[j];
        if (choice[b] < 0) continue;  /* the default Split: num_vertices = LONG_MAX > N */
        int32_t n = 0;
        for (int32_t v = 0; v < N; ++v)
            if (assign[v] == b) vertices[n++] = v;
        if (n >= 2) {
            split_sides(seed, epoch, b, choice[b], vertices, n, side);
            for (int32_t i = 0; i < n; ++i)
                if (side[i] == 1) assign[vertices[i]] = num_blocks;
        }
        num_blocks++;
        if (num_blocks >= target) break;
    }
    free(order); free(tmp); free(delta); free(choice); free(size); free(mapping); free(vertices); free(side);
    free(zeros); free(out_off); free(in_off); free(out_adj); free(in_adj);
    return num_blocks;
}
