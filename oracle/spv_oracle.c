/*
 * pp_oracle — CPU oracle for the SparsePauliVector evolution path.
 *
 * TEST INFRASTRUCTURE ONLY: a single-threaded C restatement of the reference
 * algorithm (PauliOperators.jl v3.1.0: src/spv_kernels.jl, src/spv_evolve.jl,
 * src/channels.jl, src/type_SparsePauliVector.jl, src/clip.jl,
 * src/truncation.jl).  Only the tests, the smoke check and the bench's
 * cpu_baseline leg load it; the product library never does.
 *
 * The generic body lives in spv_body.h and is instantiated twice:
 *   _w1 : UInt64 words  (N <= 64),  type_SparsePauliVector.jl:23-26
 *   _w2 : UInt128 words (N <= 128)
 * Keys cross this ABI as little-endian uint64 words per term (W words for z,
 * W words for x), i.e. a Julia Vector{UInt128} passes unchanged.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define OR_EXPTAB 257

/* spv_kernels.jl:33-42 (MergeFilter), plus host-computed exp tables for the
 * damped criteria: exptab_w[w] = exp((-alpha_w) * w). */
typedef struct {
    int64_t wmax, xwmax, mwmax;
    double thresh, alpha_w, dthresh_w, alpha_xw, dthresh_xw;
    double exptab_w[OR_EXPTAB];
    double exptab_xw[OR_EXPTAB];
} or_filter_t;

enum { OR_STRAT_COMPILED = 0, OR_STRAT_ADAPTIVE = 1 };
/* The boundary strategy plus the correction accumulators
 * (truncation.jl:262-310): energy != 0 (EnergyCorrection) measures
 * <ket|O|ket> over live + pending appends before the merge and over the
 * result after it, and adds after - before to energy_acc; variance != 0
 * (EnergyVarianceCorrection) measures energy and variance on the merged state
 * (spv_evolve.jl:246-270) and accumulates both. */
typedef struct {
    int kind;
    int64_t max_terms;
    double min_thresh;
    int energy;
    int variance;
    uint64_t ket[2];
    double energy_acc;
    double variance_acc;
} or_strategy_t;

/* Sort used at merge boundaries with more than one contributing rotation.
 * 0: the reference's unstable quicksort (spv_kernels.jl:202-247).
 * 1: a stable merge sort — equal keys keep their append order.  The two
 *    orders differ only in how equal keys are summed (a last-ulp effect);
 *    mode 1 is the order the device's stable radix sort produces. */
static int or_sort_mode = 0;

typedef struct {
    int64_t rotations, terms_swept, terms_created, merge_in, merge_out, early_merges;
} or_counters_t;

static void or_filter_fill_tables(or_filter_t *f) {
    for (int w = 0; w < OR_EXPTAB; w++) {
        f->exptab_w[w] = exp(-f->alpha_w * (double)w);
        f->exptab_xw[w] = exp(-f->alpha_xw * (double)w);
    }
}

/* spv_kernels.jl:44-45 NOFILTER */
static void or_filter_init_nofilter(or_filter_t *f) {
    f->wmax = INT64_MAX;
    f->xwmax = INT64_MAX;
    f->mwmax = INT64_MAX;
    f->thresh = -1.0;
    f->alpha_w = 0.0;
    f->dthresh_w = -1.0;
    f->alpha_xw = 0.0;
    f->dthresh_xw = -1.0;
    or_filter_fill_tables(f);
}

static inline int popc_w1(uint64_t v) { return __builtin_popcountll(v); }
static inline int popc_w2(u128 v) {
    return __builtin_popcountll((uint64_t)v) + __builtin_popcountll((uint64_t)(v >> 64));
}

#define WT uint64_t
#define SFX _w1
#include "spv_body.h"
#undef WT
#undef SFX
#undef F
#undef CAT
#undef CAT2

#define WT u128
#define SFX _w2
#include "spv_body.h"
#undef WT
#undef SFX

/* ------------------------------------------------------------------------ */
/* Julia's Float64^Int (base/math.jl pow_body): compensated power by squaring.
 * Julia's muladd is taken as fused, which is what a native-CPU JIT emits on
 * FMA hardware.  Used for the channel tables λ^k. */
static double two_mul_lo(double a, double b, double p) { return fma(a, b, -p); }

double or_julia_powi(double x, int64_t n) {
    if (n == 0) return 1.0;
    if (n < -4096 || n > 3 * 8192) return pow(x, (double)n); /* outside the table range */
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    if (n == 3) return x * x * x;
    if (n < 0) {
        double rx = 1.0 / x;
        if (n == -2) return rx * rx;
        if (isfinite(x)) xnlo = -fma(x, rx, -1.0) * rx;
        x = rx;
        n = -n;
    }
    while (n > 1) {
        if (n & 1) {
            double err = fma(y, xnlo, x * ynlo);
            double p = x * y;
            double plo = two_mul_lo(x, y, p);
            y = p;
            ynlo = plo + err;
        }
        double err = x * 2 * xnlo;
        double p = x * x;
        double plo = two_mul_lo(x, x, p);
        x = p;
        xnlo = plo + err;
        n >>= 1;
    }
    double err = fma(y, xnlo, x * ynlo);
    return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

/* ------------------------------------------------------------------------ */
/* Exported oracle API (ctypes).  Handles carry the word count. */

typedef struct {
    int W;
    void *v;
} or_handle;

static int words_for(int N) { return N <= 64 ? 1 : 2; }

static inline u128 ld2(const uint64_t *p) { return ((u128)p[1] << 64) | (u128)p[0]; }
static inline void st2(uint64_t *p, u128 v) {
    p[0] = (uint64_t)v;
    p[1] = (uint64_t)(v >> 64);
}

void or_filter_make(or_filter_t *f, int64_t wmax, int64_t xwmax, int64_t mwmax, double thresh,
                    double alpha_w, double dthresh_w, double alpha_xw, double dthresh_xw) {
    f->wmax = wmax;
    f->xwmax = xwmax;
    f->mwmax = mwmax;
    f->thresh = thresh;
    f->alpha_w = alpha_w;
    f->dthresh_w = dthresh_w;
    f->alpha_xw = alpha_xw;
    f->dthresh_xw = dthresh_xw;
    or_filter_fill_tables(f);
}

size_t or_filter_size(void) { return sizeof(or_filter_t); }

/* type_SparsePauliVector.jl:167-188: capacities, fill, construction sort. */
or_handle *or_create(int N, int64_t live_cap, int64_t app_cap) {
    if (N < 1 || N > 128 || live_cap < 1 || app_cap < 1) return NULL;
    or_handle *h = (or_handle *)malloc(sizeof(or_handle));
    h->W = words_for(N);
    h->v = h->W == 1 ? (void *)or_alloc_w1(N, live_cap, app_cap) : (void *)or_alloc_w2(N, live_cap, app_cap);
    return h;
}

void or_destroy(or_handle *h) {
    if (!h) return;
    if (h->W == 1) or_free_w1((or_spv_w1 *)h->v);
    else or_free_w2((or_spv_w2 *)h->v);
    free(h);
}

/* Load n terms (any order, no duplicates) and sort them as the constructor does. */
int or_load(or_handle *h, const uint64_t *z, const uint64_t *x, const double *c, int64_t n) {
    if (h->W == 1) {
        or_spv_w1 *v = (or_spv_w1 *)h->v;
        if (n > v->live_cap) grow_live_w1(v, n);
        for (int64_t i = 0; i < n; i++) { v->z[i] = z[i]; v->x[i] = x[i]; v->c[i] = c[i]; }
        v->n = n;
        v->an = 0;
        sort_live_w1(v);
    } else {
        or_spv_w2 *v = (or_spv_w2 *)h->v;
        if (n > v->live_cap) grow_live_w2(v, n);
        for (int64_t i = 0; i < n; i++) { v->z[i] = ld2(z + 2 * i); v->x[i] = ld2(x + 2 * i); v->c[i] = c[i]; }
        v->n = n;
        v->an = 0;
        sort_live_w2(v);
    }
    return 0;
}

int64_t or_size(or_handle *h) {
    return h->W == 1 ? ((or_spv_w1 *)h->v)->n : ((or_spv_w2 *)h->v)->n;
}

int64_t or_pending(or_handle *h) {
    return h->W == 1 ? ((or_spv_w1 *)h->v)->an : ((or_spv_w2 *)h->v)->an;
}

void or_capacities(or_handle *h, int64_t *live_cap, int64_t *app_cap) {
    if (h->W == 1) { *live_cap = ((or_spv_w1 *)h->v)->live_cap; *app_cap = ((or_spv_w1 *)h->v)->app_cap; }
    else { *live_cap = ((or_spv_w2 *)h->v)->live_cap; *app_cap = ((or_spv_w2 *)h->v)->app_cap; }
}

void or_get(or_handle *h, uint64_t *z, uint64_t *x, double *c) {
    if (h->W == 1) {
        or_spv_w1 *v = (or_spv_w1 *)h->v;
        for (int64_t i = 0; i < v->n; i++) { z[i] = v->z[i]; x[i] = v->x[i]; c[i] = v->c[i]; }
    } else {
        or_spv_w2 *v = (or_spv_w2 *)h->v;
        for (int64_t i = 0; i < v->n; i++) { st2(z + 2 * i, v->z[i]); st2(x + 2 * i, v->x[i]); c[i] = v->c[i]; }
    }
}

/* spv_evolve.jl:228-236 — single rotation, NOFILTER, unshuffle merge. */
int or_evolve_single(or_handle *h, const uint64_t *gz, const uint64_t *gx, double cosv, double sinv) {
    or_filter_t nof;
    or_filter_init_nofilter(&nof);
    int64_t n_in, n_out, created;
    if (h->W == 1) {
        or_spv_w1 *v = (or_spv_w1 *)h->v;
        int ng = popc_w1(gz[0] & gx[0]);
        if (v->n + 2 * v->an > v->app_cap) grow_append_w1(v, v->n + 2 * v->an);
        if (rotate_spv_w1(v, gz[0], gx[0], ng, cosv, sinv, &nof, &created)) return -2;
        int64_t m = gather_append_w1(v);
        unshuffle_ws_w1(v, m, gz[0], gx[0]);
        merge_spv_w1(v, m, &nof, &n_in, &n_out);
    } else {
        or_spv_w2 *v = (or_spv_w2 *)h->v;
        u128 z = ld2(gz), x = ld2(gx);
        int ng = popc_w2(z & x);
        if (v->n + 2 * v->an > v->app_cap) grow_append_w2(v, v->n + 2 * v->an);
        if (rotate_spv_w2(v, z, x, ng, cosv, sinv, &nof, &created)) return -2;
        int64_t m = gather_append_w2(v);
        unshuffle_ws_w2(v, m, z, x);
        merge_spv_w2(v, m, &nof, &n_in, &n_out);
    }
    return 0;
}

/* spv_evolve.jl:312-387 — windowed sequence driver.  gz/gx hold L generators
 * of W words each; cosv/sinv are the caller's cos/sin of each angle. */
static int or_evolve_strat(or_handle *h, const uint64_t *gz, const uint64_t *gx, const double *cosv,
                           const double *sinv, int64_t L, int64_t window, const or_filter_t *strict,
                           or_strategy_t *sp, const or_filter_t *local, or_counters_t *cnt);

/* variance_acc != NULL (with energy_ket) is EnergyVarianceCorrection. */
int or_evolve_ex(or_handle *h, const uint64_t *gz, const uint64_t *gx, const double *cosv,
                 const double *sinv, int64_t L, int64_t window, const or_filter_t *strict,
                 int strat_kind, int64_t max_terms, double min_thresh, const or_filter_t *local,
                 const uint64_t *energy_ket, double *energy_acc, double *variance_acc, or_counters_t *cnt) {
    if (window < 1) return -1;
    if (L == 0) return 0;
    or_strategy_t s;
    memset(&s, 0, sizeof(s));
    s.kind = strat_kind;
    s.max_terms = max_terms;
    s.min_thresh = min_thresh;
    if (energy_ket) {
        s.energy = 1;
        s.variance = variance_acc != NULL;
        s.ket[0] = energy_ket[0];
        s.ket[1] = h->W == 2 ? energy_ket[1] : 0;
        s.energy_acc = energy_acc ? *energy_acc : 0.0;
        s.variance_acc = variance_acc ? *variance_acc : 0.0;
    }
    int rc = or_evolve_strat(h, gz, gx, cosv, sinv, L, window, strict, &s, local, cnt);
    if (energy_acc) *energy_acc = s.energy_acc;
    if (variance_acc) *variance_acc = s.variance_acc;
    return rc;
}

/* variance(v, ψ) (statistics.jl:6-22) for a computational-basis ket. */
double or_variance(or_handle *h, const uint64_t *ket) {
    if (h->W == 1) return variance_w1((or_spv_w1 *)h->v, (uint64_t)ket[0]);
    return variance_w2((or_spv_w2 *)h->v, ld2(ket));
}

void or_set_sort_mode(int mode) { or_sort_mode = mode; }

int or_evolve(or_handle *h, const uint64_t *gz, const uint64_t *gx, const double *cosv,
              const double *sinv, int64_t L, int64_t window, const or_filter_t *strict,
              int strat_kind, int64_t max_terms, double min_thresh, const or_filter_t *local,
              or_counters_t *cnt) {
    return or_evolve_ex(h, gz, gx, cosv, sinv, L, window, strict, strat_kind, max_terms, min_thresh,
                        local, NULL, NULL, NULL, cnt);
}

static int or_evolve_strat(or_handle *h, const uint64_t *gz, const uint64_t *gx, const double *cosv,
                           const double *sinv, int64_t L, int64_t window, const or_filter_t *strict,
                           or_strategy_t *sp, const or_filter_t *local, or_counters_t *cnt) {
    const int strat_kind = sp->kind;
    or_strategy_t s = *sp;
    or_filter_t nof;
    or_filter_init_nofilter(&nof);
    const or_filter_t *f = strat_kind == OR_STRAT_COMPILED ? strict : &nof;
    if (h->W == 1) {
        int rc1 = evolve_seq_w1((or_spv_w1 *)h->v, gz, gx, cosv, sinv, L, window, f, &s, local, cnt);
        *sp = s;
        return rc1;
    }
    u128 *z = (u128 *)malloc((size_t)L * sizeof(u128));
    u128 *x = (u128 *)malloc((size_t)L * sizeof(u128));
    for (int64_t i = 0; i < L; i++) { z[i] = ld2(gz + 2 * i); x[i] = ld2(gx + 2 * i); }
    int rc = evolve_seq_w2((or_spv_w2 *)h->v, z, x, cosv, sinv, L, window, f, &s, local, cnt);
    free(z);
    free(x);
    *sp = s;
    return rc;
}

/* spv_kernels.jl:395-408 via the clip family (spv_evolve.jl:13-26). */
void or_compact(or_handle *h, const or_filter_t *f) {
    if (h->W == 1) compact_w1((or_spv_w1 *)h->v, f);
    else compact_w2((or_spv_w2 *)h->v, f);
}

void or_adaptive(or_handle *h, int64_t max_terms, double min_thresh) {
    if (h->W == 1) apply_adaptive_w1((or_spv_w1 *)h->v, max_terms, min_thresh);
    else apply_adaptive_w2((or_spv_w2 *)h->v, max_terms, min_thresh);
}

/* channels.jl:203-265 — per-term popcount damping.
 * kind: 0 depolarizing (z|x), 1 dephasing (x), 2 bit flip (z), 3 bit-phase
 * flip (x^z), 4 general Pauli channel (lamX^nX * lamY^nY * lamZ^nZ).
 * The λ values are the caller's; powers use Julia's Float64^Int. */
void or_channel(or_handle *h, int kind, const uint64_t *mask, double lam, double lamX,
                double lamY, double lamZ) {
    if (h->W == 1) {
        or_spv_w1 *v = (or_spv_w1 *)h->v;
        uint64_t M = mask[0];
        for (int64_t i = 0; i < v->n; i++) {
            uint64_t z = v->z[i], x = v->x[i];
            if (kind == 4) {
                uint64_t xM = x & M, zM = z & M;
                int nX = popc_w1(xM & ~z), nY = popc_w1(xM & z), nZ = popc_w1(zM & ~x);
                double fac = or_julia_powi(lamX, nX) * or_julia_powi(lamY, nY);
                fac = fac * or_julia_powi(lamZ, nZ);
                v->c[i] = v->c[i] * fac;
            } else {
                uint64_t sel = kind == 0 ? (z | x) : kind == 1 ? x : kind == 2 ? z : (x ^ z);
                v->c[i] = v->c[i] * or_julia_powi(lam, popc_w1(sel & M));
            }
        }
    } else {
        or_spv_w2 *v = (or_spv_w2 *)h->v;
        u128 M = ld2(mask);
        for (int64_t i = 0; i < v->n; i++) {
            u128 z = v->z[i], x = v->x[i];
            if (kind == 4) {
                u128 xM = x & M, zM = z & M;
                int nX = popc_w2(xM & ~z), nY = popc_w2(xM & z), nZ = popc_w2(zM & ~x);
                double fac = or_julia_powi(lamX, nX) * or_julia_powi(lamY, nY);
                fac = fac * or_julia_powi(lamZ, nZ);
                v->c[i] = v->c[i] * fac;
            } else {
                u128 sel = kind == 0 ? (z | x) : kind == 1 ? x : kind == 2 ? z : (x ^ z);
                v->c[i] = v->c[i] * or_julia_powi(lam, popc_w2(sel & M));
            }
        }
    }
}

/* spv_evolve.jl:133-134 -> spv_kernels.jl:441-452 */
double or_expectation(or_handle *h, const uint64_t *ket) {
    if (h->W == 1) return expectation_w1((or_spv_w1 *)h->v, ket[0]);
    return expectation_w2((or_spv_w2 *)h->v, ld2(ket));
}

/* spv_kernels.jl:51-62, exported for the weight-measure known-answer tests. */
int or_majorana_weight(int W, const uint64_t *z, const uint64_t *x) {
    if (W == 1) return majorana_bits_w1(z[0], x[0]);
    return majorana_bits_w2(ld2(z), ld2(x));
}

/* spv_kernels.jl:64-79, exported for the filter known-answer tests. */
int or_should_drop(int W, const or_filter_t *f, const uint64_t *z, const uint64_t *x, double absc) {
    if (W == 1) return should_drop_w1(f, z[0], x[0], absc);
    return should_drop_w2(f, ld2(z), ld2(x), absc);
}

/* type_SparsePauliVector.jl:263-278 (key-order part of check_spv). */
int or_check(or_handle *h) {
    if (h->W == 1) {
        or_spv_w1 *v = (or_spv_w1 *)h->v;
        for (int64_t i = 1; i < v->n; i++)
            if (!key_lt_w1(v->z[i - 1], v->x[i - 1], v->z[i], v->x[i])) return 0;
    } else {
        or_spv_w2 *v = (or_spv_w2 *)h->v;
        for (int64_t i = 1; i < v->n; i++)
            if (!key_lt_w2(v->z[i - 1], v->x[i - 1], v->z[i], v->x[i])) return 0;
    }
    return 1;
}
