/*
 * pp_oracle — CPU oracle for the SparsePauliVector evolution path.
 *
 * TEST INFRASTRUCTURE ONLY: a single-threaded C restatement of the reference
 * algorithm (PauliOperators.jl v3.1.0: src/spv_kernels.jl, src/spv_evolve.jl,
 * src/channels.jl, src/type_SparsePauliVector.jl, src/clip.jl,
 * src/truncation.jl).  Only the tests, the smoke check and the bench's
 * cpu_baseline leg load it; the product library never does.
 *
 * The generic body lives in spv_body.h and is instantiated four times:
 *   _w1, _w1c : UInt64 words  (N <= 64),  type_SparsePauliVector.jl:23-26
 *   _w2, _w2c : UInt128 words (N <= 128)
 * with Float64 coefficients (_w1, _w2) or ComplexF64 (_w1c, _w2c).
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
static inline int popc_w1c(uint64_t v) { return popc_w1(v); }
static inline int popc_w2c(u128 v) { return popc_w2(v); }

double or_julia_powi(double x, int64_t n);

/* ------------------------------------------------------------------------ */
/* ComplexF64 arithmetic as Julia performs it (base/complex.jl, base/math.jl). */
typedef struct {
    double re, im;
} or_cplx;

/* hypot(x, y) = abs(::ComplexF64) (base/math.jl, the branch taken on CPUs
 * with a hardware fma: Borges' corrected hypot; muladd taken as fused). */
double or_hypot(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    if (isinf(ax) || isinf(ay)) return INFINITY;
    if (ay > ax) {
        const double t = ax;
        ax = ay;
        ay = t;
    }
    if (ay <= ax * sqrt(0x1p-53)) return ax;            /* sqrt(eps(Float64) / 2) */
    double scale = 0x1p-52 * sqrt(0x1p-1022);            /* eps * sqrt(floatmin) */
    if (ax > sqrt(0x1.fffffffffffffp+1023 / 2)) {        /* sqrt(floatmax / 2) */
        ax = ax * scale;
        ay = ay * scale;
        scale = 1.0 / scale;
    } else if (ay < sqrt(0x1p-1022)) {                   /* sqrt(floatmin) */
        ax = ax / scale;
        ay = ay / scale;
    } else {
        scale = 1.0;
    }
    double h = sqrt(fma(ax, ax, ay * ay));
    const double hsq = h * h;
    const double axsq = ax * ax;
    h -= (fma(-ay, ay, hsq - axsq) + fma(h, h, -hsq) - fma(ax, ax, -axsq)) / (2 * h);
    return h * scale;
}

/* T(k - 2) * sinθ = Complex(a, 0.0 * sinθ) times c, Complex * Complex
 * (spv_kernels.jl:163): (a re - b im, a im + b re). */
static inline or_cplx cx_rot(double a, double s, or_cplx c) {
    const double b = 0.0 * s;
    or_cplx r = {a * c.re - b * c.im, a * c.im + b * c.re};
    return r;
}
static inline or_cplx cx_scale(or_cplx c, double s) {  /* Complex * Real */
    or_cplx r = {c.re * s, c.im * s};
    return r;
}
static inline or_cplx cx_add(or_cplx a, or_cplx b) {
    or_cplx r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline double cx_abs2(or_cplx c) { return c.re * c.re + c.im * c.im; }
static inline or_cplx cx_ld(const double *p, int64_t i) {
    or_cplx r = {p[2 * i], p[2 * i + 1]};
    return r;
}
static inline void cx_st(double *p, int64_t i, or_cplx c) {
    p[2 * i] = c.re;
    p[2 * i + 1] = c.im;
}

/* Float64 */
#define CT double
#define C_ABS(c) fabs(c)
#define C_ABS2(c) ((c) * (c))
#define C_ROT(a, s, c) ((a) * (c))
#define C_SCALE(c, s) ((c) * (s))
#define C_ADD(a, b) ((a) + (b))
#define C_RE(c) (c)
#define C_IM(c) 0.0
#define C_ZERO 0.0
#define C_LD(p, i) ((p)[i])
#define C_ST(p, i, v) ((p)[i] = (v))

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
#undef F
#undef CAT
#undef CAT2

#undef CT
#undef C_ABS
#undef C_ABS2
#undef C_ROT
#undef C_SCALE
#undef C_ADD
#undef C_RE
#undef C_IM
#undef C_ZERO
#undef C_LD
#undef C_ST

/* ComplexF64 */
static const or_cplx or_czero = {0.0, 0.0};
#define CT or_cplx
#define C_ABS(c) or_hypot((c).re, (c).im)
#define C_ABS2(c) cx_abs2(c)
#define C_ROT(a, s, c) cx_rot((a), (s), (c))
#define C_SCALE(c, s) cx_scale((c), (s))
#define C_ADD(a, b) cx_add((a), (b))
#define C_RE(c) ((c).re)
#define C_IM(c) ((c).im)
#define C_ZERO or_czero
#define C_LD(p, i) cx_ld((p), (i))
#define C_ST(p, i, v) cx_st((p), (i), (v))

#define WT uint64_t
#define SFX _w1c
#include "spv_body.h"
#undef WT
#undef SFX
#undef F
#undef CAT
#undef CAT2

#define WT u128
#define SFX _w2c
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
/* Exported oracle API (ctypes).  Handles carry the word count and the
 * coefficient type; OR_CALL picks the instance. */

typedef struct {
    int W;
    int cplx;
    void *v;
} or_handle;

#define OR_CALL(h, fn, ...)                                                              \
    ((h)->W == 1 ? ((h)->cplx ? fn##_w1c((h)->v, ##__VA_ARGS__) : fn##_w1((h)->v, ##__VA_ARGS__)) \
                 : ((h)->cplx ? fn##_w2c((h)->v, ##__VA_ARGS__) : fn##_w2((h)->v, ##__VA_ARGS__)))

static int words_for(int N) { return N <= 64 ? 1 : 2; }

static inline u128 ld2(const uint64_t *p) { return ((u128)p[1] << 64) | (u128)p[0]; }

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

/* SparsePauliVector(N, T) with T = Float64 (cplx = 0) or ComplexF64 (cplx = 1):
 * type_SparsePauliVector.jl:138-141,167-188. */
or_handle *or_create_typed(int N, int64_t live_cap, int64_t app_cap, int cplx) {
    if (N < 1 || N > 128 || live_cap < 1 || app_cap < 1) return NULL;
    or_handle *h = (or_handle *)malloc(sizeof(or_handle));
    h->W = words_for(N);
    h->cplx = cplx ? 1 : 0;
    if (h->W == 1) h->v = h->cplx ? api_create_w1c(N, live_cap, app_cap) : api_create_w1(N, live_cap, app_cap);
    else h->v = h->cplx ? api_create_w2c(N, live_cap, app_cap) : api_create_w2(N, live_cap, app_cap);
    return h;
}

or_handle *or_create(int N, int64_t live_cap, int64_t app_cap) { return or_create_typed(N, live_cap, app_cap, 0); }

int or_is_complex(or_handle *h) { return h->cplx; }

void or_destroy(or_handle *h) {
    if (!h) return;
    OR_CALL(h, api_destroy);
    free(h);
}

/* Load n terms (any order, no duplicates) and sort them as the constructor does. */
int or_load(or_handle *h, const uint64_t *z, const uint64_t *x, const double *c, int64_t n) {
    return OR_CALL(h, api_load, z, x, c, n);
}

int64_t or_size(or_handle *h) { return OR_CALL(h, api_size); }
int64_t or_pending(or_handle *h) { return OR_CALL(h, api_pending); }
void or_capacities(or_handle *h, int64_t *live_cap, int64_t *app_cap) { OR_CALL(h, api_capacities, live_cap, app_cap); }
void or_get(or_handle *h, uint64_t *z, uint64_t *x, double *c) { OR_CALL(h, api_get, z, x, c); }

int or_evolve_single(or_handle *h, const uint64_t *gz, const uint64_t *gx, double cosv, double sinv) {
    return OR_CALL(h, api_evolve_single, gz, gx, cosv, sinv);
}

/* spv_evolve.jl:312-387 — windowed sequence driver.  gz/gx hold L generators
 * of W words each; cosv/sinv are the caller's cos/sin of each angle.
 * variance_acc != NULL (with energy_ket) is EnergyVarianceCorrection. */
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
    or_filter_t nof;
    or_filter_init_nofilter(&nof);
    const or_filter_t *f = strat_kind == OR_STRAT_COMPILED ? strict : &nof;
    int rc = OR_CALL(h, api_evolve_seq, gz, gx, cosv, sinv, L, window, f, &s, local, cnt);
    if (energy_acc) *energy_acc = s.energy_acc;
    if (variance_acc) *variance_acc = s.variance_acc;
    return rc;
}

/* variance(v, ψ) (statistics.jl:6-22) for a computational-basis ket. */
double or_variance(or_handle *h, const uint64_t *ket) { return OR_CALL(h, api_variance, ket); }

void or_set_sort_mode(int mode) { or_sort_mode = mode; }

int or_evolve(or_handle *h, const uint64_t *gz, const uint64_t *gx, const double *cosv,
              const double *sinv, int64_t L, int64_t window, const or_filter_t *strict,
              int strat_kind, int64_t max_terms, double min_thresh, const or_filter_t *local,
              or_counters_t *cnt) {
    return or_evolve_ex(h, gz, gx, cosv, sinv, L, window, strict, strat_kind, max_terms, min_thresh,
                        local, NULL, NULL, NULL, cnt);
}

/* spv_kernels.jl:395-408 via the clip family (spv_evolve.jl:13-26). */
void or_compact(or_handle *h, const or_filter_t *f) { OR_CALL(h, api_compact, f); }

void or_adaptive(or_handle *h, int64_t max_terms, double min_thresh) { OR_CALL(h, api_adaptive, max_terms, min_thresh); }

/* channels.jl:203-265 — per-term popcount damping; the λ values are the
 * caller's, powers use Julia's Float64^Int. */
void or_channel(or_handle *h, int kind, const uint64_t *mask, double lam, double lamX,
                double lamY, double lamZ) {
    OR_CALL(h, api_channel, kind, mask, lam, lamX, lamY, lamZ);
}

/* spv_evolve.jl:133-134 -> spv_kernels.jl:441-452 (the real part). */
double or_expectation(or_handle *h, const uint64_t *ket) {
    double out[2];
    OR_CALL(h, api_expectation, ket, out);
    return out[0];
}

/* The same, out[0] real and out[1] imaginary part. */
void or_expectation_c(or_handle *h, const uint64_t *ket, double *out) { OR_CALL(h, api_expectation, ket, out); }

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
int or_check(or_handle *h) { return OR_CALL(h, api_check); }
