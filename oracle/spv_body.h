/*
 * CPU oracle body, instantiated once per packed word type.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded restatement of the
 * reference SparsePauliVector engine (PauliOperators.jl v3.1.0).  It is the
 * checker for the HIP kernels and the CPU baseline timed beside them; the
 * product library never links or calls it.
 *
 * Instantiation: define WT (uint64_t or u128), the coefficient type CT with
 * its arithmetic (C_ABS, C_ABS2, C_ROT, C_SCALE, C_ADD, C_RE, C_IM, C_ZERO,
 * C_LD, C_ST: see spv_oracle.c) and SFX (a name suffix), then include this
 * file.  Every function cites the reference file:line it restates.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

typedef struct {
    WT z, x;
    CT c;
} F(or_trip);

/* type_SparsePauliVector.jl:69-84 — live / append / scratch SoA buffers plus
 * the AoS sort workspace and its ping-pong twin. */
typedef struct {
    int N;
    WT *z, *x;
    CT *c;
    int64_t n, live_cap;
    WT *az, *ax;
    CT *ac;
    int64_t an, app_cap;
    WT *sz, *sx;
    CT *sc;
    F(or_trip) *ws;
    int64_t ws_cap;
    F(or_trip) *ws2;
    int64_t ws2_cap;
} F(or_spv);

/* spv_kernels.jl:14-16 — lexicographic (z, x) order, z-major. */
static inline int F(key_lt)(WT az, WT ax, WT bz, WT bx) {
    return (az < bz) | ((az == bz) & (ax < bx));
}
static inline int F(trip_lt)(const F(or_trip) *a, const F(or_trip) *b) {
    return F(key_lt)(a->z, a->x, b->z, b->x);
}

/* spv_kernels.jl:51-62 — branchless suffix-parity Majorana weight. */
static inline int F(majorana_bits)(WT z, WT x) {
    WT zonly = z & ~x;
    WT S = x;
    for (int shift = 1; shift < 8 * (int)sizeof(WT); shift <<= 1) S ^= S >> shift;
    WT ctrl = ~(S ^ x);
    return F(popc)(x) + 2 * F(popc)(zonly & ctrl) + 2 * F(popc)(~(z | x) & ~ctrl);
}

/* spv_kernels.jl:64-79 — per-term drop predicate.  The exp() factors come from
 * the filter's host-computed tables (exp(-alpha*w) for w = 0..256). */
static inline int F(should_drop)(const or_filter_t *f, WT z, WT x, double absc) {
    if (absc <= f->thresh) return 1;
    int w = F(popc)(z | x);
    if ((int64_t)w > f->wmax) return 1;
    int xw = F(popc)(x);
    if ((int64_t)xw > f->xwmax) return 1;
    if (f->mwmax != INT64_MAX) {
        if ((int64_t)F(majorana_bits)(z, x) > f->mwmax) return 1;
    }
    if (f->dthresh_w >= 0.0) {
        if (absc * f->exptab_w[w] <= f->dthresh_w) return 1;
    }
    if (f->dthresh_xw >= 0.0) {
        if (absc * f->exptab_xw[xw] <= f->dthresh_xw) return 1;
    }
    return 0;
}

/* type_SparsePauliVector.jl:94-104 — allocate all buffers. */
static F(or_spv) *F(or_alloc)(int N, int64_t live_cap, int64_t app_cap) {
    F(or_spv) *v = (F(or_spv) *)calloc(1, sizeof(F(or_spv)));
    v->N = N;
    v->live_cap = live_cap;
    v->app_cap = app_cap;
    v->z = (WT *)calloc((size_t)live_cap, sizeof(WT));
    v->x = (WT *)calloc((size_t)live_cap, sizeof(WT));
    v->c = (CT *)calloc((size_t)live_cap, sizeof(CT));
    v->sz = (WT *)calloc((size_t)live_cap, sizeof(WT));
    v->sx = (WT *)calloc((size_t)live_cap, sizeof(WT));
    v->sc = (CT *)calloc((size_t)live_cap, sizeof(CT));
    v->az = (WT *)calloc((size_t)app_cap, sizeof(WT));
    v->ax = (WT *)calloc((size_t)app_cap, sizeof(WT));
    v->ac = (CT *)calloc((size_t)app_cap, sizeof(CT));
    v->ws_cap = live_cap > app_cap ? live_cap : app_cap;
    v->ws = (F(or_trip) *)malloc((size_t)v->ws_cap * sizeof(F(or_trip)));
    v->ws2 = NULL;
    v->ws2_cap = 0;
    return v;
}

static void F(or_free)(F(or_spv) *v) {
    free(v->z); free(v->x); free(v->c);
    free(v->sz); free(v->sx); free(v->sc);
    free(v->az); free(v->ax); free(v->ac);
    free(v->ws); free(v->ws2);
    free(v);
}

static void *F(xrealloc)(void *p, size_t n) {
    void *q = realloc(p, n);
    if (!q && n) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return q;
}

/* type_SparsePauliVector.jl:107-117 — chunked doubling of live + scratch. */
static void F(grow_live)(F(or_spv) *v, int64_t need) {
    int64_t newcap = 2 * v->live_cap > need ? 2 * v->live_cap : need;
    v->z = (WT *)F(xrealloc)(v->z, (size_t)newcap * sizeof(WT));
    v->x = (WT *)F(xrealloc)(v->x, (size_t)newcap * sizeof(WT));
    v->c = (CT *)F(xrealloc)(v->c, (size_t)newcap * sizeof(CT));
    v->sz = (WT *)F(xrealloc)(v->sz, (size_t)newcap * sizeof(WT));
    v->sx = (WT *)F(xrealloc)(v->sx, (size_t)newcap * sizeof(WT));
    v->sc = (CT *)F(xrealloc)(v->sc, (size_t)newcap * sizeof(CT));
    v->live_cap = newcap;
    if (v->ws_cap < newcap) {
        v->ws = (F(or_trip) *)F(xrealloc)(v->ws, (size_t)newcap * sizeof(F(or_trip)));
        v->ws_cap = newcap;
    }
}

/* type_SparsePauliVector.jl:119-126 — chunked doubling of the append buffer. */
static void F(grow_append)(F(or_spv) *v, int64_t need) {
    int64_t newcap = 2 * v->app_cap > need ? 2 * v->app_cap : need;
    v->az = (WT *)F(xrealloc)(v->az, (size_t)newcap * sizeof(WT));
    v->ax = (WT *)F(xrealloc)(v->ax, (size_t)newcap * sizeof(WT));
    v->ac = (CT *)F(xrealloc)(v->ac, (size_t)newcap * sizeof(CT));
    v->app_cap = newcap;
    if (v->ws_cap < newcap) {
        v->ws = (F(or_trip) *)F(xrealloc)(v->ws, (size_t)newcap * sizeof(F(or_trip)));
        v->ws_cap = newcap;
    }
}

/* spv_kernels.jl:146-177 — rotation sweep of terms [lo, hi) (0-based) of one
 * buffer set, appending sin-branch children at *cur. */
static void F(rotate_range)(WT *z, WT *x, CT *c, int64_t lo, int64_t hi,
                            WT gz, WT gx, int ng, double cosv, double sinv,
                            WT *dz, WT *dx, CT *dc, int64_t *cur, int64_t cap_end,
                            const or_filter_t *f, int64_t *created, int *overflowed) {
    for (int64_t i = lo; i < hi; i++) {
        WT zi = z[i], xi = x[i];
        int m1 = F(popc)(gx & zi);
        int m2 = F(popc)(gz & xi);
        if (((m1 - m2) & 1) == 0) continue;
        WT zp = gz ^ zi;
        WT xp = gx ^ xi;
        int k = (F(popc)(zp & xp) - ng - F(popc)(zi & xi) + 2 * m1) & 3;
        CT cnew = C_ROT((double)(k - 2) * sinv, sinv, c[i]);
        c[i] = C_SCALE(c[i], cosv);
        if (F(should_drop)(f, zp, xp, C_ABS(cnew))) continue;
        if (*cur >= cap_end) {
            *overflowed = 1;
            continue;
        }
        dz[*cur] = zp;
        dx[*cur] = xp;
        dc[*cur] = cnew;
        *cur += 1;
        *created += 1;
    }
}

/* spv_kernels.jl:187-193 */
static int64_t F(gather_append)(F(or_spv) *v) {
    int64_t m = v->an;
    for (int64_t i = 0; i < m; i++) {
        v->ws[i].z = v->az[i];
        v->ws[i].x = v->ax[i];
        v->ws[i].c = v->ac[i];
    }
    return m;
}

/* spv_kernels.jl:202-247 — in-place quicksort, median of 3, insertion sort
 * below 24 elements; lo, hi inclusive (0-based). */
static void F(sort_ws)(F(or_trip) *ws, int64_t lo, int64_t hi) {
    F(or_trip) t;
#define SWAPT(a, b) do { t = ws[a]; ws[a] = ws[b]; ws[b] = t; } while (0)
    while (hi - lo >= 24) {
        int64_t mid = (int64_t)(((uint64_t)lo + (uint64_t)hi) >> 1);
        if (F(trip_lt)(&ws[mid], &ws[lo])) SWAPT(mid, lo);
        if (F(trip_lt)(&ws[hi], &ws[mid])) {
            SWAPT(hi, mid);
            if (F(trip_lt)(&ws[mid], &ws[lo])) SWAPT(mid, lo);
        }
        F(or_trip) pivot = ws[mid];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (F(trip_lt)(&ws[i], &pivot)) i++;
            while (F(trip_lt)(&pivot, &ws[j])) j--;
            if (i <= j) {
                SWAPT(i, j);
                i++;
                j--;
            }
        }
        if (j - lo < hi - i) {
            F(sort_ws)(ws, lo, j);
            lo = i;
        } else {
            F(sort_ws)(ws, i, hi);
            hi = j;
        }
    }
#undef SWAPT
    for (int64_t k = lo + 1; k <= hi; k++) {
        F(or_trip) val = ws[k];
        int64_t m = k - 1;
        while (m >= lo && F(trip_lt)(&val, &ws[m])) {
            ws[m + 1] = ws[m];
            m--;
        }
        ws[m + 1] = val;
    }
}

/* Stable alternative to sort_ws (or_sort_mode == 1): top-down merge sort of
 * ws[lo..hi] (inclusive) through tmp; equal keys keep their order.  Not the
 * reference's order for duplicates: it is the device radix sort's, so that
 * the two can be compared bit for bit. */
static void F(msort_ws)(F(or_trip) *ws, F(or_trip) *tmp, int64_t lo, int64_t hi) {
    if (hi - lo < 1) return;
    int64_t mid = lo + (hi - lo) / 2;
    F(msort_ws)(ws, tmp, lo, mid);
    F(msort_ws)(ws, tmp, mid + 1, hi);
    int64_t i = lo, j = mid + 1, k = lo;
    while (i <= mid && j <= hi) {
        if (F(trip_lt)(&ws[j], &ws[i])) tmp[k++] = ws[j++];
        else tmp[k++] = ws[i++];
    }
    while (i <= mid) tmp[k++] = ws[i++];
    while (j <= hi) tmp[k++] = ws[j++];
    memcpy(ws + lo, tmp + lo, (size_t)(hi - lo + 1) * sizeof(F(or_trip)));
}

static void F(sort_pending)(F(or_trip) *ws, int64_t m) {
    if (m <= 1) return;
    if (or_sort_mode == 1) {
        F(or_trip) *tmp = (F(or_trip) *)malloc((size_t)m * sizeof(F(or_trip)));
        F(msort_ws)(ws, tmp, 0, m - 1);
        free(tmp);
    } else {
        F(sort_ws)(ws, 0, m - 1);
    }
}

/* spv_kernels.jl:253-286 — one stable block-swap pass of the unshuffle cascade
 * for bit b of word `which` (1 = z, 2 = x). */
static void F(unshuffle_pass)(const F(or_trip) *src, F(or_trip) *dst, int64_t m, int b, int which) {
    /* bits above b; Julia's shift by the full width yields 0, so hb = 0 there */
    WT hb = (b + 1 >= 8 * (int)sizeof(WT)) ? (WT)0 : ~((((WT)1) << (b + 1)) - (WT)1);
    int64_t i = 0;
    while (i < m) {
        WT hp1 = which == 1 ? (src[i].z & hb) : src[i].z;
        WT hp2 = which == 1 ? (WT)0 : (src[i].x & hb);
        int64_t j = i;
        while (j + 1 < m) {
            WT q1 = which == 1 ? (src[j + 1].z & hb) : src[j + 1].z;
            WT q2 = which == 1 ? (WT)0 : (src[j + 1].x & hb);
            if (q1 == hp1 && q2 == hp2) j++;
            else break;
        }
        int64_t f = i;
        while (f <= j) {
            WT w = which == 1 ? src[f].z : src[f].x;
            if (((w >> b) & (WT)1) == (WT)1) f++;
            else break;
        }
        int64_t k = i;
        for (int64_t s = f; s <= j; s++) dst[k++] = src[s];
        for (int64_t s = i; s < f; s++) dst[k++] = src[s];
        i = j + 1;
    }
}

/* spv_kernels.jl:300-320 */
static void F(unshuffle_ws)(F(or_spv) *v, int64_t m, WT gz, WT gx) {
    if (m <= 1) return;
    if (F(popc)(gz) + F(popc)(gx) == 0) return;
    if (v->ws2_cap < m) {
        v->ws2 = (F(or_trip) *)F(xrealloc)(v->ws2, (size_t)v->ws_cap * sizeof(F(or_trip)));
        v->ws2_cap = v->ws_cap;
    }
    F(or_trip) *src = v->ws, *dst = v->ws2, *tmp;
    int flips = 0;
    for (int b = 8 * (int)sizeof(WT) - 1; b >= 0; b--) {
        if (((gz >> b) & (WT)1) != (WT)1) continue;
        F(unshuffle_pass)(src, dst, m, b, 1);
        tmp = src; src = dst; dst = tmp;
        flips++;
    }
    for (int b = 8 * (int)sizeof(WT) - 1; b >= 0; b--) {
        if (((gx >> b) & (WT)1) != (WT)1) continue;
        F(unshuffle_pass)(src, dst, m, b, 2);
        tmp = src; src = dst; dst = tmp;
        flips++;
    }
    if (flips & 1) memcpy(v->ws, v->ws2, (size_t)m * sizeof(F(or_trip)));
}

/* spv_kernels.jl:326-363 — two-pointer merge of live + m sorted workspace
 * triples into scratch; ties take the live term first, then every equal
 * workspace triple in order.  Returns 1 on overflow (live untouched). */
static int F(try_merge)(F(or_spv) *v, int64_t m, const or_filter_t *f, int64_t *out_n) {
    int64_t n = v->n, cap = v->live_cap, out = 0, i = 0, j = 0;
    WT *z = v->z, *x = v->x, *sz = v->sz, *sx = v->sx;
    CT *c = v->c, *sc = v->sc;
    F(or_trip) *ws = v->ws;
    while (i < n || j < m) {
        WT kz, kx;
        CT acc;
        if (j >= m || (i < n && !F(key_lt)(ws[j].z, ws[j].x, z[i], x[i]))) {
            kz = z[i]; kx = x[i]; acc = c[i];
            i++;
        } else {
            kz = ws[j].z; kx = ws[j].x; acc = ws[j].c;
            j++;
        }
        while (j < m && ws[j].z == kz && ws[j].x == kx) {
            acc = C_ADD(acc, ws[j].c);
            j++;
        }
        if (F(should_drop)(f, kz, kx, C_ABS(acc))) continue;
        out++;
        if (out > cap) { *out_n = out; return 1; }
        sz[out - 1] = kz;
        sx[out - 1] = kx;
        sc[out - 1] = acc;
    }
    v->z = sz; v->sz = z;
    v->x = sx; v->sx = x;
    v->c = sc; v->sc = c;
    v->n = out;
    *out_n = out;
    return 0;
}

/* spv_kernels.jl:377-386 */
static void F(merge_spv)(F(or_spv) *v, int64_t m, const or_filter_t *f, int64_t *n_in, int64_t *n_out) {
    *n_in = v->n + m;
    int64_t out;
    while (F(try_merge)(v, m, f, &out)) F(grow_live)(v, v->n + m);
    v->an = 0;
    *n_out = out;
}

/* spv_kernels.jl:395-408 — order-preserving in-place compaction. */
static void F(compact)(F(or_spv) *v, const or_filter_t *f) {
    int64_t out = 0;
    for (int64_t i = 0; i < v->n; i++) {
        if (F(should_drop)(f, v->z[i], v->x[i], C_ABS(v->c[i]))) continue;
        if (out != i) {
            v->z[out] = v->z[i];
            v->x[out] = v->x[i];
            v->c[out] = v->c[i];
        }
        out++;
    }
    v->n = out;
}

/* spv_kernels.jl:441-452 — sequential ⟨k|O|k⟩ over live then appends;
 * (1 - 2 parity) * c is Int * T (componentwise for ComplexF64). */
static CT F(expectation)(const F(or_spv) *v, WT kv) {
    CT acc = C_ZERO;
    for (int64_t i = 0; i < v->n; i++) {
        if (v->x[i] != (WT)0) continue;
        acc = C_ADD(acc, C_SCALE(v->c[i], (double)(1 - 2 * (F(popc)(v->z[i] & kv) & 1))));
    }
    for (int64_t i = 0; i < v->an; i++) {
        if (v->ax[i] != (WT)0) continue;
        acc = C_ADD(acc, C_SCALE(v->ac[i], (double)(1 - 2 * (F(popc)(v->az[i] & kv) & 1))));
    }
    return acc;
}

/* type_SparsePauliVector.jl:195-204 — construction-time quicksort of live. */
static void F(sort_live)(F(or_spv) *v) {
    if (v->ws_cap < v->n) {
        v->ws = (F(or_trip) *)F(xrealloc)(v->ws, (size_t)v->n * sizeof(F(or_trip)));
        v->ws_cap = v->n;
    }
    for (int64_t i = 0; i < v->n; i++) {
        v->ws[i].z = v->z[i];
        v->ws[i].x = v->x[i];
        v->ws[i].c = v->c[i];
    }
    if (v->n > 0) F(sort_ws)(v->ws, 0, v->n - 1);
    for (int64_t i = 0; i < v->n; i++) {
        v->z[i] = v->ws[i].z;
        v->x[i] = v->ws[i].x;
        v->c[i] = v->ws[i].c;
    }
}

/* spv_evolve.jl:207-218 — rotate live, then the pending appends present at
 * entry; children land in the append buffer. */
static int F(rotate_spv)(F(or_spv) *v, WT gz, WT gx, int ng, double cosv, double sinv,
                         const or_filter_t *f, int64_t *created) {
    int64_t hi = v->an;
    int64_t cur = v->an;
    int ovf = 0;
    *created = 0;
    F(rotate_range)(v->z, v->x, v->c, 0, v->n, gz, gx, ng, cosv, sinv,
                    v->az, v->ax, v->ac, &cur, v->app_cap, f, created, &ovf);
    F(rotate_range)(v->az, v->ax, v->ac, 0, hi, gz, gx, ng, cosv, sinv,
                    v->az, v->ax, v->ac, &cur, v->app_cap, f, created, &ovf);
    v->an = cur;
    return ovf;
}

/* truncation.jl:222-233 / spv_evolve.jl:103-114 — AdaptiveTruncation on
 * merged state: exact sort-based top-k threshold. */
static int F(cmp_dbl)(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}
static void F(apply_adaptive)(F(or_spv) *v, int64_t max_terms, double min_thresh) {
    or_filter_t f;
    or_filter_init_nofilter(&f);
    if (v->n > max_terms) {
        double *coeffs = (double *)malloc((size_t)v->n * sizeof(double));
        for (int64_t i = 0; i < v->n; i++) coeffs[i] = C_ABS(v->c[i]);
        qsort(coeffs, (size_t)v->n, sizeof(double), F(cmp_dbl));
        /* Julia coeffs[end - max_terms] (1-based) = index n - max_terms - 1 (0-based) */
        f.thresh = coeffs[v->n - max_terms - 1];
        free(coeffs);
    } else {
        f.thresh = min_thresh;
    }
    F(compact)(v, &f);
}

/* statistics.jl:6-22 — variance(O, ψ) of the merged live set for a
 * computational-basis ket: σ = O|ψ⟩ accumulated per target ket in the
 * operator's iteration (key) order, P|k⟩ = i^((symplectic_phase + 2 sign) % 4)
 * |k ⊻ x⟩ (multiplication.jl:48-52, type_Pauli.jl:64), then
 * Σ_k' |σ[k']|^2 - σ[ψ]^2.  σ is a Dict whose iteration order is Julia's
 * hash order; the sum over k' is taken here in increasing k' (= x ⊻ ψ order
 * of the groups by x), a last-ulp difference. */
typedef struct {
    WT x;
    int64_t i;
} F(or_xi);
static int F(cmp_xi)(const void *pa, const void *pb) {
    const F(or_xi) *a = (const F(or_xi) *)pa, *b = (const F(or_xi) *)pb;
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    return a->i < b->i ? -1 : (a->i > b->i ? 1 : 0);
}
static double F(variance)(const F(or_spv) *v, WT kv) {
    const int64_t n = v->n;
    if (n == 0) return 0.0;
    F(or_xi) *ix = (F(or_xi) *)malloc((size_t)n * sizeof(F(or_xi)));
    for (int64_t i = 0; i < n; i++) {
        ix[i].x = v->x[i];
        ix[i].i = i;
    }
    qsort(ix, (size_t)n, sizeof(F(or_xi)), F(cmp_xi));  /* by x, then key order */
    double e2 = 0.0, e1r = 0.0, e1i = 0.0;
    for (int64_t g = 0; g < n;) {
        const WT x = ix[g].x;
        double re = 0.0, im = 0.0;
        int64_t j = g;
        for (; j < n && ix[j].x == x; j++) {
            const int64_t q = ix[j].i;
            const WT z = v->z[q];
            const int sign = F(popc)(z & (x ^ kv)) & 1;
            const int sp = (4 - F(popc)(z & x) % 4) % 4;
            const int ph = (sp + 2 * sign) % 4;  /* PHASE_TBL = [1, i, -1, -i] */
            /* phase * c: (1) c, (i) (-im, re), (-1) (-re, -im), (-i) (im, -re) */
            const double cr = C_RE(v->c[q]), ci = C_IM(v->c[q]);
            if (ph == 0) { re = re + cr; im = im + ci; }
            else if (ph == 1) { re = re + (-ci); im = im + cr; }
            else if (ph == 2) { re = re + (-cr); im = im + (-ci); }
            else { re = re + ci; im = im + (-cr); }
        }
        e2 = e2 + (re * re + im * im);
        if (x == (WT)0) { e1r = re; e1i = im; }
        g = j;
    }
    free(ix);
    /* real(e2 - e1 * e1), statistics.jl:19-21 */
    return e2 - (e1r * e1r - e1i * e1i);
}

/* spv_evolve.jl:246-284 — merge boundary with the correction accumulators.
 * EnergyCorrection measures the pre-merge state (live + pending appends);
 * EnergyVarianceCorrection needs the merged state: merge with NOFILTER,
 * measure, then the strict filter as a compaction (or the strategy). */
static void F(boundary)(F(or_spv) *v, const or_filter_t *f, or_strategy_t *s,
                        int have_mask, WT lz, WT lx, int64_t *n_in, int64_t *n_out) {
    const WT kv = (WT)((((u128)s->ket[1]) << 64) | (u128)s->ket[0]);
    if (s->variance) {
        or_filter_t nof;
        or_filter_init_nofilter(&nof);
        int64_t m = F(gather_append)(v);
        if (have_mask) F(unshuffle_ws)(v, m, lz, lx);
        else F(sort_pending)(v->ws, m);
        F(merge_spv)(v, m, &nof, n_in, n_out);
        const double e0 = C_RE(F(expectation)(v, kv)), v0 = F(variance)(v, kv);
        if (s->kind == OR_STRAT_ADAPTIVE) F(apply_adaptive)(v, s->max_terms, s->min_thresh);
        else F(compact)(v, f);
        s->energy_acc += C_RE(F(expectation)(v, kv)) - e0;  /* truncation.jl:308-310 */
        s->variance_acc += F(variance)(v, kv) - v0;
        return;
    }
    /* real(expectation_value(O, ψ)), truncation.jl:288 */
    const double before = s->energy ? C_RE(F(expectation)(v, kv)) : 0.0;
    int64_t m = F(gather_append)(v);
    if (have_mask) F(unshuffle_ws)(v, m, lz, lx);
    else F(sort_pending)(v->ws, m);
    F(merge_spv)(v, m, f, n_in, n_out);
    if (s->kind == OR_STRAT_ADAPTIVE) F(apply_adaptive)(v, s->max_terms, s->min_thresh);
    if (s->energy) {
        const double after = C_RE(F(expectation)(v, kv));
        s->energy_acc += after - before;  /* truncation.jl:304-306 */
    }
}

/* spv_evolve.jl:312-387 — the windowed driver. */
static int F(evolve_seq)(F(or_spv) *v, const WT *gz, const WT *gx, const double *cosv,
                         const double *sinv, int64_t L, int64_t window,
                         const or_filter_t *strict, or_strategy_t *strat,
                         const or_filter_t *local, or_counters_t *cnt) {
    if (window < 1) return -1;
    int since = 0;
    WT lz = 0, lx = 0;
    for (int64_t i = 0; i < L; i++) {
        int64_t w = i / window; /* 0-based window index */
        int64_t n_in, n_out;
        if (2 * v->an + v->n > v->app_cap) {
            F(boundary)(v, strict, strat, since == 1, lz, lx, &n_in, &n_out);
            since = 0;
            if (cnt) {
                cnt->early_merges += 1;
                cnt->merge_in += n_in;
                cnt->merge_out += n_out;
            }
            (void)w;
            if (v->n > v->app_cap) F(grow_append)(v, v->n);
        }
        int ng = F(popc)(gz[i] & gx[i]);
        int64_t swept = v->n + v->an;
        int64_t created;
        int ovf = F(rotate_spv)(v, gz[i], gx[i], ng, cosv[i], sinv[i], local, &created);
        if (ovf) return -2;
        lz = gz[i];
        lx = gx[i];
        since += 1;
        if (cnt) {
            cnt->terms_created += created;
            cnt->rotations += 1;
            cnt->terms_swept += swept; /* terms entering the sweep (window=1: live n_in) */
        }
        if ((i + 1) % window == 0 || i == L - 1) {
            F(boundary)(v, strict, strat, since == 1, lz, lx, &n_in, &n_out);
            since = 0;
            if (cnt) {
                cnt->merge_in += n_in;
                cnt->merge_out += n_out;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-instance bodies of the exported API (spv_oracle.c dispatches on the
 * handle's word count and coefficient type).  Keys cross as W little-endian
 * uint64 words per term; coefficients as one double per term (Float64) or
 * two, (re, im) (ComplexF64). */

static inline WT F(ldw)(const uint64_t *p, int64_t i) {
    if (sizeof(WT) == 8) return (WT)p[i];
    return (WT)((((u128)p[2 * i + 1]) << 64) | (u128)p[2 * i]);
}
static inline void F(stw)(uint64_t *p, int64_t i, WT v) {
    if (sizeof(WT) == 8) {
        p[i] = (uint64_t)v;
    } else {
        p[2 * i] = (uint64_t)v;
        p[2 * i + 1] = (uint64_t)((u128)v >> 64);
    }
}

static void *F(api_create)(int N, int64_t live_cap, int64_t app_cap) { return F(or_alloc)(N, live_cap, app_cap); }
static void F(api_destroy)(void *vv) { F(or_free)((F(or_spv) *)vv); }

/* type_SparsePauliVector.jl:167-188: fill, then the construction sort. */
static int F(api_load)(void *vv, const uint64_t *z, const uint64_t *x, const double *c, int64_t n) {
    F(or_spv) *v = (F(or_spv) *)vv;
    if (n > v->live_cap) F(grow_live)(v, n);
    for (int64_t i = 0; i < n; i++) {
        v->z[i] = F(ldw)(z, i);
        v->x[i] = F(ldw)(x, i);
        v->c[i] = C_LD(c, i);
    }
    v->n = n;
    v->an = 0;
    F(sort_live)(v);
    return 0;
}

static int64_t F(api_size)(void *vv) { return ((F(or_spv) *)vv)->n; }
static int64_t F(api_pending)(void *vv) { return ((F(or_spv) *)vv)->an; }
static void F(api_capacities)(void *vv, int64_t *live_cap, int64_t *app_cap) {
    *live_cap = ((F(or_spv) *)vv)->live_cap;
    *app_cap = ((F(or_spv) *)vv)->app_cap;
}

static void F(api_get)(void *vv, uint64_t *z, uint64_t *x, double *c) {
    F(or_spv) *v = (F(or_spv) *)vv;
    for (int64_t i = 0; i < v->n; i++) {
        F(stw)(z, i, v->z[i]);
        F(stw)(x, i, v->x[i]);
        C_ST(c, i, v->c[i]);
    }
}

/* spv_evolve.jl:228-236 — single rotation, NOFILTER, unshuffle merge. */
static int F(api_evolve_single)(void *vv, const uint64_t *gz, const uint64_t *gx, double cosv, double sinv) {
    F(or_spv) *v = (F(or_spv) *)vv;
    or_filter_t nof;
    or_filter_init_nofilter(&nof);
    int64_t n_in, n_out, created;
    const WT z = F(ldw)(gz, 0), x = F(ldw)(gx, 0);
    const int ng = F(popc)(z & x);
    if (v->n + 2 * v->an > v->app_cap) F(grow_append)(v, v->n + 2 * v->an);
    if (F(rotate_spv)(v, z, x, ng, cosv, sinv, &nof, &created)) return -2;
    int64_t m = F(gather_append)(v);
    F(unshuffle_ws)(v, m, z, x);
    F(merge_spv)(v, m, &nof, &n_in, &n_out);
    return 0;
}

static int F(api_evolve_seq)(void *vv, const uint64_t *gz, const uint64_t *gx, const double *cosv,
                             const double *sinv, int64_t L, int64_t window, const or_filter_t *f,
                             or_strategy_t *s, const or_filter_t *local, or_counters_t *cnt) {
    WT *z = (WT *)malloc((size_t)(L > 0 ? L : 1) * sizeof(WT));
    WT *x = (WT *)malloc((size_t)(L > 0 ? L : 1) * sizeof(WT));
    for (int64_t i = 0; i < L; i++) {
        z[i] = F(ldw)(gz, i);
        x[i] = F(ldw)(gx, i);
    }
    int rc = F(evolve_seq)((F(or_spv) *)vv, z, x, cosv, sinv, L, window, f, s, local, cnt);
    free(z);
    free(x);
    return rc;
}

static double F(api_variance)(void *vv, const uint64_t *ket) { return F(variance)((F(or_spv) *)vv, F(ldw)(ket, 0)); }
static void F(api_compact)(void *vv, const or_filter_t *f) { F(compact)((F(or_spv) *)vv, f); }
static void F(api_adaptive)(void *vv, int64_t max_terms, double min_thresh) {
    F(apply_adaptive)((F(or_spv) *)vv, max_terms, min_thresh);
}

/* channels.jl:203-265 — c *= factor (Complex * Real componentwise).
 * kind: 0 depolarizing (z|x), 1 dephasing (x), 2 bit flip (z), 3 bit-phase
 * flip (x^z), 4 general Pauli channel (lamX^nX * lamY^nY) * lamZ^nZ. */
static void F(api_channel)(void *vv, int kind, const uint64_t *mask, double lam, double lamX, double lamY,
                           double lamZ) {
    F(or_spv) *v = (F(or_spv) *)vv;
    const WT M = F(ldw)(mask, 0);
    for (int64_t i = 0; i < v->n; i++) {
        const WT z = v->z[i], x = v->x[i];
        double fac;
        if (kind == 4) {
            const WT xM = x & M, zM = z & M;
            const int nX = F(popc)(xM & ~z), nY = F(popc)(xM & z), nZ = F(popc)(zM & ~x);
            fac = or_julia_powi(lamX, nX) * or_julia_powi(lamY, nY);
            fac = fac * or_julia_powi(lamZ, nZ);
        } else {
            const WT sel = kind == 0 ? (z | x) : kind == 1 ? x : kind == 2 ? z : (x ^ z);
            fac = or_julia_powi(lam, F(popc)(sel & M));
        }
        v->c[i] = C_SCALE(v->c[i], fac);
    }
}

/* spv_evolve.jl:133-134 -> spv_kernels.jl:441-452: out[0] real part, out[1]
 * imaginary part (0 for Float64). */
static void F(api_expectation)(void *vv, const uint64_t *ket, double *out) {
    const CT e = F(expectation)((F(or_spv) *)vv, F(ldw)(ket, 0));
    out[0] = C_RE(e);
    out[1] = C_IM(e);
}

/* type_SparsePauliVector.jl:263-278 (key-order part of check_spv). */
static int F(api_check)(void *vv) {
    F(or_spv) *v = (F(or_spv) *)vv;
    for (int64_t i = 1; i < v->n; i++)
        if (!F(key_lt)(v->z[i - 1], v->x[i - 1], v->z[i], v->x[i])) return 0;
    return 1;
}
