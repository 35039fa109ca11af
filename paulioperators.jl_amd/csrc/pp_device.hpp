// Device helpers of the SparsePauliVector engine: packed keys, the
// reference's key order, filters, the rotation algebra, channel factors and
// compensated sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_common.hpp"

namespace pp {

template <int W>
struct Key {
    uint64_t z[W];
    uint64_t x[W];
};

__device__ __forceinline__ int popc64(uint64_t v) { return __popcll(v); }

// Device (global) memory accessors.  Buffers reached through descriptor
// structs are generic pointers to the compiler, whose accesses become FLAT
// instructions; FLAT loads also count against the LDS wait counter, so an
// LDS wait would wait for them too.  Accessing through a global-address-space
// pointer gives global instructions (every buffer here is device memory).
#define PP_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ldg(const T* p, int64_t i) {
    return ((const PP_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void stg(T* p, int64_t i, T v) {
    ((PP_GLOBAL T*)p)[i] = v;
}
// 16-byte key words through the builtin vector type (HIP's ulonglong2 is a
// class, which does not copy across address spaces)
typedef unsigned long long pp_u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ ulonglong2 ldg(const ulonglong2* p, int64_t i) {
    const pp_u64x2 v = ((const PP_GLOBAL pp_u64x2*)p)[i];
    return make_ulonglong2(v.x, v.y);
}
__device__ __forceinline__ void stg(ulonglong2* p, int64_t i, ulonglong2 v) {
    pp_u64x2 w;
    w.x = v.x;
    w.y = v.y;
    ((PP_GLOBAL pp_u64x2*)p)[i] = w;
}

template <int W>
__device__ __forceinline__ Key<W> load_key(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t i) {
    Key<W> k;
    if constexpr (W == 1) {
        k.z[0] = ldg(Z, i);
        k.x[0] = ldg(X, i);
    } else {
        const ulonglong2 a = ldg(reinterpret_cast<const ulonglong2*>(Z), i);
        const ulonglong2 b = ldg(reinterpret_cast<const ulonglong2*>(X), i);
        k.z[0] = a.x; k.z[1] = a.y;
        k.x[0] = b.x; k.x[1] = b.y;
    }
    return k;
}

// Streaming loads and stores (nontemporal hint: the line is not kept in the
// caches), for inputs read once and outputs far larger than the caches.
template <typename T>
__device__ __forceinline__ void stg_nt(T* p, int64_t i, T v) {
    __builtin_nontemporal_store(v, (PP_GLOBAL T*)p + i);
}
template <typename T>
__device__ __forceinline__ T ldg_nt(const T* p, int64_t i) {
    return __builtin_nontemporal_load((const PP_GLOBAL T*)p + i);
}
template <int W>
__device__ __forceinline__ Key<W> load_key_nt(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X,
                                              int64_t i) {
    Key<W> k;
    if constexpr (W == 1) {
        k.z[0] = ldg_nt(Z, i);
        k.x[0] = ldg_nt(X, i);
    } else {
        const pp_u64x2 a = __builtin_nontemporal_load((const PP_GLOBAL pp_u64x2*)Z + i);
        const pp_u64x2 b = __builtin_nontemporal_load((const PP_GLOBAL pp_u64x2*)X + i);
        k.z[0] = a.x; k.z[1] = a.y;
        k.x[0] = b.x; k.x[1] = b.y;
    }
    return k;
}
template <int W>
__device__ __forceinline__ void store_key_nt(uint64_t* __restrict__ Z, uint64_t* __restrict__ X, int64_t i,
                                             const Key<W>& k) {
    if constexpr (W == 1) {
        stg_nt(Z, i, k.z[0]);
        stg_nt(X, i, k.x[0]);
    } else {
        pp_u64x2 a, b;
        a.x = k.z[0]; a.y = k.z[1];
        b.x = k.x[0]; b.y = k.x[1];
        __builtin_nontemporal_store(a, (PP_GLOBAL pp_u64x2*)Z + i);
        __builtin_nontemporal_store(b, (PP_GLOBAL pp_u64x2*)X + i);
    }
}

template <int W>
__device__ __forceinline__ void store_key(uint64_t* __restrict__ Z, uint64_t* __restrict__ X, int64_t i, const Key<W>& k) {
    if constexpr (W == 1) {
        stg(Z, i, k.z[0]);
        stg(X, i, k.x[0]);
    } else {
        stg(reinterpret_cast<ulonglong2*>(Z), i, make_ulonglong2(k.z[0], k.z[1]));
        stg(reinterpret_cast<ulonglong2*>(X), i, make_ulonglong2(k.x[0], k.x[1]));
    }
}

// spv_kernels.jl:14-16: z-major, high word first.
// Lexicographic order (z words from the top, then x words), as one borrow
// chain over the 4W 32-bit limbs of (z, x) read as a single unsigned number
// with z's top word most significant: a < b iff a - b borrows.
template <int W>
__device__ __forceinline__ bool key_lt(const Key<W>& a, const Key<W>& b) {
    unsigned long long br = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) (void)__builtin_subcll(a.x[w], b.x[w], br, &br);
#pragma unroll
    for (int w = 0; w < W; ++w) (void)__builtin_subcll(a.z[w], b.z[w], br, &br);
    return br != 0;
}
template <int W>
__device__ __forceinline__ bool key_le(const Key<W>& a, const Key<W>& b) { return !key_lt<W>(b, a); }
template <int W>
__device__ __forceinline__ bool key_eq(const Key<W>& a, const Key<W>& b) {
    uint64_t d = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) d |= (a.z[w] ^ b.z[w]) | (a.x[w] ^ b.x[w]);
    return d == 0;
}

// Highest differing bit of the combined key (x bits at [0, 64W), z bits above);
// -1 when equal.
template <int W>
__device__ __forceinline__ int hdb(const Key<W>& a, const Key<W>& b) {
#pragma unroll
    for (int w = W - 1; w >= 0; --w) {
        const uint64_t d = a.z[w] ^ b.z[w];
        if (d) return 64 * W + 64 * w + 63 - __clzll((long long)d);
    }
#pragma unroll
    for (int w = W - 1; w >= 0; --w) {
        const uint64_t d = a.x[w] ^ b.x[w];
        if (d) return 64 * w + 63 - __clzll((long long)d);
    }
    return -1;
}

// Word q of the combined key (x words 0..W-1, then z words).  Selects, not
// an indexed access: a key indexed by a run-time word number would be placed
// in per-thread LDS, whose addressing reads the workgroup shape from the
// dispatch packet (a host-memory read at every launch).
template <int W>
__device__ __forceinline__ uint64_t key_word(const Key<W>& k, int q) {
    uint64_t v = k.x[0];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        v = q == w ? k.x[w] : v;
        v = q == W + w ? k.z[w] : v;
    }
    return v;
}

template <int W>
__device__ __forceinline__ int key_bit(const Key<W>& k, int b) {
    return (int)((key_word<W>(k, b >> 6) >> (b & 63)) & 1ull);
}

template <int W>
__device__ __forceinline__ int popc_or(const Key<W>& k) {
    int s = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += popc64(k.z[w] | k.x[w]);
    return s;
}
template <int W>
__device__ __forceinline__ int popc_x(const Key<W>& k) {
    int s = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += popc64(k.x[w]);
    return s;
}

// spv_kernels.jl:51-62 on the word type of width 64 W.
template <int W>
__device__ __forceinline__ int majorana_weight(const Key<W>& k) {
    if constexpr (W == 1) {
        const uint64_t z = k.z[0], x = k.x[0];
        const uint64_t zonly = z & ~x;
        uint64_t S = x;
#pragma unroll
        for (int sh = 1; sh < 64; sh <<= 1) S ^= S >> sh;
        const uint64_t ctrl = ~(S ^ x);
        return popc64(x) + 2 * popc64(zonly & ctrl) + 2 * popc64(~(z | x) & ~ctrl);
    } else {
        typedef unsigned __int128 u128;
        const u128 z = ((u128)k.z[1] << 64) | k.z[0];
        const u128 x = ((u128)k.x[1] << 64) | k.x[0];
        const u128 zonly = z & ~x;
        u128 S = x;
#pragma unroll
        for (int sh = 1; sh < 128; sh <<= 1) S ^= S >> sh;
        const u128 ctrl = ~(S ^ x);
        auto pc = [](u128 v) { return popc64((uint64_t)v) + popc64((uint64_t)(v >> 64)); };
        return pc(x) + 2 * pc(zonly & ctrl) + 2 * pc(~(z | x) & ~ctrl);
    }
}

// spv_kernels.jl:64-79.
// The scalar criteria of a FilterDev held in registers (loaded once per
// kernel; the exp tables stay in memory, read only by damped filters).
struct FilterReg {
    int64_t wmax, xwmax, mwmax;
    double thresh, dthresh_w, dthresh_xw;
    const double* exptab_w;
    const double* exptab_xw;
    __device__ __forceinline__ void load(const FilterDev* f) {
        wmax = f->wmax;
        xwmax = f->xwmax;
        mwmax = f->mwmax;
        thresh = f->thresh;
        dthresh_w = f->dthresh_w;
        dthresh_xw = f->dthresh_xw;
        exptab_w = f->exptab_w;
        exptab_xw = f->exptab_xw;
    }
};

template <int W, typename F>
__device__ __forceinline__ bool should_drop_f(const F& f, const Key<W>& k, double absc) {
    if (absc <= f.thresh) return true;
    const int w = popc_or<W>(k);
    if ((int64_t)w > f.wmax) return true;
    const int xw = popc_x<W>(k);
    if ((int64_t)xw > f.xwmax) return true;
    if (f.mwmax != INT64_MAX) {
        if ((int64_t)majorana_weight<W>(k) > f.mwmax) return true;
    }
    if (f.dthresh_w >= 0.0) {
        if (absc * ldg((const double*)f.exptab_w, w) <= f.dthresh_w) return true;
    }
    if (f.dthresh_xw >= 0.0) {
        if (absc * ldg((const double*)f.exptab_xw, xw) <= f.dthresh_xw) return true;
    }
    return false;
}
template <int W>
__device__ __forceinline__ bool should_drop(const FilterDev* __restrict__ f, const Key<W>& k, double absc) {
    return should_drop_f<W>(*f, k, absc);
}
template <int W>
__device__ __forceinline__ bool should_drop(const FilterReg& f, const Key<W>& k, double absc) {
    return should_drop_f<W>(f, k, absc);
}

// ---------------------------------------------------------------------------
// Coefficient types.  A vector's coefficients are Float64 (`double`) or
// ComplexF64 (`pp_cplx`: (re, im), Julia's interleaved layout, 16 bytes).
// Kernels are templated on the type; the vector operators of the clang
// vector extension make `c * s` (s real) and `a + b` componentwise, which is
// Julia's Complex * Real and Complex + Complex bit for bit.
typedef double pp_cplx __attribute__((ext_vector_type(2)));

template <typename CT>
struct CoefTraits;
template <>
struct CoefTraits<double> {
    static constexpr bool kComplex = false;
};
template <>
struct CoefTraits<pp_cplx> {
    static constexpr bool kComplex = true;
};

__device__ __forceinline__ double coef_re(double c) { return c; }
__device__ __forceinline__ double coef_re(pp_cplx c) { return c.x; }
__device__ __forceinline__ double coef_im(double) { return 0.0; }
__device__ __forceinline__ double coef_im(pp_cplx c) { return c.y; }
template <typename CT>
__device__ __forceinline__ CT coef_zero() {
    return CT(0.0);
}

// Julia's hypot(x, y) (base/math.jl, the branch taken where the CPU has a
// hardware fma: Borges' corrected hypot, correctly rounded), which is
// abs(::ComplexF64).  muladd is taken as fused, as Julia lowers it there.
__host__ __device__ __forceinline__ double julia_hypot(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    if (isinf(ax) || isinf(ay)) return INFINITY;
    if (ay > ax) {
        const double t = ax;
        ax = ay;
        ay = t;
    }
    // sqrt(eps(Float64) / 2) = sqrt(2^-53), correctly rounded
    if (ay <= ax * 0x1.6a09e667f3bcdp-27) return ax;
    double scale = 0x1p-563;  // eps(Float64) * sqrt(floatmin(Float64)) = 2^-52 * 2^-511
    if (ax > 0x1.6a09e667f3bccp+511) {  // sqrt(floatmax(Float64) / 2)
        ax = ax * scale;
        ay = ay * scale;
        scale = 0x1p+563;  // inv(scale)
    } else if (ay < 0x1p-511) {  // sqrt(floatmin(Float64))
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

// abs(c) and abs2(c) (Julia's abs2(z) = real(z)*real(z) + imag(z)*imag(z)).
__device__ __forceinline__ double coef_abs(double c) { return fabs(c); }
__device__ __forceinline__ double coef_abs(pp_cplx c) { return julia_hypot(c.x, c.y); }
__device__ __forceinline__ double coef_abs2(double c) { return c * c; }
__device__ __forceinline__ double coef_abs2(pp_cplx c) { return c.x * c.x + c.y * c.y; }

// The rotation child's coefficient (T(k - 2) * sinθ) * c (spv_kernels.jl:163).
// For ComplexF64, T(k - 2) * sinθ is Complex((k - 2) sinθ, 0.0 sinθ) and the
// product with c is Julia's Complex * Complex,
// (a re - b im, a im + b re): the b terms are signed zeros, kept so that
// zero components carry Julia's signs.
__device__ __forceinline__ double coef_rot(double a, double s, double c) {
    (void)s;
    return a * c;
}
__device__ __forceinline__ pp_cplx coef_rot(double a, double s, pp_cplx c) {
    const double b = 0.0 * s;
    pp_cplx r;
    r.x = a * c.x - b * c.y;
    r.y = a * c.y + b * c.x;
    return r;
}

template <int W>
struct GenReg {
    uint64_t gz[W], gx[W];
    double cosv, sinv;
    int ng;
    __device__ __forceinline__ void load(const GenDev* g) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            gz[w] = g->gz[w];
            gx[w] = g->gx[w];
        }
        cosv = g->cosv;
        sinv = g->sinv;
        ng = g->ng;
    }
    // popcount(gx & z) - popcount(gz & x) odd  (spv_kernels.jl:157-159)
    __device__ __forceinline__ bool anti(const Key<W>& k) const {
        int m = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) m += popc64(gx[w] & k.z[w]) + popc64(gz[w] & k.x[w]);
        return (m & 1) != 0;
    }
    // spv_kernels.jl:160-163: child key and coefficient from the OLD c.
    template <typename CT>
    __device__ __forceinline__ CT child(const Key<W>& k, CT c, Key<W>& out) const {
        int m1 = 0, nyp = 0, ny = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            out.z[w] = gz[w] ^ k.z[w];
            out.x[w] = gx[w] ^ k.x[w];
            m1 += popc64(gx[w] & k.z[w]);
            nyp += popc64(out.z[w] & out.x[w]);
            ny += popc64(k.z[w] & k.x[w]);
        }
        const int kk = (nyp - ng - ny + 2 * m1) & 3;
        return coef_rot((double)(kk - 2) * sinv, sinv, c);
    }
};


// Global exclusive A-prefix at live position p (0 <= p <= n).
template <int T>
__device__ __forceinline__ int64_t a_prefix(int64_t p, int64_t n, int64_t ntiles, const uint64_t* __restrict__ flags,
                                            const uint16_t* __restrict__ wpre, const int64_t* __restrict__ tileoff) {
    if (p >= n) return ldg(tileoff, ntiles);
    const int64_t w = p >> 6;
    const uint64_t lowmask = (p & 63) ? ((~0ull) >> (64 - (p & 63))) : 0ull;
    return ldg(tileoff, p / T) + (int64_t)ldg(wpre, w) + popc64(ldg(flags, w) & lowmask);
}

// First q <= p whose key agrees with K_p on every bit >= b.
template <int W>
__device__ int64_t seg_start(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t p, int b) {
    const Key<W> kp = load_key<W>(Z, X, p);
    int64_t in = p, out = -1, step = 1;
    while (true) {
        const int64_t c = p - step;
        if (c < 0) { out = -1; break; }
        if (hdb<W>(load_key<W>(Z, X, c), kp) < b) {
            in = c;
            step <<= 1;
        } else {
            out = c;
            break;
        }
    }
    while (in - out > 1) {
        const int64_t mid = out + (in - out) / 2;
        if (hdb<W>(load_key<W>(Z, X, mid), kp) < b) in = mid;
        else out = mid;
    }
    return in;
}

// First r > p whose key differs from K_p on some bit >= b (n if none).
template <int W>
__device__ int64_t seg_end(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t p, int64_t n, int b) {
    const Key<W> kp = load_key<W>(Z, X, p);
    int64_t in = p, out = n, step = 1;
    while (true) {
        const int64_t c = p + step;
        if (c >= n) { out = n; break; }
        if (hdb<W>(load_key<W>(Z, X, c), kp) < b) {
            in = c;
            step <<= 1;
        } else {
            out = c;
            break;
        }
    }
    while (out - in > 1) {
        const int64_t mid = in + (out - in) / 2;
        if (hdb<W>(load_key<W>(Z, X, mid), kp) < b) in = mid;
        else out = mid;
    }
    return out;
}


// K_q agrees with kp on every combined-key bit >= b (hdb(K_q, kp) < b).  The
// x words (combined bits [0, 64 W)) are read only when b falls among them:
// for a level in the z words, a probe fetches half the bytes.
template <int W>
__device__ __forceinline__ bool agrees_above(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X,
                                             int64_t q, const Key<W>& kp, int b) {
    uint64_t z[W], x[W];
    const bool need_x = b < 64 * W;
    if constexpr (W == 1) {
        z[0] = ldg(Z, q);
        x[0] = need_x ? ldg(X, q) : 0ull;
    } else {
        const ulonglong2 u = ldg(reinterpret_cast<const ulonglong2*>(Z), q);
        z[0] = u.x; z[1] = u.y;
        if (need_x) {
            const ulonglong2 v = ldg(reinterpret_cast<const ulonglong2*>(X), q);
            x[0] = v.x; x[1] = v.y;
        } else {
            x[0] = x[1] = 0ull;
        }
    }
    bool ok = true;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int lo = 64 * W + 64 * w;  // combined position of bit 0 of z[w]
        if (b < lo + 64) {
            const uint64_t m = b <= lo ? ~0ull : (~0ull << (b - lo));
            ok = ok && ((z[w] ^ kp.z[w]) & m) == 0;
        }
    }
    if (need_x) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int lo = 64 * w;
            if (b < lo + 64) {
                const uint64_t m = b <= lo ? ~0ull : (~0ull << (b - lo));
                ok = ok && ((x[w] ^ kp.x[w]) & m) == 0;
            }
        }
    }
    return ok;
}

// Wave-cooperative versions: all 64 lanes call with the same arguments and
// get the same result.  Each step probes 64 positions at once (a galloping
// search with stride 64^k, then a 64-ary refinement), so a segment of length
// L costs about 2 log_64(L) dependent load rounds instead of 2 log_2(L).

// First q <= p whose key agrees with kp = K_p on every bit >= b.
template <int W>
__device__ int64_t seg_start_wave(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t p, int b,
                                  const Key<W>& kp) {
    const int lane = threadIdx.x % kWave;
    int64_t in = p, out = -1, stride = 1;
    while (true) {
        const int64_t q = in - (int64_t)(lane + 1) * stride;
        const bool inside = q >= 0 && agrees_above<W>(Z, X, q, kp, b);
        const uint64_t outm = __ballot(!inside);
        if (outm == 0) {
            in -= (int64_t)kWave * stride;
            stride *= kWave;
            continue;
        }
        const int f = __ffsll((unsigned long long)outm) - 1;
        const int64_t qf = in - (int64_t)(f + 1) * stride;
        out = qf < 0 ? -1 : qf;
        in = in - (int64_t)f * stride;
        break;
    }
    while (in - out > 1) {
        const int64_t step = (in - out - 1 + kWave - 1) / kWave;
        const int64_t q = out + step * (int64_t)(lane + 1);
        const bool valid = q < in;
        const bool inside = valid && agrees_above<W>(Z, X, q, kp, b);
        const uint64_t inm = __ballot(inside);
        if (inm == 0) {
            int64_t kmax = (in - out - 1) / step;
            if (kmax > kWave) kmax = kWave;
            out = out + step * kmax;
        } else {
            const int g = __ffsll((unsigned long long)inm) - 1;
            const int64_t qg = out + step * (int64_t)(g + 1);
            out = g == 0 ? out : out + step * (int64_t)g;
            in = qg;
        }
    }
    return in;
}

// First r > p whose key differs from kp = K_p on some bit >= b (n if none).
template <int W>
__device__ int64_t seg_end_wave(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t p, int64_t n,
                                int b, const Key<W>& kp) {
    const int lane = threadIdx.x % kWave;
    int64_t in = p, out = n, stride = 1;
    while (true) {
        const int64_t q = in + (int64_t)(lane + 1) * stride;
        const bool inside = q < n && agrees_above<W>(Z, X, q, kp, b);
        const uint64_t outm = __ballot(!inside);
        if (outm == 0) {
            in += (int64_t)kWave * stride;
            stride *= kWave;
            continue;
        }
        const int f = __ffsll((unsigned long long)outm) - 1;
        const int64_t qf = in + (int64_t)(f + 1) * stride;
        out = qf > n ? n : qf;
        in = in + (int64_t)f * stride;
        break;
    }
    while (out - in > 1) {
        const int64_t step = (out - in - 1 + kWave - 1) / kWave;
        const int64_t q = in + step * (int64_t)(lane + 1);
        const bool valid = q < out;
        const bool outside = valid && !(agrees_above<W>(Z, X, q, kp, b));
        const uint64_t om = __ballot(outside);
        if (om == 0) {
            int64_t kmax = (out - in - 1) / step;
            if (kmax > kWave) kmax = kWave;
            in = in + step * kmax;
        } else {
            const int h = __ffsll((unsigned long long)om) - 1;
            const int64_t qh = in + step * (int64_t)(h + 1);
            in = h == 0 ? in : in + step * (int64_t)h;
            out = qh;
        }
    }
    return out;
}

struct ChanDev {
    int kind;
    uint64_t mask[2];
    const double* t0;
    const double* t1;
    const double* t2;
};


template <int W>
__device__ __forceinline__ double chan_factor(const ChanDev& ch, const Key<W>& k) {
    int a = 0, b = 0, c = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t z = k.z[w], x = k.x[w], M = ch.mask[w];
        switch (ch.kind) {
            case 0: a += popc64((z | x) & M); break;
            case 1: a += popc64(x & M); break;
            case 2: a += popc64(z & M); break;
            case 3: a += popc64((x ^ z) & M); break;
            default:
                a += popc64(x & M & ~z);
                b += popc64(x & M & z);
                c += popc64(z & M & ~x);
                break;
        }
    }
    if (ch.kind != 4) return ch.t0[a];
    double f = ch.t0[a] * ch.t1[b];
    return f * ch.t2[c];
}

// Merge look-back words: [flag:2][epoch:22][count:40].  The epoch is the
// launch number, so words left by earlier launches read as "not ready" and
// the array never needs clearing.
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr int kEpochShift = 40;
constexpr uint64_t kEpochMask = (1ull << 22) - 1;
constexpr uint64_t kCountMask = (1ull << kEpochShift) - 1;

// Key order A[ia] <= B[ib] reading the z words first and the x words only
// where the z words are equal: a search probe then fetches two scattered
// cache lines instead of four unless its keys share z.
template <int W>
__device__ __forceinline__ bool key_le_zfirst(const uint64_t* __restrict__ Az, const uint64_t* __restrict__ Ax,
                                              int64_t ia, const uint64_t* __restrict__ Bz,
                                              const uint64_t* __restrict__ Bx, int64_t ib) {
    uint64_t a[W], b[W];
    if constexpr (W == 1) {
        a[0] = ldg(Az, ia);
        b[0] = ldg(Bz, ib);
    } else {
        const ulonglong2 u = ldg(reinterpret_cast<const ulonglong2*>(Az), ia);
        const ulonglong2 v = ldg(reinterpret_cast<const ulonglong2*>(Bz), ib);
        a[0] = u.x; a[1] = u.y;
        b[0] = v.x; b[1] = v.y;
    }
    bool tie = true, lt = false;
#pragma unroll
    for (int w = W - 1; w >= 0; --w) {
        if (tie && a[w] != b[w]) {
            lt = a[w] < b[w];
            tie = false;
        }
    }
    if (!tie) return lt;
    if constexpr (W == 1) {
        a[0] = ldg(Ax, ia);
        b[0] = ldg(Bx, ib);
    } else {
        const ulonglong2 u = ldg(reinterpret_cast<const ulonglong2*>(Ax), ia);
        const ulonglong2 v = ldg(reinterpret_cast<const ulonglong2*>(Bx), ib);
        a[0] = u.x; a[1] = u.y;
        b[0] = v.x; b[1] = v.y;
    }
#pragma unroll
    for (int w = W - 1; w >= 0; --w)
        if (a[w] != b[w]) return a[w] < b[w];
    return true;
}

// Merge-path split: the number of live terms among the first d outputs of
// the merge of L (n) and C (m), live first on equal keys.  Wave-cooperative
// 64-ary search over the monotone predicate L[i] <= C[d-1-i].  ZFIRST: probes
// read the x words only on z ties (half the scattered lines where z words
// differ, one more dependent load where they tie; measured faster for every
// merge tile size of kb_merge).
template <int W, bool ZFIRST = false>
__device__ int64_t merge_split_wave(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx, int64_t n,
                                    const uint64_t* __restrict__ Cz, const uint64_t* __restrict__ Cx, int64_t m,
                                    int64_t d) {
    const int lane = threadIdx.x % kWave;
    int64_t lo = d - m > 0 ? d - m : 0;
    int64_t hi = d < n ? d : n;
    while (lo < hi) {
        const int64_t step = (hi - lo + kWave - 1) / kWave;
        const int64_t i = lo + step * (int64_t)lane;
        const bool valid = i < hi;
        bool pred = false;
        if (valid) {
            if constexpr (ZFIRST) pred = key_le_zfirst<W>(Lz, Lx, i, Cz, Cx, d - 1 - i);
            else pred = key_le<W>(load_key<W>(Lz, Lx, i), load_key<W>(Cz, Cx, d - 1 - i));
        }
        const uint64_t tm = __ballot(pred);
        const uint64_t vm = __ballot(valid);
        if (tm == vm) {
            lo = lo + step * (int64_t)(__popcll(vm) - 1) + 1;
        } else {
            const int g = __ffsll((unsigned long long)(~tm & vm)) - 1;
            const int64_t ig = lo + step * (int64_t)g;
            lo = g == 0 ? lo : lo + step * (int64_t)(g - 1) + 1;
            hi = ig;
        }
    }
    return lo;
}


__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_add(double& h, double& l, double bh, double bl) {
    double s, e;
    two_sum(h, bh, s, e);
    e = e + (l + bl);
    h = s + e;
    l = e - (h - s);
}


}  // namespace pp
