// Lookups and reductions on device-resident vectors: the Dict-idiom access
// of type_SparsePauliVector.jl:286-311 (get / getindex / haskey by binary
// search of the sorted keys) and the comparison reductions of spv_ops.jl:
// norm (:57-84), isapprox (:88-107), inner_product (:378-396).  Sums use a
// double-double accumulator per thread and per block, so they equal the
// exactly rounded sum to about one ulp (the reference sums left to right).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_common.hpp"
#include "pp_device.hpp"

namespace pp {

// Index of key k in the sorted array (z, x, n), or -1 (spv: _find).
template <int W>
__device__ __forceinline__ int64_t find_key(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t n,
                                            const Key<W>& k) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        const Key<W> m = load_key<W>(Z, X, mid);
        if (key_lt<W>(m, k)) lo = mid + 1;
        else hi = mid;
    }
    if (lo < n && key_eq<W>(load_key<W>(Z, X, lo), k)) return lo;
    return -1;
}

template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) ko_lookup(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X,
                                                    const CT* __restrict__ C, int64_t n,
                                                    const uint64_t* __restrict__ Qz, const uint64_t* __restrict__ Qx,
                                                    int64_t count, CT* __restrict__ out,
                                                    int32_t* __restrict__ found, int64_t* __restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock) {
        const int64_t j = find_key<W>(Z, X, n, load_key<W>(Qz, Qx, i));
        found[i] = j >= 0 ? 1 : 0;
        out[i] = j >= 0 ? C[j] : coef_zero<CT>();
        if (idx) idx[i] = j;
    }
}

// setindex! on present keys: C[idx[i]] = v[i]
template <typename CT>
__global__ void __launch_bounds__(kBlock) ko_scatter(CT* __restrict__ C, const int64_t* __restrict__ idx,
                                                     const CT* __restrict__ v, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock)
        C[idx[i]] = v[i];
}

// delete!: mark[idx[i]] = 1
__global__ void __launch_bounds__(kBlock) ko_mark(uint8_t* __restrict__ mark, const int64_t* __restrict__ idx,
                                                  int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock)
        mark[idx[i]] = 1;
}

__device__ __forceinline__ void block_dd_reduce(double h, double l, double2* __restrict__ partial) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    s_h[threadIdx.x] = h;
    s_l[threadIdx.x] = l;
    __syncthreads();
    for (int sh = kBlock / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) {
            double a = s_h[threadIdx.x], c = s_l[threadIdx.x];
            dd_add(a, c, s_h[threadIdx.x + sh], s_l[threadIdx.x + sh]);
            s_h[threadIdx.x] = a;
            s_l[threadIdx.x] = c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s_h[0], s_l[0]);
}

// Σ over the terms of A of:
//   mode 0 (inner product)   conj(c_a) c_b           for keys also in B
//                            (component comp: 0 real, 1 imaginary)
//   mode 1 (difference, A)   abs2(c_a - c_b) if in B, else abs2(c_a)
//   mode 2 (difference, B)   abs2(c_a)               for keys not in B
// conj(a) b is Julia's Complex * Complex of (a.re, -a.im) and b.
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) ko_join(const uint64_t* __restrict__ Az, const uint64_t* __restrict__ Ax,
                                                  const CT* __restrict__ Ac, int64_t na,
                                                  const uint64_t* __restrict__ Bz, const uint64_t* __restrict__ Bx,
                                                  const CT* __restrict__ Bc, int64_t nb, int mode, int comp,
                                                  double2* __restrict__ partial) {
    double h = 0.0, l = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < na; i += (int64_t)gridDim.x * kBlock) {
        const int64_t j = find_key<W>(Bz, Bx, nb, load_key<W>(Az, Ax, i));
        const CT a = Ac[i];
        double v = 0.0;
        if (mode == 0) {
            if (j >= 0) {
                const CT b = Bc[j];
                const double ar = coef_re(a), ai = -coef_im(a), br = coef_re(b), bi = coef_im(b);
                if constexpr (CoefTraits<CT>::kComplex) v = comp ? ar * bi + ai * br : ar * br - ai * bi;
                else v = ar * br;
            }
        } else if (mode == 1) {
            v = coef_abs2(j >= 0 ? a - Bc[j] : a);
        } else {
            if (j < 0) v = coef_abs2(a);
        }
        dd_add(h, l, v, 0.0);
    }
    block_dd_reduce(h, l, partial);
}

// kind 0: Σ abs2(c), 1: Σ |c|, 2: max |c|, 3: Σ |c|^p
template <typename CT>
__global__ void __launch_bounds__(kBlock) ko_norm(const CT* __restrict__ C, int64_t n, int kind, double p,
                                                  double2* __restrict__ partial) {
    double h = 0.0, l = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double a = kind == 0 ? 0.0 : coef_abs(C[i]);
        if (kind == 2) {
            h = a > h ? a : h;
        } else {
            const double v = kind == 0 ? coef_abs2(C[i]) : (kind == 1 ? a : pow(a, p));
            dd_add(h, l, v, 0.0);
        }
    }
    if (kind != 2) {
        block_dd_reduce(h, l, partial);
        return;
    }
    __shared__ double s_m[kBlock];
    s_m[threadIdx.x] = h;
    __syncthreads();
    for (int sh = kBlock / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh && s_m[threadIdx.x + sh] > s_m[threadIdx.x]) s_m[threadIdx.x] = s_m[threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s_m[0], 0.0);
}

// ---------------------------------------------------------------------------
// largest / largest_diag (spv_ops.jl:520-553): the first term with the largest
// |c| (a strict > scan from the front), among all terms or those with x == 0.
// Per block (best |c|, index; index -1: none), then one block over the
// partials; ties keep the lower index, a NaN |c| is never taken.
struct ArgMax {
    double a;
    int64_t i;
};
__device__ __forceinline__ ArgMax argmax_join(const ArgMax& p, const ArgMax& q) {
    if (q.i < 0) return p;
    if (p.i < 0) return q;
    if (q.a > p.a || (q.a == p.a && q.i < p.i)) return q;
    return p;
}
__device__ __forceinline__ ArgMax block_argmax(ArgMax m) {
    __shared__ double s_a[kBlock];
    __shared__ int64_t s_i[kBlock];
    s_a[threadIdx.x] = m.a;
    s_i[threadIdx.x] = m.i;
    __syncthreads();
    for (int sh = kBlock / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) {
            const ArgMax r = argmax_join(ArgMax{s_a[threadIdx.x], s_i[threadIdx.x]},
                                         ArgMax{s_a[threadIdx.x + sh], s_i[threadIdx.x + sh]});
            s_a[threadIdx.x] = r.a;
            s_i[threadIdx.x] = r.i;
        }
        __syncthreads();
    }
    return ArgMax{s_a[0], s_i[0]};
}
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) ko_argmax(const uint64_t* __restrict__ X, const CT* __restrict__ C, int64_t n,
                                                    int diag_only, double2* __restrict__ partial) {
    ArgMax m{0.0, -1};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        if (diag_only) {
            uint64_t xo = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) xo |= X[i * W + w];
            if (xo != 0) continue;
        }
        const double a = coef_abs(C[i]);
        if (!(a == a)) continue;  // NaN: `a > best` is never true
        m = argmax_join(m, ArgMax{a, i});
    }
    const ArgMax r = block_argmax(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(r.a, __longlong_as_double((long long)r.i));
}
__global__ void __launch_bounds__(kBlock) ko_argmax_final(const double2* __restrict__ partial, int64_t np,
                                                          int64_t* __restrict__ out) {
    ArgMax m{0.0, -1};
    for (int64_t i = threadIdx.x; i < np; i += kBlock)
        m = argmax_join(m, ArgMax{partial[i].x, (int64_t)__double_as_longlong(partial[i].y)});
    const ArgMax r = block_argmax(m);
    if (threadIdx.x == 0) *out = r.i;
}

// ---------------------------------------------------------------------------
// otimes(v1, v2) (spv_ops.jl:342-358): every pair (i, j) gives the key
// (z1 | z2 << N1, x1 | x2 << N1) (type_PauliBasis.jl:130-132) and c1 * c2
// (Julia's T * T; Complex * Complex is (a re b re - a im b im, a re b im +
// a im b re)).  Pair i * nb + j; the caller sorts by key.  Keys of any word
// count are handled as 128-bit values.
__device__ __forceinline__ unsigned __int128 ld_u128(const uint64_t* p, int64_t i, int W) {
    return W == 1 ? (unsigned __int128)p[i] : (((unsigned __int128)p[2 * i + 1] << 64) | p[2 * i]);
}
__device__ __forceinline__ double coef_mul(double a, double b) { return a * b; }
__device__ __forceinline__ pp_cplx coef_mul(pp_cplx a, pp_cplx b) {
    pp_cplx r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename CT>
__global__ void __launch_bounds__(kBlock) ko_otimes(const uint64_t* __restrict__ Az, const uint64_t* __restrict__ Ax,
                                                    const CT* __restrict__ Ac, int64_t na, int Wa,
                                                    const uint64_t* __restrict__ Bz, const uint64_t* __restrict__ Bx,
                                                    const CT* __restrict__ Bc, int64_t nb, int Wb, int Na,
                                                    uint64_t* __restrict__ Oz, uint64_t* __restrict__ Ox,
                                                    CT* __restrict__ Oc, int Wo) {
    const int64_t total = na * nb;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < total; p += (int64_t)gridDim.x * kBlock) {
        const int64_t i = p / nb, j = p - i * nb;
        const unsigned __int128 z = ld_u128(Az, i, Wa) | (ld_u128(Bz, j, Wb) << Na);
        const unsigned __int128 x = ld_u128(Ax, i, Wa) | (ld_u128(Bx, j, Wb) << Na);
        if (Wo == 1) {
            Oz[p] = (uint64_t)z;
            Ox[p] = (uint64_t)x;
        } else {
            Oz[2 * p] = (uint64_t)z;
            Oz[2 * p + 1] = (uint64_t)(z >> 64);
            Ox[2 * p] = (uint64_t)x;
            Ox[2 * p + 1] = (uint64_t)(x >> 64);
        }
        Oc[p] = coef_mul(Ac[i], Bc[j]);
    }
}

// ---------------------------------------------------------------------------
// Matrix(v) (spv_ops.jl:109-115): out = Σ over terms in key order of
// Matrix(op) .* c, Matrix(PauliBasis) with site 1 the last kron factor
// (type_PauliBasis.jl:71-95), so its only nonzero in row r is column r ⊻ x,
// with value i^nY (-1)^popcount(r & z & ~x) Π_Y (2 r_q - 1).  The terms come
// sorted by (x, z) (Sz holds the x words, Sx the z words), so entry (r, col)
// sums the group x = r ⊻ col in increasing z, the reference's key order.
// out is column-major (Julia's layout), D = 2^N, cw doubles per entry.
// A Float64 matrix cannot hold an imaginary entry: err is set (InexactError).
template <typename CT>
__global__ void __launch_bounds__(kBlock) ko_matrix(const uint64_t* __restrict__ Sz, const uint64_t* __restrict__ Sx,
                                                    const CT* __restrict__ Sc, int64_t n, int W, int N,
                                                    double* __restrict__ out, int32_t* __restrict__ err) {
    const int64_t D = (int64_t)1 << N;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D * D; e += (int64_t)gridDim.x * kBlock) {
        const uint64_t r = (uint64_t)(e % D), col = (uint64_t)(e / D);
        const uint64_t x = r ^ col;
        // first term of the x group (the x words beyond word 0 are zero: N <= 13)
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            bool less = Sz[mid * W] < x;
            if (W == 2) less = Sz[mid * W + 1] == 0 && less;
            if (less) lo = mid + 1;
            else hi = mid;
        }
        double re = 0.0, im = 0.0;
        for (int64_t t = lo; t < n; ++t) {
            if (Sz[t * W] != x || (W == 2 && Sz[t * W + 1] != 0)) break;
            const uint64_t z = Sx[t * W];
            const uint64_t y = z & x;
            const int ny = popc64(y);
            int neg = popc64(r & z & ~x) + popc64(~r & y);  // Z sites at r_q = 1, Y sites at r_q = 0
            // i^ny * (-1)^neg: (er, ei) in {(±1, 0), (0, ±1)}
            const int q = (ny + 2 * (neg & 1)) & 3;
            const double er = q == 0 ? 1.0 : (q == 2 ? -1.0 : 0.0);
            const double ei = q == 1 ? 1.0 : (q == 3 ? -1.0 : 0.0);
            const CT c = Sc[t];
            // Matrix(op) .* c, then out .+= (Complex{Int} * T in Julia)
            if constexpr (CoefTraits<CT>::kComplex) {
                re = re + (er * coef_re(c) - ei * coef_im(c));
                im = im + (er * coef_im(c) + ei * coef_re(c));
            } else {
                if (ei != 0.0 && coef_re(c) != 0.0) *err = 1;
                re = re + er * coef_re(c);
            }
        }
        if constexpr (CoefTraits<CT>::kComplex) {
            out[2 * e] = re;
            out[2 * e + 1] = im;
        } else {
            out[e] = re;
        }
    }
}

__global__ void __launch_bounds__(kBlock) ko_max_final(const double2* __restrict__ partial, int64_t np,
                                                       double* __restrict__ out) {
    __shared__ double s_m[kBlock];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += kBlock) m = partial[i].x > m ? partial[i].x : m;
    s_m[threadIdx.x] = m;
    __syncthreads();
    for (int sh = kBlock / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh && s_m[threadIdx.x + sh] > s_m[threadIdx.x]) s_m[threadIdx.x] = s_m[threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s_m[0];
}

}  // namespace pp
