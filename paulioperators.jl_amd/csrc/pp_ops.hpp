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
