// Kernels of the Pauli-sum products (spv_ops.jl:255-484): A·B with optional
// adjoints (_mul_spv), the commutator [A, B] and the anticommutator {A, B}.
//
//   ka_pairs   every pair (i, j) in generation order p = i nB + j (the order
//              the reference fills its workspace in): product key
//              (zA xor zB, xA xor xB), phase i^k from the fused-phase
//              identity, the coefficient by the reference's own complex or
//              real arithmetic; the commutator keeps anticommuting pairs, the
//              anticommutator pairs with an even phase.  Two passes (block
//              counts, then stable writes), as kw_rotate.
//   ks_*       the windowed path's stable LSD radix sort of (key, index)
//   ka_runs    equal-key runs summed left to right by the run's first thread
//              (_triples_to_live!, spv_ops.jl:229-251: zero sums are kept),
//              written in key order with the same two-pass compaction.
//
// Coefficients are complex (separate real and imaginary arrays) or real.
// The reference sums a run in the order its unstable quicksort leaves the
// pairs; here a run is summed in generation order (the sort is stable).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_common.hpp"
#include "pp_device.hpp"
#include "pp_window.hpp"

namespace pp {

enum { kAlgMul = 0, kAlgComm = 1, kAlgAnti = 2 };

// One operand on the device.
struct AlgOperand {
    const uint64_t* z;
    const uint64_t* x;
    const double* re;
    const double* im;   // unused for real coefficients
    int64_t n;
    int conj;           // adjoint (A†: conjugated coefficients, same basis)
};

// Product of terms a·b (ca, cb already conjugated where asked).  Returns
// false for a pair the operation skips.  Arithmetic follows the reference
// expression by expression, so results agree to the bit:
//   mul   PHASE_TBL[k+1] * ca * cb      (spv_ops.jl:274-276), PHASE_TBL = i^k
//   anti  (2 (1 - k)) * ca * cb         (spv_ops.jl:436-437), k even
//   comm  (2 (2 - k)) * im * ca * cb    (spv_ops.jl:439), anticommuting pairs
// evaluated left to right; a complex times a complex is
// (ar br - ai bi, ar bi + ai br), an integer or real times a complex scales
// both parts.  With real coefficients the product's imaginary part is
// returned in im (the reference's convert to a real type requires it zero).
template <int W>
__device__ __forceinline__ bool alg_pair(const Key<W>& a, double ar, double ai, const Key<W>& b, double br,
                                         double bi, int op, int cplx, Key<W>& o, double& re, double& im) {
    int n_a = 0, n_b = 0, m1 = 0, m2 = 0, n_p = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        o.z[w] = a.z[w] ^ b.z[w];
        o.x[w] = a.x[w] ^ b.x[w];
        n_a += popc64(a.z[w] & a.x[w]);
        n_b += popc64(b.z[w] & b.x[w]);
        m1 += popc64(a.x[w] & b.z[w]);
        m2 += popc64(a.z[w] & b.x[w]);
        n_p += popc64(o.z[w] & o.x[w]);
    }
    if (op == kAlgComm && ((m1 - m2) & 1) == 0) return false;  // commuting pair
    const int k = (n_p - n_a - n_b + 2 * m1) & 3;               // mod(., 4)
    if (op == kAlgAnti && (k & 1)) return false;
    if (op == kAlgMul) {
        const double pr = k == 0 ? 1.0 : (k == 2 ? -1.0 : 0.0);
        const double pi = k == 1 ? 1.0 : (k == 3 ? -1.0 : 0.0);
        if (cplx) {
            const double zr = pr * ar - pi * ai, zi = pr * ai + pi * ar;
            re = zr * br - zi * bi;
            im = zr * bi + zi * br;
        } else {
            re = (pr * ar) * br;
            im = (pi * ar) * br;
        }
    } else if (op == kAlgAnti) {
        const double f = (double)(2 * (1 - k));
        if (cplx) {
            const double zr = f * ar, zi = f * ai;
            re = zr * br - zi * bi;
            im = zr * bi + zi * br;
        } else {
            re = (f * ar) * br;
            im = 0.0;
        }
    } else {
        // (f * im) is the integer complex (0, f); promoted and multiplied by ca
        const double f = (double)(2 * (2 - k));
        const double zr = 0.0 * ar - f * ai, zi = 0.0 * ai + f * ar;
        re = zr * br - zi * bi;
        im = zr * bi + zi * br;
    }
    return true;
}

// Pass 1 (count_only = 1): kept pairs per block of kRB pair indices.  Pass 2:
// the kept pairs' keys at their stable positions, with the position itself
// as the sort payload, and the coefficients at the same positions.  err bit
// 16: a real-coefficient product with a nonzero imaginary part.
template <int W>
__global__ void __launch_bounds__(kBlock) ka_pairs(AlgOperand A, AlgOperand B, int op, int cplx, int count_only,
                                                   int64_t* __restrict__ bcount, const int64_t* __restrict__ boff,
                                                   uint64_t* __restrict__ Oz, uint64_t* __restrict__ Ox,
                                                   double* __restrict__ Oidx, double* __restrict__ Ore,
                                                   double* __restrict__ Oim, int32_t* __restrict__ err) {
    __shared__ int s_w[4];
    const int64_t total = A.n * B.n;
    const int64_t base = (int64_t)blockIdx.x * kRB;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    int64_t run = count_only ? 0 : boff[blockIdx.x];
    int blockcnt = 0;
    bool inexact = false;
    for (int r = 0; r < kRB / kBlock; ++r) {
        const int64_t p = base + (int64_t)r * kBlock + tid;
        bool keep = false;
        Key<W> o;
        double re = 0.0, im = 0.0;
        if (p < total) {
            const int64_t i = p / B.n, j = p - i * B.n;
            const Key<W> ka = load_key<W>(A.z, A.x, i);
            const Key<W> kb = load_key<W>(B.z, B.x, j);
            const double ar = A.re[i], br = B.re[j];
            double ai = cplx ? A.im[i] : 0.0, bi = cplx ? B.im[j] : 0.0;
            if (A.conj) ai = -ai;
            if (B.conj) bi = -bi;
            keep = alg_pair<W>(ka, ar, ai, kb, br, bi, op, cplx, o, re, im);
            if (keep && !cplx && im != 0.0) inexact = true;
        }
        const uint64_t mask = __ballot(keep);
        const int before = __popcll(mask & ((lane ? (~0ull >> (64 - lane)) : 0ull)));
        if (lane == 0) s_w[wave] = __popcll(mask);
        __syncthreads();
        int wpre = 0, rtot = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) wpre += s_w[w];
            rtot += s_w[w];
        }
        if (!count_only && keep) {
            const int64_t dst = run + wpre + before;
            store_key<W>(Oz, Ox, dst, o);
            Oidx[dst] = (double)dst;
            Ore[dst] = re;
            if (cplx) Oim[dst] = im;
        }
        run += rtot;
        blockcnt += rtot;
        __syncthreads();
    }
    if (count_only && tid == 0) bcount[blockIdx.x] = blockcnt;
    if (!count_only && inexact) atomicOr(err, 16);
}

// Runs of equal keys in the sorted pairs (keys S, payload = index into the
// coefficient arrays).  Pass 1: run heads per block.  Pass 2: each head sums
// its run left to right (reading past the block where the run continues)
// and writes (key, sum) at its stable position.
template <int W>
__global__ void __launch_bounds__(kBlock) ka_runs(const uint64_t* __restrict__ Sz, const uint64_t* __restrict__ Sx,
                                                  const double* __restrict__ Sidx, int64_t m,
                                                  const double* __restrict__ Pre, const double* __restrict__ Pim,
                                                  int cplx, int count_only, int64_t* __restrict__ bcount,
                                                  const int64_t* __restrict__ boff, uint64_t* __restrict__ Oz,
                                                  uint64_t* __restrict__ Ox, double* __restrict__ Ore,
                                                  double* __restrict__ Oim) {
    __shared__ int s_w[4];
    const int64_t base = (int64_t)blockIdx.x * kRB;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    int64_t run = count_only ? 0 : boff[blockIdx.x];
    int blockcnt = 0;
    for (int r = 0; r < kRB / kBlock; ++r) {
        const int64_t p = base + (int64_t)r * kBlock + tid;
        bool head = false;
        Key<W> k;
        if (p < m) {
            k = load_key<W>(Sz, Sx, p);
            head = p == 0 || !key_eq<W>(load_key<W>(Sz, Sx, p - 1), k);
        }
        const uint64_t mask = __ballot(head);
        const int before = __popcll(mask & ((lane ? (~0ull >> (64 - lane)) : 0ull)));
        if (lane == 0) s_w[wave] = __popcll(mask);
        __syncthreads();
        int wpre = 0, rtot = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) wpre += s_w[w];
            rtot += s_w[w];
        }
        if (!count_only && head) {
            const int64_t q0 = (int64_t)Sidx[p];
            double sr = Pre[q0];
            double si = cplx ? Pim[q0] : 0.0;
            for (int64_t q = p + 1; q < m; ++q) {
                if (!key_eq<W>(load_key<W>(Sz, Sx, q), k)) break;
                const int64_t qi = (int64_t)Sidx[q];
                sr = sr + Pre[qi];
                if (cplx) si = si + Pim[qi];
            }
            const int64_t dst = run + wpre + before;
            store_key<W>(Oz, Ox, dst, k);
            Ore[dst] = sr;
            if (cplx) Oim[dst] = si;
        }
        run += rtot;
        blockcnt += rtot;
        __syncthreads();
    }
    if (count_only && tid == 0) bcount[blockIdx.x] = blockcnt;
}

}  // namespace pp
