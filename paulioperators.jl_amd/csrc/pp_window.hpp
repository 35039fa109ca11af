// Kernels of the windowed evolution path (window > 1, spv_evolve.jl:312-387):
// rotations sweep the live terms and the pending (unsorted) appends and
// append children; every `window` rotations the appends are sorted and
// merged into the live set with duplicate runs summed.
//
//   kw_rotate     flags + per-block counts over [live ; appends]; cos-scales
//                 anticommuting terms in place and writes children at the
//                 end of the append buffer in the reference's order (children
//                 of live terms in live order, then of old appends in append
//                 order: spv_evolve.jl:207-218)
//   ks_*          stable LSD radix sort of (key, c) records, 8-bit digits,
//                 digits constant over the whole buffer skipped
//   kw_merge      merge-path merge of sorted live + sorted appends; ties take
//                 the live term first, then every equal append left to right
//                 (spv_kernels.jl:335-357); strict filter on every output;
//                 decoupled look-back compaction
//   kq_*          radix select of the k-th smallest |c| (AdaptiveTruncation,
//                 spv_evolve.jl:103-114)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_common.hpp"
#include "pp_device.hpp"

namespace pp {

constexpr int kRB = 2048;   // radix-sort / rotate block (elements)
constexpr int kRadix = 256;

// ---------------------------------------------------------------------------
// Generic int64 exclusive scan over `count` values: pass 1 block sums, pass 2
// scans the block sums (one workgroup), pass 3 adds them back.
__global__ void __launch_bounds__(kBlock) kx_scan_blocks(const int64_t* __restrict__ in, int64_t count,
                                                         int64_t* __restrict__ out, int64_t* __restrict__ bsum) {
    __shared__ int64_t s_w[4];
    const int64_t base = (int64_t)blockIdx.x * kBlock * 8;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    int64_t v[8];
    int64_t tot = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t i = base + (int64_t)tid * 8 + k;
        v[k] = i < count ? in[i] : 0;
        tot += v[k];
    }
    int64_t incl = tot;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int64_t u = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += u;
    }
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    int64_t pre = 0;
    for (int w = 0; w < wave; ++w) pre += s_w[w];
    int64_t run = pre + incl - tot;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t i = base + (int64_t)tid * 8 + k;
        if (i < count) out[i] = run;
        run += v[k];
    }
    if (tid == kBlock - 1) bsum[blockIdx.x] = pre + incl;
}

__global__ void __launch_bounds__(1024) kx_scan_sums(int64_t* __restrict__ bsum, int64_t nb, int64_t* __restrict__ total) {
    __shared__ int64_t s_w[16];
    __shared__ int64_t s_carry;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + tid;
        const int64_t v = i < nb ? bsum[i] : 0;
        int64_t incl = v;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const int64_t u = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += u;
        }
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        int64_t pre = 0;
        for (int w = 0; w < wave; ++w) pre += s_w[w];
        const int64_t carry = s_carry;
        if (i < nb) bsum[i] = carry + pre + incl - v;
        __syncthreads();
        if (tid == 1023) s_carry = carry + pre + incl;
        __syncthreads();
    }
    if (tid == 0 && total) *total = s_carry;
}

__global__ void __launch_bounds__(kBlock) kx_scan_add(int64_t* __restrict__ out, int64_t count,
                                                      const int64_t* __restrict__ bsum) {
    const int64_t base = (int64_t)blockIdx.x * kBlock * 8;
    const int64_t add = bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t i = base + (int64_t)threadIdx.x * 8 + k;
        if (i < count) out[i] += add;
    }
}

// ---------------------------------------------------------------------------
// Rotation over [live (n) ; appends (an)].  Pass 1 (count=1): per-block number
// of children kept by the local filter.  Pass 2 (count=0): given the block
// offsets, cos-scales anticommuting terms in place and writes the children
// at appends[an + offset].
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) kw_rotate(uint64_t* __restrict__ Lz, uint64_t* __restrict__ Lx,
                                                    CT* __restrict__ Lc, int64_t n, uint64_t* __restrict__ Az,
                                                    uint64_t* __restrict__ Ax, CT* __restrict__ Ac, int64_t an,
                                                    const GenDev* __restrict__ g, const FilterDev* __restrict__ floc,
                                                    int count_only, int64_t* __restrict__ bcount,
                                                    const int64_t* __restrict__ boff, int64_t cap_app) {
    __shared__ int s_w[4];
    const int64_t total = n + an;
    const int64_t base = (int64_t)blockIdx.x * kRB;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    GenReg<W> G;
    G.load(g);
    int64_t run = count_only ? 0 : boff[blockIdx.x];
    int blockcnt = 0;
    for (int r = 0; r < kRB / kBlock; ++r) {
        const int64_t i = base + (int64_t)r * kBlock + tid;
        bool keep = false;
        Key<W> k, ch;
        CT cn = coef_zero<CT>();
        if (i < total) {
            const bool live = i < n;
            const int64_t q = live ? i : i - n;
            k = live ? load_key<W>(Lz, Lx, q) : load_key<W>(Az, Ax, q);
            if (G.anti(k)) {
                CT* cp = live ? Lc + q : Ac + q;
                const CT c = *cp;
                cn = G.child(k, c, ch);
                if (!count_only) *cp = c * G.cosv;
                keep = !should_drop<W>(floc, ch, coef_abs(cn));
            }
        }
        // ordered offsets inside the block: waves in order, lanes in order
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
            const int64_t dst = an + run + wpre + before;
            if (dst < cap_app) {
                store_key<W>(Az, Ax, dst, ch);
                Ac[dst] = cn;
            }
        }
        run += rtot;
        blockcnt += rtot;
        __syncthreads();
    }
    if (count_only && tid == 0) bcount[blockIdx.x] = blockcnt;
}

// ---------------------------------------------------------------------------
// Radix sort of (key, c) records; digit d = bits [8 dd, 8 dd + 8) of the
// combined key (x words low, z words high), least significant first.
template <int W>
__device__ __forceinline__ int key_digit(const Key<W>& k, int dd) {
    const int bit = dd * 8;
    return (int)((key_word<W>(k, bit >> 6) >> (bit & 63)) & 0xffull);
}

// OR and AND of every key word (constant digits are skipped), in two
// stages without atomics (same-address atomics from every wave of a large
// grid serialise): per-block partials part[block][4 W], then ks_orand_fin.
constexpr int kOrandBlocks = 1024;
template <int W>
__global__ void __launch_bounds__(kBlock) ks_orand(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X,
                                                   int64_t m, unsigned long long* __restrict__ part) {
    __shared__ unsigned long long s_p[kBlock / kWave][4 * W];
    uint64_t o[2 * W], a[2 * W];
#pragma unroll
    for (int w = 0; w < 2 * W; ++w) {
        o[w] = 0;
        a[w] = ~0ull;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const Key<W> k = load_key<W>(Z, X, i);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            o[w] |= k.x[w];
            a[w] &= k.x[w];
            o[W + w] |= k.z[w];
            a[W + w] &= k.z[w];
        }
    }
#pragma unroll
    for (int w = 0; w < 2 * W; ++w) {
#pragma unroll
        for (int dd = kWave / 2; dd > 0; dd >>= 1) {
            o[w] |= __shfl_xor(o[w], dd, kWave);
            a[w] &= __shfl_xor(a[w], dd, kWave);
        }
    }
    const int wave = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
#pragma unroll
        for (int w = 0; w < 2 * W; ++w) {
            s_p[wave][w] = o[w];
            s_p[wave][2 * W + w] = a[w];
        }
    }
    __syncthreads();
    if (threadIdx.x < 4 * W) {
        const int k = threadIdx.x;
        unsigned long long v = s_p[0][k];
        for (int q = 1; q < kBlock / kWave; ++q) v = k < 2 * W ? (v | s_p[q][k]) : (v & s_p[q][k]);
        part[(int64_t)blockIdx.x * 4 * W + k] = v;
    }
}

// One workgroup: acc[0, 2W) = OR and acc[2W, 4W) = AND over nb partials.
template <int W>
__global__ void __launch_bounds__(kBlock) ks_orand_fin(const unsigned long long* __restrict__ part, int nb,
                                                       unsigned long long* __restrict__ acc) {
    __shared__ unsigned long long s_p[kBlock / kWave][4 * W];
    unsigned long long v[4 * W];
#pragma unroll
    for (int k = 0; k < 4 * W; ++k) v[k] = k < 2 * W ? 0ull : ~0ull;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
        for (int k = 0; k < 4 * W; ++k) {
            const unsigned long long x = part[(int64_t)b * 4 * W + k];
            v[k] = k < 2 * W ? (v[k] | x) : (v[k] & x);
        }
    }
#pragma unroll
    for (int k = 0; k < 4 * W; ++k) {
#pragma unroll
        for (int dd = kWave / 2; dd > 0; dd >>= 1) {
            const unsigned long long y = __shfl_xor(v[k], dd, kWave);
            v[k] = k < 2 * W ? (v[k] | y) : (v[k] & y);
        }
    }
    const int wave = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
#pragma unroll
        for (int k = 0; k < 4 * W; ++k) s_p[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 4 * W) {
        const int k = threadIdx.x;
        unsigned long long r = s_p[0][k];
        for (int q = 1; q < kBlock / kWave; ++q) r = k < 2 * W ? (r | s_p[q][k]) : (r & s_p[q][k]);
        acc[k] = r;
    }
}

// Per-block digit histogram, digit-major: hist[digit * nblocks + block].
template <int W>
__global__ void __launch_bounds__(kBlock) ks_hist(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X,
                                                  int64_t m, int dd, int64_t* __restrict__ hist, int64_t nblocks) {
    __shared__ int s_h[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kBlock) s_h[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRB;
    for (int r = 0; r < kRB / kBlock; ++r) {
        const int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
        if (i < m) atomicAdd(&s_h[key_digit<W>(load_key<W>(Z, X, i), dd)], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadix; i += kBlock) hist[(int64_t)i * nblocks + blockIdx.x] = s_h[i];
}

// Stable scatter: within a block, elements keep their order per digit
// (waves in order, lanes in order, rounds in order).
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) ks_scatter(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X,
                                                     const CT* __restrict__ C, int64_t m, int dd,
                                                     const int64_t* __restrict__ off, int64_t nblocks,
                                                     uint64_t* __restrict__ OZ, uint64_t* __restrict__ OX,
                                                     CT* __restrict__ OC) {
    __shared__ int64_t s_base[kRadix];
    __shared__ int s_cnt[4][kRadix];
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    for (int i = tid; i < kRadix; i += kBlock) s_base[i] = off[(int64_t)i * nblocks + blockIdx.x];
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRB;
    for (int r = 0; r < kRB / kBlock; ++r) {
        const int64_t i = base + (int64_t)r * kBlock + tid;
        const bool valid = i < m;
        Key<W> k;
        CT c = coef_zero<CT>();
        int dg = -1;
        if (valid) {
            k = load_key<W>(Z, X, i);
            c = C[i];
            dg = key_digit<W>(k, dd);
        }
        // lanes of this wave with the same digit (8 ballots)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t bm = __ballot(valid && ((dg >> bit) & 1));
            peers &= ((dg >> bit) & 1) ? bm : ~bm;
        }
        if (!valid) peers = 0;
        const int rank_in_wave = __popcll(peers & (lane ? (~0ull >> (64 - lane)) : 0ull));
        for (int i2 = tid; i2 < 4 * kRadix; i2 += kBlock) (&s_cnt[0][0])[i2] = 0;
        __syncthreads();
        // the lowest lane of each digit group publishes the group size
        if (valid && rank_in_wave == 0) s_cnt[wave][dg] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pre = 0;
            for (int w = 0; w < wave; ++w) pre += s_cnt[w][dg];
            const int64_t dst = s_base[dg] + pre + rank_in_wave;
            store_key<W>(OZ, OX, dst, k);
            OC[dst] = c;
        }
        __syncthreads();
        // advance the per-digit bases by this round's counts
        for (int d2 = tid; d2 < kRadix; d2 += kBlock)
            s_base[d2] += s_cnt[0][d2] + s_cnt[1][d2] + s_cnt[2][d2] + s_cnt[3][d2];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// kw_merge: sorted live L (unique keys) + sorted appends C (runs of equal
// keys) -> O.  A run's value is L's coefficient (if present) plus the run's
// appends left to right; the thread that holds the run's first element sums
// it (reading past its range when needed).
template <int W, int TM, typename CT>
__global__ void __launch_bounds__(kBlock) kw_merge(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                   const CT* __restrict__ Lc, int64_t n,
                                                   const uint64_t* __restrict__ Cz, const uint64_t* __restrict__ Cx,
                                                   const CT* __restrict__ Cc, int64_t m,
                                                   const FilterDev* __restrict__ f, uint64_t* __restrict__ Oz,
                                                   uint64_t* __restrict__ Ox, CT* __restrict__ Oc, int64_t capOut,
                                                   uint64_t* __restrict__ status, uint32_t* __restrict__ ticket,
                                                   uint32_t epoch, int64_t* __restrict__ d_nout,
                                                   int32_t* __restrict__ d_err) {
    constexpr int E = TM / kBlock;
    __shared__ Key<W> s_key[TM];
    __shared__ CT s_c[TM];
    __shared__ int64_t s_tile;
    __shared__ int64_t s_base;
    __shared__ int64_t s_split[2];
    __shared__ int s_wsum[4];
    const int64_t tot = n + m;
    const int64_t ntm = (tot + TM - 1) / TM;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const uint64_t ep = ((uint64_t)epoch & kEpochMask) << kEpochShift;
    if (ntm == 0) {
        if (blockIdx.x == 0 && tid == 0) *d_nout = 0;
        return;
    }
    while (true) {
        if (tid == 0) s_tile = (int64_t)atomicAdd(ticket, 1u);
        __syncthreads();
        const int64_t t = s_tile;
        if (t >= ntm) break;
        const int64_t d0 = t * TM;
        const int64_t d1 = d0 + TM < tot ? d0 + TM : tot;
        if (wave < 2) {
            const int64_t sp = merge_split_wave<W, (TM >= 1024)>(Lz, Lx, n, Cz, Cx, m, wave == 0 ? d0 : d1);
            if (lane == 0) s_split[wave] = sp;
        }
        __syncthreads();
        int64_t i0 = s_split[0], i1 = s_split[1];
        int64_t j0 = d0 - i0, j1 = d1 - i1;
        if (i0 < 0 || i1 < i0 || i1 > n || j0 < 0 || j1 < j0 || j1 > m) {
            if (tid == 0) atomicOr(d_err, 2);
            i0 = i1 = j0 = j1 = 0;
        }
        const int nL = (int)(i1 - i0), nC = (int)(j1 - j0);
        {
            // every load of the tile first (one memory round trip), then LDS
            Key<W> lk[E];
            CT lcv[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = tid + e * kBlock;
#pragma unroll
                for (int w = 0; w < W; ++w) lk[e].z[w] = lk[e].x[w] = 0;
                lcv[e] = coef_zero<CT>();
                if (k < nL) {
                    lk[e] = load_key<W>(Lz, Lx, i0 + k);
                    lcv[e] = ldg(Lc, i0 + k);
                } else if (k < nL + nC) {
                    lk[e] = load_key<W>(Cz, Cx, j0 + (k - nL));
                    lcv[e] = ldg(Cc, j0 + (k - nL));
                }
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = tid + e * kBlock;
                if (k < nL + nC) {
                    s_key[k] = lk[e];
                    s_c[k] = lcv[e];
                }
            }
        }
        __syncthreads();
        const Key<W>* sL = s_key;
        const Key<W>* sC = s_key + nL;
        const CT* sLc = s_c;
        const CT* sCc = s_c + nL;
        const int dk = tid * E < nL + nC ? tid * E : nL + nC;
        int lo = dk - nC > 0 ? dk - nC : 0;
        int hi = dk < nL ? dk : nL;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (key_le<W>(sL[mid], sC[dk - 1 - mid])) lo = mid + 1;
            else hi = mid;
        }
        int ia = lo, jb = dk - lo;
        // outputs in registers indexed by the unrolled position (a run-time
        // index would place the arrays in scratch memory)
        Key<W> outK[E];
        CT outC[E];
        bool outE[E];
        const int stop = dk + E < nL + nC ? dk + E : nL + nC;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            outE[e] = false;
#pragma unroll
            for (int w = 0; w < W; ++w) outK[e].z[w] = outK[e].x[w] = 0;
            outC[e] = coef_zero<CT>();
            if (dk + e >= stop) continue;
            const bool takeL = ia < nL && (jb >= nC || key_le<W>(sL[ia], sC[jb]));
            Key<W> K;
            CT acc;
            bool head = true;
            int64_t jg;  // global index of the next append after this element
            if (takeL) {
                K = sL[ia];
                acc = sLc[ia];
                ++ia;
                jg = j0 + jb;
            } else {
                K = sC[jb];
                acc = sCc[jb];
                // continuation of a run begun by the previous merged element?
                if (jb > 0) {
                    if (key_eq<W>(sC[jb - 1], K)) head = false;
                } else if (j0 > 0) {
                    if (key_eq<W>(load_key<W>(Cz, Cx, j0 - 1), K)) head = false;
                }
                if (head) {
                    if (ia > 0) {
                        if (key_eq<W>(sL[ia - 1], K)) head = false;
                    } else if (i0 > 0) {
                        if (key_eq<W>(load_key<W>(Lz, Lx, i0 - 1), K)) head = false;
                    }
                }
                ++jb;
                jg = j0 + jb;
            }
            if (!head) continue;
            // sum the rest of the run, left to right
            while (jg < m) {
                const int lj = (int)(jg - j0);
                const bool inl = lj < nC;
                const Key<W> kn = inl ? sC[lj] : load_key<W>(Cz, Cx, jg);
                if (!key_eq<W>(kn, K)) break;
                acc = acc + (inl ? sCc[lj] : Cc[jg]);
                ++jg;
            }
            outK[e] = K;
            outC[e] = acc;
            outE[e] = !should_drop<W>(f, K, coef_abs(acc));
        }
        int cnt = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) cnt += outE[e] ? 1 : 0;
        int incl = cnt;
#pragma unroll
        for (int dd = 1; dd < kWave; dd <<= 1) {
            const int u = __shfl_up(incl, dd, kWave);
            if (lane >= dd) incl += u;
        }
        if (lane == kWave - 1) s_wsum[wave] = incl;
        __syncthreads();
        int wpref = 0, total = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) wpref += s_wsum[w];
            total += s_wsum[w];
        }
        const int myoff = wpref + incl - cnt;
        if (wave == 0) {
            if (t == 0) {
                if (lane == 0) {
                    __hip_atomic_store(&status[0], kFlagPre | ep | (uint64_t)total, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    s_base = 0;
                }
            } else {
                if (lane == 0)
                    __hip_atomic_store(&status[t], kFlagAgg | ep | (uint64_t)total, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                int64_t excl = 0, top = t - 1, spins = 0;
                while (true) {
                    const int64_t q = top - lane;
                    const uint64_t sv = q >= 0 ? __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                               : (kFlagPre | ep);
                    const bool fresh = (sv & (kEpochMask << kEpochShift)) == ep;
                    const uint64_t flag = fresh ? (sv & (3ull << 62)) : 0ull;
                    const uint64_t pre = __ballot(flag == kFlagPre);
                    const uint64_t notready = __ballot(flag == 0);
                    const int first = pre ? __ffsll((unsigned long long)pre) - 1 : kWave;
                    const uint64_t need = first >= kWave - 1 ? ~0ull : ((1ull << (first + 1)) - 1);
                    if (notready & need) {
                        if (++spins > (1ll << 26)) {
                            if (lane == 0) atomicOr(d_err, 4);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    int64_t v = lane <= first ? (int64_t)(sv & kCountMask) : 0;
#pragma unroll
                    for (int dd = kWave / 2; dd > 0; dd >>= 1) v += __shfl_xor(v, dd, kWave);
                    excl += v;
                    if (first < kWave) break;
                    top -= kWave;
                }
                if (lane == 0) {
                    __hip_atomic_store(&status[t], kFlagPre | ep | (uint64_t)(excl + total), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    s_base = excl;
                }
            }
            if (lane == 0 && t == ntm - 1) *d_nout = s_base + total;
        }
        __syncthreads();
        const int64_t base = s_base;
        {
            int k = myoff;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (outE[e]) {
                    s_key[k] = outK[e];
                    s_c[k] = outC[e];
                    ++k;
                }
            }
        }
        __syncthreads();
        if (base + total <= capOut) {
            for (int k = tid; k < total; k += kBlock) {
                store_key<W>(Oz, Ox, base + k, s_key[k]);
                Oc[base + k] = s_c[k];
            }
        } else if (tid == 0) {
            atomicOr(d_err, 8);  // output overflow: the host sized the buffer for n + m
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Radix select over |c| bit patterns (non-negative doubles order as uint64):
// histogram of one byte among the values whose higher bytes equal `prefix`.
template <typename CT>
__global__ void __launch_bounds__(kBlock) kq_hist(const CT* __restrict__ C, int64_t n, int shift, uint64_t prefix,
                                                  uint64_t pmask, unsigned long long* __restrict__ hist) {
    __shared__ int s_h[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kBlock) s_h[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t u = (uint64_t)__double_as_longlong(coef_abs(C[i]));
        if ((u & pmask) == prefix) atomicAdd(&s_h[(int)((u >> shift) & 0xffull)], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadix; i += kBlock)
        if (s_h[i]) atomicAdd(hist + i, (unsigned long long)s_h[i]);
}

// ---------------------------------------------------------------------------
// Order-preserving compaction of one buffer into another (spv_kernels.jl:
// 395-408 and stochastic_clip!, spv_evolve.jl:35-52).  Pass 1 (count_only)
// writes per-block survivor counts; pass 2 scatters with the block offsets.
//   mode 0: keep unless should_drop(f, key, |c|)
//   mode 1: stochastic clip: |c| < eps survives with probability |c| / eps
//           and becomes eps * sign(c); the uniform variate of term i is
//           counter_uniform(seed, i)
//   mode 2: sampling selection (SampleSel); mode 3: delete marked terms
//   mode 4: diag (x == 0, spv_ops.jl:571-587); mode 5: offdiag (x != 0,
//           spv_ops.jl:555-569)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    // SplitMix64 finaliser (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double counter_uniform(uint64_t seed, uint64_t i) {
    const uint64_t h = mix64(seed ^ mix64(i + 0x9e3779b97f4a7c15ull));
    return (double)(h >> 11) * 0x1.0p-53;
}

// Selection of mode 2: the sampling key of term i (monotone encoding) and the
// threshold pair: keep when (key, n - 1 - i) >= (eT, jT) lexicographically.
struct SampleSel {
    const uint64_t* key;
    uint64_t eT;
    int64_t jT;
    int64_t n;
    const uint8_t* mark;  // mode 3: terms to delete
};

template <int W, typename CT>
__device__ __forceinline__ bool compact_keep(int mode, const FilterDev* f, double eps, uint64_t seed, int64_t i,
                                             const Key<W>& k, CT& c, const SampleSel& ss) {
    if (mode == 0) return !should_drop<W>(f, k, coef_abs(c));
    if (mode == 2) {
        const uint64_t e = ss.key[i];
        return e > ss.eT || (e == ss.eT && ss.n - 1 - i >= ss.jT);
    }
    if (mode == 3) return ss.mark[i] == 0;
    if (mode == 4 || mode == 5) {
        uint64_t xo = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) xo |= k.x[w];
        return (xo == 0) == (mode == 4);
    }
    const double ac = coef_abs(c);
    if (ac < eps) {
        if (!(counter_uniform(seed, (uint64_t)i) < ac / eps)) return false;
        c = eps * (c / ac);
    }
    return true;
}

template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) kw_compact(const uint64_t* __restrict__ Iz, const uint64_t* __restrict__ Ix,
                                                     const CT* __restrict__ Ic, int64_t n,
                                                     const FilterDev* __restrict__ f, int mode, double eps,
                                                     uint64_t seed, int count_only, int64_t* __restrict__ bcount,
                                                     const int64_t* __restrict__ boff, uint64_t* __restrict__ Oz,
                                                     uint64_t* __restrict__ Ox, CT* __restrict__ Oc, SampleSel ss) {
    __shared__ int s_w[4];
    const int64_t base = (int64_t)blockIdx.x * kRB;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    int64_t run = count_only ? 0 : boff[blockIdx.x];
    int blockcnt = 0;
    for (int r = 0; r < kRB / kBlock; ++r) {
        const int64_t i = base + (int64_t)r * kBlock + tid;
        bool keep = false;
        Key<W> k;
        CT c = coef_zero<CT>();
        if (i < n) {
            k = load_key<W>(Iz, Ix, i);
            c = Ic[i];
            keep = compact_keep<W, CT>(mode, f, eps, seed, i, k, c, ss);
        }
        const uint64_t mask = __ballot(keep);
        const int before = __popcll(mask & (lane ? (~0ull >> (64 - lane)) : 0ull));
        if (lane == 0) s_w[wave] = __popcll(mask);
        __syncthreads();
        int wpre = 0, rtot = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) wpre += s_w[w];
            rtot += s_w[w];
        }
        if (!count_only && keep) {
            const int64_t dst = run + wpre + before;
            store_key<W>(Oz, Ox, dst, k);
            Oc[dst] = c;
        }
        run += rtot;
        blockcnt += rtot;
        __syncthreads();
    }
    if (count_only && tid == 0) bcount[blockIdx.x] = blockcnt;
}

// StochasticSamplingTruncation keys (spv_evolve.jl:69-99): term i gets
// log(u_i) / |c_i|^2, which orders the terms as u_i^(1/|c_i|^2) does (zero
// weight: -inf), encoded as an unsigned integer with the same order.
template <typename CT>
__global__ void __launch_bounds__(kBlock) kw_sample_keys(const CT* __restrict__ C, int64_t n, uint64_t seed,
                                                         uint64_t* __restrict__ key) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double w = coef_abs2(C[i]);
        const double u = counter_uniform(seed, (uint64_t)i);
        const double k = w > 0.0 ? log(u) / w : -INFINITY;
        const uint64_t b = (uint64_t)__double_as_longlong(k);
        key[i] = (b >> 63) ? ~b : (b | (1ull << 63));
    }
}

// Byte histogram for the radix select over (key, n - 1 - i): stage 0 bins
// key byte `shift` of the keys matching `prefix` on `pmask`; stage 1 bins byte
// `shift` of n - 1 - i among the terms whose key equals eT.
__global__ void __launch_bounds__(kBlock) kq_hist_sel(const uint64_t* __restrict__ key, int64_t n, int stage,
                                                      uint64_t eT, int shift, uint64_t prefix, uint64_t pmask,
                                                      unsigned long long* __restrict__ hist) {
    __shared__ int s_h[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kBlock) s_h[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = key[i];
        if (stage == 0) {
            if ((e & pmask) == prefix) atomicAdd(&s_h[(int)((e >> shift) & 0xffull)], 1);
        } else if (e == eT) {
            const uint64_t j = (uint64_t)(n - 1 - i);
            if ((j & pmask) == prefix) atomicAdd(&s_h[(int)((j >> shift) & 0xffull)], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadix; i += kBlock)
        if (s_h[i]) atomicAdd(hist + i, (unsigned long long)s_h[i]);
}

__global__ void __launch_bounds__(kBlock) kw_scale(double* __restrict__ C, int64_t n, double f) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) C[i] *= f;
}

// mul!(v, a) with a ComplexF64 scalar on ComplexF64 coefficients
// (spv_ops.jl:12-17): Julia's Complex * Complex, (c.re a.re - c.im a.im,
// c.re a.im + c.im a.re).  A real scalar scales componentwise (kw_scale over
// 2n values), as Complex * Real does.
__global__ void __launch_bounds__(kBlock) kw_scale_c(pp_cplx* __restrict__ C, int64_t n, double ar, double ai) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const pp_cplx c = C[i];
        pp_cplx r;
        r.x = c.x * ar - c.y * ai;
        r.y = c.x * ai + c.y * ar;
        C[i] = r;
    }
}

// <ket|O|ket> partial sums over [A (na) ; B (nb)] (live then pending appends,
// spv_kernels.jl:441-452), double-double per block.
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) kw_expect(const uint64_t* __restrict__ Az, const uint64_t* __restrict__ Ax,
                                                    const CT* __restrict__ Ac, int64_t na,
                                                    const uint64_t* __restrict__ Bz, const uint64_t* __restrict__ Bx,
                                                    const CT* __restrict__ Bc, int64_t nb, Key<W> ket,
                                                    double2* __restrict__ partial, int comp) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    double h = 0.0, l = 0.0;
    const int64_t tot = na + nb;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < tot; i += (int64_t)gridDim.x * kBlock) {
        const bool inA = i < na;
        const int64_t q = inA ? i : i - na;
        const Key<W> k = inA ? load_key<W>(Az, Ax, q) : load_key<W>(Bz, Bx, q);
        bool diag = true;
        int par = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            diag = diag && k.x[w] == 0;
            par += popc64(k.z[w] & ket.z[w]);
        }
        if (diag) {
            const CT cv = inA ? Ac[q] : Bc[q];
            const double c = comp ? coef_im(cv) : coef_re(cv);
            dd_add(h, l, (par & 1) ? -c : c, 0.0);
        }
    }
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

// variance(O, ψ) of the live set (statistics.jl:6-22), step 1: the terms
// with their words swapped (x first), so a sort by key orders them by x and,
// within one x, by z (the operator's iteration order).
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) kv_swap(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                  const CT* __restrict__ Lc, int64_t n, uint64_t* __restrict__ Az,
                                                  uint64_t* __restrict__ Ax, CT* __restrict__ Ac) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            Az[i * W + w] = Lx[i * W + w];
            Ax[i * W + w] = Lz[i * W + w];
        }
        Ac[i] = Lc[i];
    }
}

// Steps 2-4, on the terms sorted by (x, z): σ[ψ ⊻ x] = Σ i^((symplectic_phase
// + 2 sign) % 4) c over each x group (multiplication.jl:48-52: sign =
// popcount(z & (x ⊻ ψ)), symplectic_phase = (4 - popcount(z & x) % 4) % 4),
// then Σ |σ|^2.  The group sums are a segmented scan over tiles of kVarTile
// terms: kv_tiles<W, false> reduces each tile to (head seen, sum since the
// tile's last group head), kv_carry scans those into the sum open at each
// tile's start, and kv_tiles<W, true> rescans each tile from its carry, adding
// |σ|^2 at the last term of every group to a double-double block partial.  The
// group x = 0 (the first, when present) also gives σ[ψ] in out[1].  The
// addition order is a fixed tree, so results repeat bit for bit.
constexpr int kVarPer = 4;                     // consecutive terms per thread
constexpr int kVarTile = kBlock * kVarPer;     // terms per tile

struct VarSeg {
    double re, im;
    int head;  // a group starts inside the segment; (re, im) is the sum since its last start
};

__device__ __forceinline__ VarSeg var_join(const VarSeg& a, const VarSeg& b) {
    return b.head ? b : VarSeg{a.re + b.re, a.im + b.im, a.head};
}

template <int W>
__device__ __forceinline__ bool var_same_x(const uint64_t* __restrict__ Sz, int64_t i, int64_t j) {
    bool same = true;
#pragma unroll
    for (int w = 0; w < W; ++w) same = same && Sz[i * W + w] == Sz[j * W + w];
    return same;
}

// Term j's phased coefficient; Sz holds the x words, Sx the z words.
template <int W, typename CT>
__device__ __forceinline__ VarSeg var_term(const uint64_t* __restrict__ Sz, const uint64_t* __restrict__ Sx,
                                           const CT* __restrict__ Sc, int64_t j, const Key<W>& ket) {
    int sgn = 0, zx = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t x = Sz[j * W + w], z = Sx[j * W + w];
        sgn += popc64(z & (x ^ ket.z[w]));
        zx += popc64(z & x);
    }
    const int ph = ((4 - zx % 4) % 4 + 2 * (sgn & 1)) % 4;  // PHASE_TBL = [1, i, -1, -i]
    const CT c = Sc[j];
    const double cr = coef_re(c), ci = coef_im(c);
    const int head = j == 0 || !var_same_x<W>(Sz, j, j - 1);
    // phase * c: (1) c, (i) (-im, re), (-1) (-re, -im), (-i) (im, -re)
    if (ph == 0) return VarSeg{cr, ci, head};
    if (ph == 1) return VarSeg{-ci, cr, head};
    if (ph == 2) return VarSeg{-cr, -ci, head};
    return VarSeg{ci, -cr, head};
}

// Exclusive scan of the threads' segments in thread order, seeded by `seed`;
// returns this thread's prefix and leaves the tile's total in `total`.
__device__ __forceinline__ VarSeg var_block_scan(VarSeg mine, const VarSeg& seed, VarSeg& total, double* s_re,
                                                 double* s_im, int* s_hd) {
    const int t = threadIdx.x;
    for (int off = 1; off < kBlock; off <<= 1) {
        s_re[t] = mine.re;
        s_im[t] = mine.im;
        s_hd[t] = mine.head;
        __syncthreads();
        if (t >= off) mine = var_join(VarSeg{s_re[t - off], s_im[t - off], s_hd[t - off]}, mine);
        __syncthreads();
    }
    s_re[t] = mine.re;
    s_im[t] = mine.im;
    s_hd[t] = mine.head;
    __syncthreads();
    const VarSeg prev = t > 0 ? VarSeg{s_re[t - 1], s_im[t - 1], s_hd[t - 1]} : VarSeg{0.0, 0.0, 0};
    total = var_join(seed, VarSeg{s_re[kBlock - 1], s_im[kBlock - 1], s_hd[kBlock - 1]});
    __syncthreads();
    return t > 0 ? var_join(seed, prev) : seed;
}

template <int W, bool kFinal, typename CT>
__global__ void __launch_bounds__(kBlock) kv_tiles(const uint64_t* __restrict__ Sz, const uint64_t* __restrict__ Sx,
                                                   const CT* __restrict__ Sc, int64_t n, Key<W> ket,
                                                   double2* __restrict__ tile_v, int32_t* __restrict__ tile_h,
                                                   const double2* __restrict__ carry, double2* __restrict__ partial,
                                                   double* __restrict__ out) {
    __shared__ double s_re[kBlock], s_im[kBlock];
    __shared__ int s_hd[kBlock];
    __shared__ double s_h[kBlock], s_l[kBlock];
    double h = 0.0, l = 0.0;
    const int64_t ntiles = (n + kVarTile - 1) / kVarTile;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t j0 = tile * kVarTile + (int64_t)threadIdx.x * kVarPer;
        VarSeg term[kVarPer];
        VarSeg mine{0.0, 0.0, 0};
#pragma unroll
        for (int e = 0; e < kVarPer; ++e) {
            term[e] = j0 + e < n ? var_term<W, CT>(Sz, Sx, Sc, j0 + e, ket) : VarSeg{0.0, 0.0, 0};
            mine = var_join(mine, term[e]);
        }
        const VarSeg seed = kFinal ? VarSeg{carry[tile].x, carry[tile].y, 0} : VarSeg{0.0, 0.0, 0};
        VarSeg total;
        VarSeg run = var_block_scan(mine, seed, total, s_re, s_im, s_hd);
        if (!kFinal) {
            if (threadIdx.x == 0) {
                tile_v[tile] = make_double2(total.re, total.im);
                tile_h[tile] = total.head;
            }
            continue;
        }
#pragma unroll
        for (int e = 0; e < kVarPer; ++e) {
            const int64_t j = j0 + e;
            if (j >= n) break;
            run = var_join(run, term[e]);
            if (j + 1 == n || !var_same_x<W>(Sz, j, j + 1)) {
                dd_add(h, l, run.re * run.re + run.im * run.im, 0.0);
                bool zero = true;
#pragma unroll
                for (int w = 0; w < W; ++w) zero = zero && Sz[j * W + w] == 0;
                if (zero) {
                    out[1] = run.re;
                    out[2] = run.im;
                }
            }
        }
    }
    if (!kFinal) return;
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

// One workgroup: carry[t] = the group sum open at tile t's start (the
// exclusive segmented scan of the tile totals); also clears out[1], which the
// final pass sets when a group x = 0 exists.
__global__ void __launch_bounds__(kBlock) kv_carry(const double2* __restrict__ tile_v,
                                                   const int32_t* __restrict__ tile_h, int64_t ntiles,
                                                   double2* __restrict__ carry, double* __restrict__ out) {
    __shared__ double s_re[kBlock], s_im[kBlock];
    __shared__ int s_hd[kBlock];
    const int64_t per = (ntiles + kBlock - 1) / kBlock;
    const int64_t t0 = (int64_t)threadIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
    VarSeg mine{0.0, 0.0, 0};
    for (int64_t t = t0; t < t1; ++t) mine = var_join(mine, VarSeg{tile_v[t].x, tile_v[t].y, tile_h[t]});
    VarSeg total;
    VarSeg run = var_block_scan(mine, VarSeg{0.0, 0.0, 0}, total, s_re, s_im, s_hd);
    for (int64_t t = t0; t < t1; ++t) {
        carry[t] = make_double2(run.re, run.im);
        run = var_join(run, VarSeg{tile_v[t].x, tile_v[t].y, tile_h[t]});
    }
    if (threadIdx.x == 0) {
        out[1] = 0.0;
        out[2] = 0.0;
    }
}

__global__ void __launch_bounds__(kBlock) kw_expect_final(const double2* __restrict__ partial, int64_t np,
                                                          double* __restrict__ out) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    double h = 0.0, l = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += kBlock) dd_add(h, l, partial[i].x, partial[i].y);
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
    if (threadIdx.x == 0) *out = s_h[0] + s_l[0];
}

}  // namespace pp
