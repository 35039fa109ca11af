// Device kernels of the SparsePauliVector engine for gfx950 (wave64).
//
// Storage is structure-of-arrays in HBM: z and x are W little-endian uint64
// words per term (W = 1 for N <= 64, 2 for N <= 128; a 16-byte vector per
// term when W = 2), coefficients are FP64.  The live set is kept sorted by
// the reference's (z, x) z-major key (spv_kernels.jl:14-16) and duplicate-free.
//
// One rotation at window = 1 (spv_evolve.jl:207-218 + 258-284):
//   k_flags     anticommutation test per term -> bit mask + per-tile counts
//   k_scan      exclusive scan of tile counts -> child offsets, m
//   k_bounds    per (tile, mask level) boundary counts that cross tile edges
//   k_emit      child (G xor P, ((k-2) sinθ) c) written straight to its rank
//               in key order (no sort pass, see below)
//   k_partition merge-path splits of live + children
//   k_merge     merge, live term first on ties (spv_kernels.jl:338-350),
//               cos-scale of anticommuting live terms, sum of the (at most
//               one) child per key, strict filter on every output
//               (spv_kernels.jl:351), stream compaction with a decoupled
//               look-back scan.
//
// Child rank.  Children of one rotation are the anticommuting live terms P
// (kept by the local filter; call that set A) mapped by P -> P xor M, where M
// is the combined (z, x) mask of G.  XOR with M swaps, at every key bit b set
// in M, the bit-0 and bit-1 halves of each trie node at that level, and the
// swaps compose additively, so
//   rank(P xor M) = #{A before P} + sum_{b in M} (bit_b(P) == 0 ? +|S'_b| : -|S'_b|)
// where S'_b is the set of A-terms sharing P's key bits above b with bit b
// flipped (P's sibling segment at level b).  The unshuffle cascade of
// spv_kernels.jl:258-320 computes the same permutation with one stable pass
// per set bit; here each child is placed once.
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

template <int W>
__device__ __forceinline__ Key<W> load_key(const uint64_t* __restrict__ Z, const uint64_t* __restrict__ X, int64_t i) {
    Key<W> k;
    if constexpr (W == 1) {
        k.z[0] = Z[i];
        k.x[0] = X[i];
    } else {
        const ulonglong2 a = reinterpret_cast<const ulonglong2*>(Z)[i];
        const ulonglong2 b = reinterpret_cast<const ulonglong2*>(X)[i];
        k.z[0] = a.x; k.z[1] = a.y;
        k.x[0] = b.x; k.x[1] = b.y;
    }
    return k;
}

template <int W>
__device__ __forceinline__ void store_key(uint64_t* __restrict__ Z, uint64_t* __restrict__ X, int64_t i, const Key<W>& k) {
    if constexpr (W == 1) {
        Z[i] = k.z[0];
        X[i] = k.x[0];
    } else {
        reinterpret_cast<ulonglong2*>(Z)[i] = make_ulonglong2(k.z[0], k.z[1]);
        reinterpret_cast<ulonglong2*>(X)[i] = make_ulonglong2(k.x[0], k.x[1]);
    }
}

// spv_kernels.jl:14-16: z-major, high word first.
template <int W>
__device__ __forceinline__ bool key_lt(const Key<W>& a, const Key<W>& b) {
#pragma unroll
    for (int w = W - 1; w >= 0; --w)
        if (a.z[w] != b.z[w]) return a.z[w] < b.z[w];
#pragma unroll
    for (int w = W - 1; w >= 0; --w)
        if (a.x[w] != b.x[w]) return a.x[w] < b.x[w];
    return false;
}
template <int W>
__device__ __forceinline__ bool key_le(const Key<W>& a, const Key<W>& b) { return !key_lt<W>(b, a); }
template <int W>
__device__ __forceinline__ bool key_eq(const Key<W>& a, const Key<W>& b) {
    bool e = true;
#pragma unroll
    for (int w = 0; w < W; ++w) e = e && a.z[w] == b.z[w] && a.x[w] == b.x[w];
    return e;
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

template <int W>
__device__ __forceinline__ int key_bit(const Key<W>& k, int b) {
    if (b >= 64 * W) {
        const int bb = b - 64 * W;
        return (int)((k.z[bb >> 6] >> (bb & 63)) & 1ull);
    }
    return (int)((k.x[b >> 6] >> (b & 63)) & 1ull);
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
template <int W>
__device__ __forceinline__ bool should_drop(const FilterDev* __restrict__ f, const Key<W>& k, double absc) {
    if (absc <= f->thresh) return true;
    const int w = popc_or<W>(k);
    if ((int64_t)w > f->wmax) return true;
    const int xw = popc_x<W>(k);
    if ((int64_t)xw > f->xwmax) return true;
    if (f->mwmax != INT64_MAX) {
        if ((int64_t)majorana_weight<W>(k) > f->mwmax) return true;
    }
    if (f->dthresh_w >= 0.0) {
        if (absc * f->exptab_w[w] <= f->dthresh_w) return true;
    }
    if (f->dthresh_xw >= 0.0) {
        if (absc * f->exptab_xw[xw] <= f->dthresh_xw) return true;
    }
    return false;
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
    __device__ __forceinline__ double child(const Key<W>& k, double c, Key<W>& out) const {
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
        return ((double)(kk - 2) * sinv) * c;
    }
};

// ---------------------------------------------------------------------------
// k_flags: one bit per live term: anticommutes with G and its child survives
// the local filter.  Writes the mask words, the in-tile exclusive prefix of
// each word and the per-tile count.
template <int W>
__global__ void __launch_bounds__(kBlock) k_flags(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                  const double* __restrict__ Lc, const int64_t* __restrict__ d_n,
                                                  const GenDev* __restrict__ g, const FilterDev* __restrict__ floc,
                                                  int local_active, uint64_t* __restrict__ flags,
                                                  uint16_t* __restrict__ wpre, int32_t* __restrict__ tilecnt) {
    __shared__ int s_wcnt[kTileWords];
    const int64_t n = *d_n;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    GenReg<W> G;
    G.load(g);
    bool need_z = false, need_x = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        need_z |= G.gx[w] != 0;
        need_x |= G.gz[w] != 0;
    }
    if (local_active) need_z = need_x = true;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
#pragma unroll
        for (int r = 0; r < kTileWords / 4; ++r) {
            const int wi = wave * (kTileWords / 4) + r;
            const int64_t i = t * kTile + wi * 64 + lane;
            bool flag = false;
            if (i < n) {
                Key<W> k;
                if constexpr (W == 1) {
                    k.z[0] = need_z ? Lz[i] : 0;
                    k.x[0] = need_x ? Lx[i] : 0;
                } else {
                    if (need_z) {
                        const ulonglong2 a = reinterpret_cast<const ulonglong2*>(Lz)[i];
                        k.z[0] = a.x; k.z[1] = a.y;
                    } else {
                        k.z[0] = k.z[1] = 0;
                    }
                    if (need_x) {
                        const ulonglong2 b = reinterpret_cast<const ulonglong2*>(Lx)[i];
                        k.x[0] = b.x; k.x[1] = b.y;
                    } else {
                        k.x[0] = k.x[1] = 0;
                    }
                }
                flag = G.anti(k);
                if (flag && local_active) {
                    Key<W> ch;
                    const double cn = G.child(k, Lc[i], ch);
                    flag = !should_drop<W>(floc, ch, fabs(cn));
                }
            }
            const uint64_t mask = __ballot(flag);
            if (lane == 0) {
                if (t * kTile + wi * 64 < n) flags[t * kTileWords + wi] = mask;
                s_wcnt[wi] = popc64(mask);
            }
        }
        __syncthreads();
        if (wave == 0) {
            const int v = lane < kTileWords ? s_wcnt[lane] : 0;
            int incl = v;
#pragma unroll
            for (int d = 1; d < kTileWords; d <<= 1) {
                const int u = __shfl_up(incl, d, kWave);
                if (lane >= d) incl += u;
            }
            if (lane < kTileWords && t * kTile + lane * 64 < n) wpre[t * kTileWords + lane] = (uint16_t)(incl - v);
            if (lane == kTileWords - 1) tilecnt[t] = incl;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_scan: exclusive scan of tile counts (one workgroup of 1024).  tileoff has
// ntiles + 1 entries; the last is m, also stored to *d_m.  Sets *d_ovf when m
// exceeds the child buffer.
__global__ void __launch_bounds__(1024) k_scan(const int32_t* __restrict__ tilecnt, const int64_t* __restrict__ d_n,
                                               int64_t* __restrict__ tileoff, int64_t* __restrict__ d_m,
                                               int64_t capC, int32_t* __restrict__ d_ovf) {
    __shared__ int64_t s_w[16];
    __shared__ int64_t s_carry;
    const int64_t n = *d_n;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < ntiles; base += 1024) {
        const int64_t t = base + tid;
        const int64_t v = t < ntiles ? tilecnt[t] : 0;
        int64_t incl = v;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const int64_t u = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += u;
        }
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        int64_t wpref = 0;
        for (int w = 0; w < wave; ++w) wpref += s_w[w];
        const int64_t carry = s_carry;
        if (t < ntiles) tileoff[t] = carry + wpref + incl - v;
        __syncthreads();
        if (tid == 1023) s_carry = carry + wpref + incl;
        __syncthreads();
    }
    if (tid == 0) {
        tileoff[ntiles] = s_carry;
        *d_m = s_carry;
        if (s_carry > capC) *d_ovf = 1;
    }
}

// Global exclusive A-prefix at live position p (0 <= p <= n).
__device__ __forceinline__ int64_t a_prefix(int64_t p, int64_t n, int64_t ntiles, const uint64_t* __restrict__ flags,
                                            const uint16_t* __restrict__ wpre, const int64_t* __restrict__ tileoff) {
    if (p >= n) return tileoff[ntiles];
    const int64_t w = p >> 6;
    const uint64_t lowmask = (p & 63) ? ((~0ull) >> (64 - (p & 63))) : 0ull;
    return tileoff[p / kTile] + (int64_t)wpre[w] + popc64(flags[w] & lowmask);
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

// ---------------------------------------------------------------------------
// k_bounds: for every (tile, level, side) the A-counts of trie segments that
// cross the tile edge.  Layout bnd[(t * nlev + li) * 4 + q]:
//   q=0 pre_same : A in [segstart_b(ts-1), ts)
//   q=1 pre_sib  : A in the segment just before segstart_b(ts-1) when it is
//                  that segment's bit-0 sibling and K_{ts-1} has bit b = 1
//   q=2 post_same: A in [te, segend_b(te))
//   q=3 post_sib : A in the segment just after segend_b(te) when it is the
//                  bit-1 sibling and K_te has bit b = 0
template <int W>
__global__ void __launch_bounds__(kBlock) k_bounds(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                   const int64_t* __restrict__ d_n, const GenDev* __restrict__ g,
                                                   const uint64_t* __restrict__ flags, const uint16_t* __restrict__ wpre,
                                                   const int64_t* __restrict__ tileoff, int64_t* __restrict__ bnd) {
    const int64_t n = *d_n;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const int nlev = g->nlev;
    const int64_t total = ntiles * nlev * 2;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
        const int side = (int)(idx & 1);
        const int li = (int)((idx >> 1) % nlev);
        const int64_t t = (idx >> 1) / nlev;
        const int b = g->lev[li];
        const int64_t ts = t * kTile;
        const int64_t te = ts + kTile < n ? ts + kTile : n;
        int64_t same = 0, sib = 0;
        if (side == 0) {
            if (ts > 0) {
                const Key<W> kp = load_key<W>(Lz, Lx, ts - 1);
                const int h = hdb<W>(kp, load_key<W>(Lz, Lx, ts));
                if (h <= b) {
                    const int64_t q = seg_start<W>(Lz, Lx, ts - 1, b);
                    same = a_prefix(ts, n, ntiles, flags, wpre, tileoff) - a_prefix(q, n, ntiles, flags, wpre, tileoff);
                    if (h < b && key_bit<W>(kp, b) == 1 && q > 0 &&
                        hdb<W>(load_key<W>(Lz, Lx, q - 1), load_key<W>(Lz, Lx, q)) == b) {
                        const int64_t q2 = seg_start<W>(Lz, Lx, q - 1, b);
                        sib = a_prefix(q, n, ntiles, flags, wpre, tileoff) - a_prefix(q2, n, ntiles, flags, wpre, tileoff);
                    }
                }
            }
            bnd[(t * nlev + li) * 4 + 0] = same;
            bnd[(t * nlev + li) * 4 + 1] = sib;
        } else {
            if (te < n) {
                const Key<W> kn = load_key<W>(Lz, Lx, te);
                const int h = hdb<W>(load_key<W>(Lz, Lx, te - 1), kn);
                if (h <= b) {
                    const int64_t r = seg_end<W>(Lz, Lx, te, n, b);
                    same = a_prefix(r, n, ntiles, flags, wpre, tileoff) - a_prefix(te, n, ntiles, flags, wpre, tileoff);
                    if (h < b && key_bit<W>(kn, b) == 0 && r < n &&
                        hdb<W>(load_key<W>(Lz, Lx, r - 1), load_key<W>(Lz, Lx, r)) == b) {
                        const int64_t r2 = seg_end<W>(Lz, Lx, r, n, b);
                        sib = a_prefix(r2, n, ntiles, flags, wpre, tileoff) - a_prefix(r, n, ntiles, flags, wpre, tileoff);
                    }
                }
            }
            bnd[(t * nlev + li) * 4 + 2] = same;
            bnd[(t * nlev + li) * 4 + 3] = sib;
        }
    }
}

// ---------------------------------------------------------------------------
// k_emit: children written at their rank in key order.
// Layout inside a tile: wave w owns elements [512 w, 512 w + 512), round r
// covers 64 consecutive elements (coalesced loads), lane = element % 64.
template <int W>
__global__ void __launch_bounds__(kBlock) k_emit(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                 const double* __restrict__ Lc, const int64_t* __restrict__ d_n,
                                                 const GenDev* __restrict__ g, const uint64_t* __restrict__ flags,
                                                 const uint16_t* __restrict__ wpre, const int64_t* __restrict__ tileoff,
                                                 const int64_t* __restrict__ bnd, uint64_t* __restrict__ Cz,
                                                 uint64_t* __restrict__ Cx, double* __restrict__ Cc,
                                                 const int32_t* __restrict__ d_ovf, int32_t* __restrict__ d_err) {
    constexpr int R = kTile / kBlock;  // 8 rounds
    __shared__ int16_t s_hdb[kTile + 1];
    __shared__ int16_t s_aloc[kTile + 1];
    __shared__ int16_t s_st[kTile];
    __shared__ int16_t s_e[kTile];
    __shared__ int s_wred[4];
    if (*d_ovf) return;
    const int64_t n = *d_n;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    GenReg<W> G;
    G.load(g);
    const int nlev = g->nlev;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t ts = t * kTile;
        const int Teff = (int)((n - ts) < kTile ? (n - ts) : kTile);
        Key<W> K[R];
        double C[R];
        int64_t rank[R];
        bool fl[R];
        const int64_t toff = tileoff[t];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = wave * (kTile / 4) + r * kWave + lane;
            const int64_t i = ts + j;
            fl[r] = false;
            rank[r] = 0;
            C[r] = 0.0;
            if (j < Teff) {
                K[r] = load_key<W>(Lz, Lx, i);
                const uint64_t word = flags[i >> 6];
                fl[r] = (word >> lane) & 1ull;
                const uint64_t lowmask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
                const int aloc = (int)wpre[i >> 6] + popc64(word & lowmask);
                s_aloc[j] = (int16_t)aloc;
                s_hdb[j] = (int16_t)(i == 0 ? kHdbNone : hdb<W>(load_key<W>(Lz, Lx, i - 1), K[r]));
                rank[r] = toff + aloc;
                if (fl[r]) C[r] = Lc[i];
            } else {
#pragma unroll
                for (int w = 0; w < W; ++w) K[r].z[w] = K[r].x[w] = 0;
            }
        }
        if (threadIdx.x == 0) {
            s_aloc[Teff] = (int16_t)(tileoff[t + 1] - toff);
            s_hdb[Teff] = (int16_t)(ts + Teff < n ? hdb<W>(load_key<W>(Lz, Lx, ts + Teff - 1), load_key<W>(Lz, Lx, ts + Teff))
                                                  : kHdbNone);
        }
        __syncthreads();
        for (int li = 0; li < nlev; ++li) {
            const int b = g->lev[li];
            const int64_t* bq = bnd + (t * nlev + li) * 4;
            // forward: st_j = last head <= j (j = 0 is always a head)
            int st[R];
            int carry = -1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                int v = (j < Teff && (j == 0 || s_hdb[j] >= b)) ? j : -1;
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const int u = __shfl_up(v, d, kWave);
                    if (lane >= d) v = v > u ? v : u;
                }
                v = v > carry ? v : carry;
                st[r] = v;
                carry = __shfl(v, kWave - 1, kWave);
            }
            if (lane == 0) s_wred[wave] = carry;
            __syncthreads();
            int pre = -1;
            for (int w = 0; w < wave; ++w) pre = pre > s_wred[w] ? pre : s_wred[w];
            __syncthreads();
            // backward: e_j = first head > j, Teff if none
            int e[R];
            int carry2 = Teff;
#pragma unroll
            for (int r = R - 1; r >= 0; --r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                int v = (j < Teff && (j == 0 || s_hdb[j] >= b)) ? j : Teff;
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const int u = __shfl_down(v, d, kWave);
                    if (lane + d < kWave) v = v < u ? v : u;
                }
                v = v < carry2 ? v : carry2;  // inclusive suffix min at j
                int nxt = __shfl_down(v, 1, kWave);
                if (lane == kWave - 1) nxt = carry2;
                e[r] = nxt;
                carry2 = __shfl(v, 0, kWave);
            }
            if (lane == 0) s_wred[wave] = carry2;
            __syncthreads();
            int post = Teff;
            for (int w = wave + 1; w < 4; ++w) post = post < s_wred[w] ? post : s_wred[w];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                st[r] = st[r] > pre ? st[r] : pre;
                e[r] = e[r] < post ? e[r] : post;
                if (j < Teff) {
                    s_st[j] = (int16_t)st[r];
                    s_e[j] = (int16_t)e[r];
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!fl[r]) continue;
                int64_t cnt = 0;
                if (key_bit<W>(K[r], b) == 0) {
                    const int ee = e[r];
                    if (ee < Teff) {
                        if (s_hdb[ee] == b) {
                            const int f = s_e[ee];
                            cnt = (int64_t)s_aloc[f] - s_aloc[ee];
                            if (f == Teff && s_hdb[Teff] < b) cnt += bq[2];
                        }
                    } else {
                        const int h = s_hdb[Teff];
                        cnt = h == b ? bq[2] : (h < b ? bq[3] : 0);
                    }
                    rank[r] += cnt;
                } else {
                    const int s = st[r];
                    if (s > 0) {
                        if (s_hdb[s] == b) {
                            const int s2 = s_st[s - 1];
                            cnt = (int64_t)s_aloc[s] - s_aloc[s2];
                            if (s2 == 0 && s_hdb[0] < b) cnt += bq[0];
                        }
                    } else {
                        const int h = s_hdb[0];
                        cnt = h == b ? bq[0] : (h < b ? bq[1] : 0);
                    }
                    rank[r] -= cnt;
                }
            }
            __syncthreads();
        }
        const int64_t mtot = tileoff[ntiles];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!fl[r]) continue;
            if (rank[r] < 0 || rank[r] >= mtot) {  // never expected: report, do not write
                atomicOr(d_err, 1);
                continue;
            }
            Key<W> ch;
            const double cn = G.child(K[r], C[r], ch);
            store_key<W>(Cz, Cx, rank[r], ch);
            Cc[rank[r]] = cn;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_partition: merge-path split of diagonal d = t * TM between live L (n) and
// children C (m), live first on equal keys.  Also clears the look-back
// status words and the tile ticket for k_merge, and writes n_out = 0 when
// there is nothing to merge.
template <int W>
__global__ void __launch_bounds__(kBlock) k_partition(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                      const int64_t* __restrict__ d_n, const uint64_t* __restrict__ Cz,
                                                      const uint64_t* __restrict__ Cx, const int64_t* __restrict__ d_m,
                                                      int64_t* __restrict__ split, uint64_t* __restrict__ status,
                                                      uint32_t* __restrict__ ticket, int64_t* __restrict__ d_nout,
                                                      int64_t max_threads, const int32_t* __restrict__ d_ovf) {
    constexpr int TM = MergeTile<W>::kSize;
    // The child buffer overflowed (m > its capacity): C holds fewer than m
    // valid terms, so no split may be searched; k_merge skips as well and the
    // host re-runs the rotation with a larger buffer.
    if (*d_ovf) return;
    const int64_t n = *d_n, m = *d_m;
    const int64_t tot = n + m;
    const int64_t ntm = (tot + TM - 1) / TM;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *ticket = 0u;
        if (ntm == 0) *d_nout = 0;
    }
    if (t >= max_threads) return;
    if (t < ntm) status[t] = 0ull;
    if (t <= ntm) {
        const int64_t d = t * TM < tot ? t * TM : tot;
        int64_t lo = d - m > 0 ? d - m : 0;
        int64_t hi = d < n ? d : n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (key_le<W>(load_key<W>(Lz, Lx, mid), load_key<W>(Cz, Cx, d - 1 - mid))) lo = mid + 1;
            else hi = mid;
        }
        split[t] = lo;
    }
}

// ---------------------------------------------------------------------------
// k_merge: live + children -> out (see the file header).  g == nullptr means
// a pure compaction (no rotation).  chan != nullptr applies a popcount
// damping to every surviving output after the filter (the channel that the
// reference runs after the last rotation of a layer, channels.jl:203-265).
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

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

template <int W>
__global__ void __launch_bounds__(kBlock) k_merge(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                  const double* __restrict__ Lc, const int64_t* __restrict__ d_n,
                                                  const uint64_t* __restrict__ Cz, const uint64_t* __restrict__ Cx,
                                                  const double* __restrict__ Cc, const int64_t* __restrict__ d_m,
                                                  const int64_t* __restrict__ split, const GenDev* __restrict__ g,
                                                  const FilterDev* __restrict__ f, ChanDev chan, int use_chan,
                                                  uint64_t* __restrict__ Oz, uint64_t* __restrict__ Ox,
                                                  double* __restrict__ Oc, int64_t capOut,
                                                  uint64_t* __restrict__ status, uint32_t* __restrict__ ticket,
                                                  int64_t* __restrict__ d_nout, int32_t* __restrict__ d_ovf,
                                                  int32_t* __restrict__ d_err) {
    constexpr int TM = MergeTile<W>::kSize;
    constexpr int E = TM / kBlock;
    __shared__ Key<W> s_key[TM];
    __shared__ double s_c[TM];
    __shared__ int64_t s_tile;
    __shared__ int64_t s_base;
    __shared__ int s_wsum[4];
    if (*d_ovf) return;
    const int64_t n = *d_n, m = *d_m;
    const int64_t tot = n + m;
    const int64_t ntm = (tot + TM - 1) / TM;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    GenReg<W> G;
    const bool rot = g != nullptr;
    if (rot) G.load(g);
    while (true) {
        if (tid == 0) s_tile = (int64_t)atomicAdd(ticket, 1u);
        __syncthreads();
        const int64_t t = s_tile;
        if (t >= ntm) break;
        const int64_t d0 = t * TM;
        const int64_t d1 = d0 + TM < tot ? d0 + TM : tot;
        int64_t i0 = split[t], i1 = split[t + 1];
        int64_t j0 = d0 - i0, j1 = d1 - i1;
        if (i0 < 0 || i1 < i0 || i1 > n || j0 < 0 || j1 < j0 || j1 > m) {  // never expected
            if (tid == 0) atomicOr(d_err, 2);
            i1 = i0 = 0;
            j1 = j0 = 0;
        }
        const int nL = (int)(i1 - i0), nC = (int)(j1 - j0);
        for (int k = tid; k < nL + nC; k += kBlock) {
            if (k < nL) {
                s_key[k] = load_key<W>(Lz, Lx, i0 + k);
                s_c[k] = Lc[i0 + k];
            } else {
                s_key[k] = load_key<W>(Cz, Cx, j0 + (k - nL));
                s_c[k] = Cc[j0 + (k - nL)];
            }
        }
        __syncthreads();
        const Key<W>* sL = s_key;
        const Key<W>* sC = s_key + nL;
        const double* sLc = s_c;
        const double* sCc = s_c + nL;
        // this thread's diagonal inside the tile
        const int dk = tid * E < nL + nC ? tid * E : nL + nC;
        int lo = dk - nC > 0 ? dk - nC : 0;
        int hi = dk < nL ? dk : nL;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (key_le<W>(sL[mid], sC[dk - 1 - mid])) lo = mid + 1;
            else hi = mid;
        }
        int ia = lo, jb = dk - lo;
        Key<W> outK[E];
        double outC[E];
        int cnt = 0;
        const int stop = dk + E < nL + nC ? dk + E : nL + nC;
        for (int pos = dk; pos < stop; ++pos) {
            const bool takeL = ia < nL && (jb >= nC || key_le<W>(sL[ia], sC[jb]));
            Key<W> K;
            double acc;
            bool emit = true;
            if (takeL) {
                K = sL[ia];
                acc = sLc[ia];
                if (rot && G.anti(K)) acc = acc * G.cosv;
                ++ia;
                // the equal-key child, if any, is the next C element
                if (jb < nC) {
                    if (key_eq<W>(sC[jb], K)) acc = acc + sCc[jb];
                } else if (j1 < m) {
                    if (key_eq<W>(load_key<W>(Cz, Cx, j1), K)) acc = acc + Cc[j1];
                }
            } else {
                K = sC[jb];
                acc = sCc[jb];
                ++jb;
                // absorbed by an equal live key, which precedes it directly
                if (ia > 0) {
                    if (key_eq<W>(sL[ia - 1], K)) emit = false;
                } else if (i0 > 0) {
                    if (key_eq<W>(load_key<W>(Lz, Lx, i0 - 1), K)) emit = false;
                }
            }
            if (emit && !should_drop<W>(f, K, fabs(acc))) {
                if (use_chan) acc = acc * chan_factor<W>(chan, K);
                outK[cnt] = K;
                outC[cnt] = acc;
                ++cnt;
            }
        }
        // block exclusive scan of survivor counts
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const int u = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += u;
        }
        if (lane == kWave - 1) s_wsum[wave] = incl;
        __syncthreads();
        int wpref = 0, total = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) wpref += s_wsum[w];
            total += s_wsum[w];
        }
        const int myoff = wpref + incl - cnt;
        // Decoupled look-back for the tile's global base, one wave scanning 64
        // predecessors per step.  Each status word carries its own payload
        // (flag | count), so relaxed atomics suffice: no other memory is
        // published through it, and no cache write-back/invalidate is needed.
        if (wave == 0) {
            if (t == 0) {
                if (lane == 0) {
                    __hip_atomic_store(&status[0], kFlagPre | (uint64_t)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_base = 0;
                }
            } else {
                if (lane == 0)
                    __hip_atomic_store(&status[t], kFlagAgg | (uint64_t)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int64_t excl = 0;
                int64_t top = t - 1;
                int64_t spins = 0;
                while (true) {
                    const int64_t q = top - lane;
                    const uint64_t s = q >= 0 ? __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                              : kFlagPre;  // before tile 0: prefix 0
                    const uint64_t flag = s & ~kValMask;
                    const uint64_t pre = __ballot(flag == kFlagPre);
                    const uint64_t notready = __ballot(flag == 0);
                    const int first = pre ? __ffsll((unsigned long long)pre) - 1 : kWave;
                    const uint64_t need = first >= kWave - 1 ? ~0ull : ((1ull << (first + 1)) - 1);
                    if (notready & need) {
                        if (++spins > (1ll << 22)) {  // bounded wait: report instead of hanging
                            if (lane == 0) atomicOr(d_err, 4);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    int64_t v = lane <= first ? (int64_t)(s & kValMask) : 0;
#pragma unroll
                    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
                    excl += v;
                    if (first < kWave) break;
                    top -= kWave;
                }
                if (lane == 0) {
                    __hip_atomic_store(&status[t], kFlagPre | (uint64_t)(excl + total), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    s_base = excl;
                }
            }
            if (lane == 0) {
                const int64_t b = s_base;
                if (t == ntm - 1) *d_nout = b + total;
                if (b + total > capOut) *d_ovf = 1;
            }
        }
        __syncthreads();
        const int64_t base = s_base;
        // stage survivors in LDS, then write coalesced
        for (int k = 0; k < cnt; ++k) {
            s_key[myoff + k] = outK[k];
            s_c[myoff + k] = outC[k];
        }
        __syncthreads();
        if (base + total <= capOut) {
            for (int k = tid; k < total; k += kBlock) {
                store_key<W>(Oz, Ox, base + k, s_key[k]);
                Oc[base + k] = s_c[k];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Channels (channels.jl:203-265): c *= table[popcount] in place.
template <int W>
__global__ void __launch_bounds__(kBlock) k_channel(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                    double* __restrict__ Lc, const int64_t* __restrict__ d_n,
                                                    ChanDev chan) {
    const int64_t n = *d_n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const Key<W> k = load_key<W>(Lz, Lx, i);
        Lc[i] = Lc[i] * chan_factor<W>(chan, k);
    }
}

// ---------------------------------------------------------------------------
// <k|O|k> (spv_kernels.jl:441-452) with a double-double accumulator so the
// parallel sum is accurate to ~2^-106 of sum |c| before its one rounding.
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

template <int W>
__global__ void __launch_bounds__(kBlock) k_expect(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                   const double* __restrict__ Lc, const int64_t* __restrict__ d_n,
                                                   uint64_t k0, uint64_t k1, double2* __restrict__ partial) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    const int64_t n = *d_n;
    double h = 0.0, l = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const Key<W> k = load_key<W>(Lz, Lx, i);
        bool diag = true;
        int par = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            diag = diag && k.x[w] == 0;
            par += popc64(k.z[w] & (w == 0 ? k0 : k1));
        }
        if (diag) {
            const double v = (par & 1) ? -Lc[i] : Lc[i];
            dd_add(h, l, v, 0.0);
        }
    }
    s_h[threadIdx.x] = h;
    s_l[threadIdx.x] = l;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double a = s_h[threadIdx.x], b = s_l[threadIdx.x];
            dd_add(a, b, s_h[threadIdx.x + s], s_l[threadIdx.x + s]);
            s_h[threadIdx.x] = a;
            s_l[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s_h[0], s_l[0]);
}

__global__ void __launch_bounds__(kBlock) k_expect_final(const double2* __restrict__ partial, int np,
                                                         double* __restrict__ out) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    double h = 0.0, l = 0.0;
    for (int i = threadIdx.x; i < np; i += kBlock) dd_add(h, l, partial[i].x, partial[i].y);
    s_h[threadIdx.x] = h;
    s_l[threadIdx.x] = l;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double a = s_h[threadIdx.x], b = s_l[threadIdx.x];
            dd_add(a, b, s_h[threadIdx.x + s], s_l[threadIdx.x + s]);
            s_h[threadIdx.x] = a;
            s_l[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s_h[0] + s_l[0];
}

// check_spv key order (type_SparsePauliVector.jl:273-276): *bad = 1 if any
// adjacent pair is not strictly increasing.
template <int W>
__global__ void __launch_bounds__(kBlock) k_check(const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                                                  const int64_t* __restrict__ d_n, int32_t* __restrict__ bad) {
    const int64_t n = *d_n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!key_lt<W>(load_key<W>(Lz, Lx, i - 1), load_key<W>(Lz, Lx, i))) *bad = 1;
    }
}

}  // namespace pp
