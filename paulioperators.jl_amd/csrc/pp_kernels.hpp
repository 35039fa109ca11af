// Device kernels of the SparsePauliVector engine for gfx950 (wave64).
//
// Storage is structure-of-arrays in HBM: z and x are W little-endian uint64
// words per term (W = 1 for N <= 64, 2 for N <= 128; a 16-byte vector per
// term when W = 2), coefficients are FP64.  The live set is kept sorted by
// the reference's (z, x) z-major key (spv_kernels.jl:14-16) and duplicate-free.
//
// Every kernel works on a batch of B independent vectors (SegDesc/SegState
// per vector; the single-vector API is B = 1).  Launch grids are sized from
// buffer capacities, never from live counts, so a whole sequence of
// rotations is enqueued without host round trips: each workgroup finds its
// vector through a prefix table over per-vector block counts, reads the live
// count from device memory and loops over that vector's tiles.
//
// One rotation at window = 1 (spv_evolve.jl:207-218 + 258-284):
//   kb_flags     anticommutation test per term -> bit mask + per-tile counts
//   kb_scan      exclusive scan of tile counts -> child offsets and m (only
//                for very long vectors or kb_bounds steps: otherwise kb_emit
//                sums the tile counts itself)
//   kb_bounds    per (tile, mask level) segment counts that cross tile edges
//   kb_emit      child (G xor P, ((k-2) sinθ) c) written straight to its rank
//                in key order (no sort pass, see below)
//   kb_merge     merge-path merge (each tile finds its own split with a
//                wave-parallel search), live term first on ties (spv_kernels.jl:338-350),
//                cos-scale of anticommuting live terms, sum of the (at most
//                one) child per key, strict filter on every output
//                (spv_kernels.jl:351), stream compaction with a decoupled
//                look-back scan.
// A vector whose children or merged output do not fit its buffers is frozen
// (SegState::stalled) with its input intact; the host regrows its buffers and
// replays it from that step.
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
#include "pp_device.hpp"

namespace pp {

// Largest b < B with prefix[b] <= g (prefix[0] = 0, g < prefix[B]).
__device__ __forceinline__ int find_seg(const int64_t* __restrict__ prefix, int B, int64_t g) {
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Grid mapping shared by every kernel.  blk holds, per launched block,
// b | lb << 20 | nb << 42: block lb of the nb blocks that serve vector b, or
// b = 2^20 - 1 for a block with no work.  A vector's blocks are laid out at
// indices of one residue modulo kXcds, so that (with workgroups dealt to the
// XCDs in turn) every kernel runs a vector on one XCD and its L2.  pre is the
// contiguous prefix of the per-vector block counts (for per-block partials).
struct BlockMap {
    const int64_t* pre;
    const uint64_t* blk;
    int B;
};

constexpr int kBlkVecBits = 20;   // vectors per batch < 2^20
constexpr int kBlkLocBits = 22;   // blocks per vector < 2^22
constexpr uint64_t kBlkNone = (1ull << kBlkVecBits) - 1;
constexpr int kXcds = 8;          // XCDs of an MI355X

// False for a block with no work (the caller returns).
__device__ __forceinline__ bool map_block(const BlockMap& bm, int& b, int64_t& lb, int64_t& nb) {
    const uint64_t e = bm.blk[blockIdx.x];
    b = (int)(e & ((1ull << kBlkVecBits) - 1));
    lb = (int64_t)((e >> kBlkVecBits) & ((1ull << kBlkLocBits) - 1));
    nb = (int64_t)(e >> (kBlkVecBits + kBlkLocBits));
    return b < bm.B;
}

// ---------------------------------------------------------------------------
// Segment summaries.  For a mask level b, a "head" is a live position j with
// hdb(K_{j-1}, K_j) >= b (position 0 always): the runs between heads are the
// trie segments of level b, and two adjacent runs are siblings when the head
// between them has hdb exactly b.  kb_flags stores, per sweep tile and level,
// the A-counts around the tile's first two and last two heads; kb_emit then
// finds the A-counts of the segments that cross its tile's edges by scanning
// the summaries of neighbouring tiles (kb_bounds does the same with key
// searches for masks of more than kAggLevels levels).
//   bits  0-11 a_first  A in [tile start, first head)   (whole tile if none)
//   bits 12-23 a_last   A in [last head, tile end)
//   bits 24-35 a_12     A in [first head, second head)
//   bits 36-47 a_l2     A in [second-last head, last head)
//   bits 48-49 heads    0, 1, or 2 for two or more
//   bit  50    eq_first hdb at the first head == b
//   bit  51    eq_last  hdb at the last head == b
constexpr int kAggLevels = 4;
constexpr int kEmitSelfTiles = 512;  // kb_emit sums the tile counts itself up to this many tiles
// Segment summaries serve calls whose vectors hold at most this many sweep
// tiles: an edge scan over the summaries grows with the segment length in
// tiles, kb_bounds' key searches only with its logarithm.
constexpr int64_t kSummMaxTiles = 256;

__device__ __forceinline__ int agg_heads(uint64_t a) { return (int)((a >> 48) & 3u); }
__device__ __forceinline__ int64_t agg_first(uint64_t a) { return (int64_t)(a & 4095u); }
__device__ __forceinline__ int64_t agg_last(uint64_t a) { return (int64_t)((a >> 12) & 4095u); }
__device__ __forceinline__ int64_t agg_12(uint64_t a) { return (int64_t)((a >> 24) & 4095u); }
__device__ __forceinline__ int64_t agg_l2(uint64_t a) { return (int64_t)((a >> 36) & 4095u); }
__device__ __forceinline__ bool agg_eq_first(uint64_t a) { return (a >> 50) & 1u; }
__device__ __forceinline__ bool agg_eq_last(uint64_t a) { return (a >> 51) & 1u; }

// One wave: the summary of one level of one tile from the per-word head,
// equal-level and A bit masks and the per-word A counts (lane k holds word k).
template <int kWords>
__device__ uint64_t tile_summary(const uint64_t* hm_w, const uint64_t* em_w, const uint64_t* am_w, const int* wcnt) {
    static_assert(kWords <= kWave, "one word per lane");
    const int lane = threadIdx.x % kWave;
    const bool in = lane < kWords;
    const uint64_t hm = in ? hm_w[lane] : 0ull;
    const uint64_t em = in ? em_w[lane] : 0ull;
    const uint64_t am = in ? am_w[lane] : 0ull;
    const int cnt = in ? wcnt[lane] : 0;
    int incl = cnt;
#pragma unroll
    for (int dd = 1; dd < kWords; dd <<= 1) {
        const int u = __shfl_up(incl, dd, kWave);
        if (lane >= dd) incl += u;
    }
    const int pre = incl - cnt;
    const int total = __shfl(incl, kWords - 1, kWave);
    // A-count of [tile start, p) for a wave-uniform p < kWords * 64
    auto apos = [&](int p) {
        const int w = p >> 6;
        const int pw = __shfl(pre, w, kWave);
        const uint64_t aw = __shfl(am, w, kWave);
        const uint64_t low = (p & 63) ? ((~0ull) >> (64 - (p & 63))) : 0ull;
        return pw + popc64(aw & low);
    };
    const uint64_t hw = __ballot(hm != 0ull);  // words holding a head
    int f1 = -1, f2 = -1, l1 = -1, l2 = -1;
    if (hw) {
        const int wf = __ffsll((unsigned long long)hw) - 1;
        const uint64_t m0 = __shfl(hm, wf, kWave);
        f1 = wf * 64 + __ffsll((unsigned long long)m0) - 1;
        const uint64_t m1 = m0 & (m0 - 1);
        if (m1) {
            f2 = wf * 64 + __ffsll((unsigned long long)m1) - 1;
        } else {
            const uint64_t hw2 = hw & (hw - 1);
            if (hw2) {
                const int w2 = __ffsll((unsigned long long)hw2) - 1;
                f2 = w2 * 64 + __ffsll((unsigned long long)__shfl(hm, w2, kWave)) - 1;
            }
        }
        const int wl = 63 - __clzll((long long)hw);
        const uint64_t n0 = __shfl(hm, wl, kWave);
        const int q = 63 - __clzll((long long)n0);
        l1 = wl * 64 + q;
        const uint64_t n1 = n0 & ~(1ull << q);
        if (n1) {
            l2 = wl * 64 + 63 - __clzll((long long)n1);
        } else {
            const uint64_t hw3 = hw & ~(1ull << wl);
            if (hw3) {
                const int w3 = 63 - __clzll((long long)hw3);
                l2 = w3 * 64 + 63 - __clzll((long long)__shfl(hm, w3, kWave));
            }
        }
    }
    const int af = f1 < 0 ? total : apos(f1);
    const int al = l1 < 0 ? 0 : total - apos(l1);
    const int a12 = f2 < 0 ? 0 : apos(f2) - af;
    const int al2 = l2 < 0 ? 0 : apos(l1) - apos(l2);
    const uint64_t heads = f1 < 0 ? 0u : (f2 < 0 ? 1u : 2u);
    const uint64_t eqf = f1 >= 0 ? ((__shfl(em, f1 >> 6, kWave) >> (f1 & 63)) & 1u) : 0u;
    const uint64_t eql = l1 >= 0 ? ((__shfl(em, l1 >> 6, kWave) >> (l1 & 63)) & 1u) : 0u;
    return (uint64_t)af | ((uint64_t)al << 12) | ((uint64_t)a12 << 24) | ((uint64_t)al2 << 36) | (heads << 48) |
           (eqf << 50) | (eql << 51);
}

struct AggHit {
    int64_t u;      // tile holding the head (ntiles when none)
    int64_t sum;    // A-counts of the tiles passed over
    uint64_t a;     // that tile's summary
};

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int dd = kWave / 2; dd > 0; dd >>= 1) v += __shfl_xor(v, dd, kWave);
    return v;
}

// Wave-cooperative: the largest tile u <= u0 with a head at level li, and the
// A-count of the tiles in (u, u0].  Tile 0 always has a head (position 0).
__device__ AggHit agg_back(const uint64_t* __restrict__ agg, const int32_t* __restrict__ tilecnt, int64_t u0, int li) {
    const int lane = threadIdx.x % kWave;
    int64_t base = u0, sum = 0;
    while (base >= 0) {
        const int64_t u = base - lane;
        uint64_t a = 0;
        int64_t c = 0;
        if (u >= 0) {
            a = ldg(agg, u * kAggLevels + li);
            c = ldg(tilecnt, u);
        }
        const uint64_t hm = __ballot(u >= 0 && agg_heads(a) != 0);
        if (hm) {
            const int f = __ffsll((unsigned long long)hm) - 1;
            sum += wave_sum64(lane < f ? c : 0);
            const uint64_t af = __shfl(a, f, kWave);
            return AggHit{base - f, sum, af};
        }
        sum += wave_sum64(c);
        base -= kWave;
    }
    return AggHit{-1, sum, 0};
}

// Wave-cooperative: the smallest tile v >= v0 with a head at level li
// (ntiles if none), and the A-count of the tiles in [v0, v).
__device__ AggHit agg_fwd(const uint64_t* __restrict__ agg, const int32_t* __restrict__ tilecnt, int64_t v0,
                          int64_t ntiles, int li) {
    const int lane = threadIdx.x % kWave;
    int64_t base = v0, sum = 0;
    while (base < ntiles) {
        const int64_t u = base + lane;
        uint64_t a = 0;
        int64_t c = 0;
        if (u < ntiles) {
            a = ldg(agg, u * kAggLevels + li);
            c = ldg(tilecnt, u);
        }
        const uint64_t hm = __ballot(u < ntiles && agg_heads(a) != 0);
        if (hm) {
            const int f = __ffsll((unsigned long long)hm) - 1;
            sum += wave_sum64(lane < f ? c : 0);
            const uint64_t af = __shfl(a, f, kWave);
            return AggHit{base + f, sum, af};
        }
        sum += wave_sum64(c);
        base += kWave;
    }
    return AggHit{ntiles, sum, 0};
}

// The summaries of 64 consecutive tiles at one level and their A-counts:
// lane l holds tile base + dir * l (when that tile exists).  An edge scan
// loads it together with the keys across the edge, so that its first step
// shares their memory round trip.
struct AggWin {
    int64_t base;
    int dir;
    uint64_t a;
    int64_t c;
    bool in;
};

__device__ __forceinline__ AggWin agg_window(const uint64_t* __restrict__ agg, const int32_t* __restrict__ tilecnt,
                                             int64_t base, int dir, int64_t ntiles, int li) {
    const int lane = threadIdx.x % kWave;
    AggWin w{base, dir, 0ull, 0, false};
    const int64_t u = base + (int64_t)dir * lane;
    if (u >= 0 && u < ntiles) {
        w.a = ldg(agg, u * kAggLevels + li);
        w.c = ldg(tilecnt, u);
        w.in = true;
    }
    return w;
}

// agg_back (dir -1) or agg_fwd (dir +1) from tile u0, through the window
// where it covers the start of the scan, then through memory.
__device__ AggHit agg_scan(const AggWin& w, const uint64_t* __restrict__ agg, const int32_t* __restrict__ tilecnt,
                           int64_t u0, int dir, int64_t ntiles, int li) {
    const int lane = threadIdx.x % kWave;
    const int64_t k0 = (u0 - w.base) * dir;  // window lane of tile u0
    if (w.dir != dir || k0 < 0 || k0 >= kWave)
        return dir < 0 ? agg_back(agg, tilecnt, u0, li) : agg_fwd(agg, tilecnt, u0, ntiles, li);
    const bool cand = lane >= k0 && w.in;
    const uint64_t hm = __ballot(cand && agg_heads(w.a) != 0);
    if (hm) {
        const int f = __ffsll((unsigned long long)hm) - 1;
        const int64_t sum = wave_sum64(cand && lane < f ? w.c : 0);
        return AggHit{w.base + (int64_t)dir * f, sum, __shfl(w.a, f, kWave)};
    }
    const int64_t sum = wave_sum64(cand ? w.c : 0);
    const int64_t next = w.base + (int64_t)dir * kWave;  // first tile past the window
    AggHit r = dir < 0 ? (next >= 0 ? agg_back(agg, tilecnt, next, li) : AggHit{-1, 0, 0})
                       : (next < ntiles ? agg_fwd(agg, tilecnt, next, ntiles, li) : AggHit{ntiles, 0, 0});
    r.sum += sum;
    return r;
}

// The four cross-edge counts kb_bounds computes (same layout: pre_same,
// pre_sib, post_same, post_sib), for one side of tile t at level lvl, from
// the summaries.  Wave-uniform; `out` receives two values.
// h_edge receives hdb(K[te - 1], K[te]) on side 1 (kHdbNone when te = n).
// WIN: load a window of summaries with the edge keys (costs registers; used
// for small tiles, whose kernels are latency-bound).
template <int W, bool WIN>
__device__ void edge_counts(int side, int li, int lvl, int64_t t, int64_t ntiles, int64_t n, int64_t ts, int64_t te,
                            const uint64_t* __restrict__ Lz, const uint64_t* __restrict__ Lx,
                            const uint64_t* __restrict__ agg, const int32_t* __restrict__ tilecnt, int64_t& same,
                            int64_t& sib, int& h_edge) {
    same = 0;
    sib = 0;
    h_edge = kHdbNone;
    if (side == 0) {
        if (ts == 0) return;
        const Key<W> ka = load_key<W>(Lz, Lx, ts - 1);
        const Key<W> kb = load_key<W>(Lz, Lx, ts);
        const AggWin win = WIN ? agg_window(agg, tilecnt, t - 1, -1, ntiles, li)  // tiles t-1, t-2, ...
                               : AggWin{0, 0, 0ull, 0, false};
        const int h = hdb<W>(ka, kb);
        if (h > lvl) return;
        const AggHit q = agg_scan(win, agg, tilecnt, t - 1, -1, ntiles, li);  // q: last head <= ts - 1
        same = agg_last(q.a) + q.sum;
        if (h < lvl && key_bit<W>(ka, lvl) == 1 && agg_eq_last(q.a)) {
            if (agg_heads(q.a) >= 2) {
                sib = agg_l2(q.a);
            } else {
                const AggHit q2 = agg_scan(win, agg, tilecnt, q.u - 1, -1, ntiles, li);
                sib = agg_first(q.a) + q2.sum + agg_last(q2.a);
            }
        }
    } else {
        if (te >= n) return;
        const Key<W> ka = load_key<W>(Lz, Lx, te);
        const Key<W> kp = load_key<W>(Lz, Lx, te - 1);
        const AggWin win = WIN ? agg_window(agg, tilecnt, t + 1, 1, ntiles, li)  // tiles t+1, t+2, ...
                               : AggWin{0, 0, 0ull, 0, false};
        const int h = hdb<W>(kp, ka);
        h_edge = h;
        if (h > lvl) return;
        if (h == lvl) {  // te is a head: the run [te, next head)
            const uint64_t a = WIN ? __shfl(win.a, 0, kWave) : ldg(agg, (t + 1) * kAggLevels + li);  // tile t + 1
            if (agg_heads(a) >= 2) {
                same = agg_12(a);
            } else {
                const AggHit r = agg_scan(win, agg, tilecnt, t + 2, 1, ntiles, li);
                same = (WIN ? __shfl(win.c, 0, kWave) : (int64_t)ldg(tilecnt, t + 1)) + r.sum +
                       (r.u < ntiles ? agg_first(r.a) : 0);
            }
            return;
        }
        const AggHit r = agg_scan(win, agg, tilecnt, t + 1, 1, ntiles, li);  // r: first head > te
        same = r.sum + (r.u < ntiles ? agg_first(r.a) : 0);
        if (r.u < ntiles && key_bit<W>(ka, lvl) == 0 && agg_eq_first(r.a)) {
            if (agg_heads(r.a) >= 2) {
                sib = agg_12(r.a);
            } else {
                const AggHit r2 = agg_scan(win, agg, tilecnt, r.u + 1, 1, ntiles, li);
                sib = agg_last(r.a) + r2.sum + (r2.u < ntiles ? agg_first(r2.a) : 0);
            }
        }
    }
}

// ---------------------------------------------------------------------------
template <int W, int T, typename CT>
__global__ void __launch_bounds__(kBlock) kb_flags(BlockMap bm, const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                   const StepDev* __restrict__ steps, int bnd_ok) {
    constexpr int kTile = T;
    constexpr int kTileWords = T / 64;
    __shared__ int s_wcnt[kTileWords];
    __shared__ uint64_t s_am[kTileWords];                   // A bits per word
    __shared__ uint64_t s_hm[kAggLevels][kTileWords];       // run heads per level
    __shared__ uint64_t s_em[kAggLevels][kTileWords];       // heads whose hdb equals the level
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    const StepDev* st = steps + b;
    SegState* s = S + b;
    const SegDesc& d = D[b];
    // every per-vector value is loaded before the first test, so the loads
    // share one memory round trip (the empty asm takes them all as operands,
    // which keeps the compiler from sinking a load past a branch)
    const int active = st->active, stalled = s->stalled, cur0 = s->cur, flip = s->flip, nlev = st->g.nlev;
    const int precheck = st->precheck;
    const int64_t n0 = s->n[0], n1 = s->n[1], appcap = s->app_cap_v;
    const uint64_t *Lz0 = d.Lz[0], *Lz1 = d.Lz[1], *Lx0 = d.Lx[0], *Lx1 = d.Lx[1];
    const CT *Lc0 = reinterpret_cast<const CT*>(d.Lc[0]), *Lc1 = reinterpret_cast<const CT*>(d.Lc[1]);
    uint64_t* __restrict__ flags = d.flags;
    uint16_t* __restrict__ wpre = d.wpre;
    int32_t* __restrict__ tilecnt = d.tilecnt;
    uint64_t* __restrict__ agg = d.agg;
    GenReg<W> G;
    G.load(&st->g);
    const int local_active = d.local_active;
    const uint32_t lev4 = *reinterpret_cast<const uint32_t*>(st->g.lev);  // the first four levels
    asm volatile("" ::"s"(active), "s"(stalled), "s"(cur0), "s"(flip), "s"(nlev), "s"(n0), "s"(n1), "s"(Lz0),
                 "s"(Lz1), "s"(Lx0), "s"(Lx1), "s"(Lc0), "s"(Lc1), "s"(flags), "s"(wpre), "s"(tilecnt), "s"(agg),
                 "s"(local_active), "s"(lev4), "s"(G.gz[0]), "s"(G.gx[0]), "s"(G.gz[W - 1]), "s"(G.gx[W - 1]),
                 "s"(precheck), "s"(appcap));
    if (!active || stalled) return;
    const int cur = cur0 ^ flip;
    const int64_t n = cur ? n1 : n0;
    if (lb == 0 && threadIdx.x == 0) {
        s->nstep = n;
        // the capacity precheck of spv_evolve.jl:356-362 for rotations after
        // the first of a sequence: the early merge it triggers drops nothing
        // (every live term passed the strict filter at the last merge)
        if (precheck && n > appcap) {
            s->early += 1;
            s->merge_in += n;
            s->merge_out += n;
            s->app_cap_v = 2 * appcap > n ? 2 * appcap : n;
        }
    }
    const uint64_t* __restrict__ Lz = cur ? Lz1 : Lz0;
    const uint64_t* __restrict__ Lx = cur ? Lx1 : Lx0;
    const CT* __restrict__ Lc = cur ? Lc1 : Lc0;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    FilterReg floc;  // the local filter, when it can drop children
    if (local_active) floc.load(d.fl);
    // segment summaries for kb_emit when the mask has 1..kAggLevels levels
    const bool summ = nlev >= 1 && nlev <= kAggLevels && !bnd_ok;  // kb_emit's summary variant runs
    bool need_z = false, need_x = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        need_z |= G.gx[w] != 0;
        need_x |= G.gz[w] != 0;
    }
    if (local_active || summ) need_z = need_x = true;
    constexpr int RW = kTileWords / 4;  // rounds of 64 terms per wave
    for (int64_t t = lb; t < ntiles; t += nb) {
        // every load of the tile first (one memory round trip): the wave's
        // keys, their coefficients when the local filter runs, and the key
        // before the wave's first term (the later rounds take theirs from
        // lane 63 of the round before)
        Key<W> kk[RW];
        CT cc[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int64_t i = t * kTile + (wave * RW + r) * 64 + lane;
#pragma unroll
            for (int w = 0; w < W; ++w) kk[r].z[w] = kk[r].x[w] = 0;
            cc[r] = coef_zero<CT>();
            if (i < n) {
                if constexpr (W == 1) {
                    kk[r].z[0] = need_z ? ldg(Lz, i) : 0;
                    kk[r].x[0] = need_x ? ldg(Lx, i) : 0;
                } else {
                    if (need_z) {
                        const ulonglong2 a = ldg(reinterpret_cast<const ulonglong2*>(Lz), i);
                        kk[r].z[0] = a.x; kk[r].z[1] = a.y;
                    }
                    if (need_x) {
                        const ulonglong2 c2 = ldg(reinterpret_cast<const ulonglong2*>(Lx), i);
                        kk[r].x[0] = c2.x; kk[r].x[1] = c2.y;
                    }
                }
                if (local_active) cc[r] = ldg(Lc, i);
            }
        }
        Key<W> kp0;
#pragma unroll
        for (int w = 0; w < W; ++w) kp0.z[w] = kp0.x[w] = 0;
        {
            const int64_t i0 = t * kTile + wave * RW * 64;
            if (summ && lane == 0 && i0 > 0 && i0 < n) kp0 = load_key<W>(Lz, Lx, i0 - 1);
        }
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int wi = wave * RW + r;
            const int64_t i = t * kTile + wi * 64 + lane;
            const bool valid = i < n;
            bool flag = false;
            const Key<W> k = kk[r];
            if (valid) {
                flag = G.anti(k);
                if (flag && local_active) {
                    Key<W> ch;
                    const CT cn = G.template child<CT>(k, cc[r], ch);
                    flag = !should_drop<W>(floc, ch, coef_abs(cn));
                }
            }
            const uint64_t mask = __ballot(flag);
            if (lane == 0) {
                if (t * kTile + wi * 64 < n) stg(flags, t * kTileWords + wi, mask);
                s_wcnt[wi] = popc64(mask);
                s_am[wi] = mask;
            }
            if (summ) {
                // hdb against the previous key: from the lane below; lane 0
                // from lane 63 of the round before, or the key loaded above
                Key<W> kp;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    kp.z[w] = __shfl_up(k.z[w], 1, kWave);
                    kp.x[w] = __shfl_up(k.x[w], 1, kWave);
                }
                if (r == 0) {
                    if (lane == 0) kp = kp0;
                } else {
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        const uint64_t pz = __shfl(kk[r - 1].z[w], kWave - 1, kWave);
                        const uint64_t px = __shfl(kk[r - 1].x[w], kWave - 1, kWave);
                        if (lane == 0) {
                            kp.z[w] = pz;
                            kp.x[w] = px;
                        }
                    }
                }
                const int hd = i == 0 ? (int)kHdbNone : hdb<W>(kp, k);
#pragma unroll
                for (int li = 0; li < kAggLevels; ++li) {
                    if (li < nlev) {
                        const int lvl = (int)((lev4 >> (8 * li)) & 255u);
                        const uint64_t hm = __ballot(valid && hd >= lvl);
                        const uint64_t em = __ballot(valid && hd == lvl);
                        if (lane == 0) {
                            s_hm[li][wi] = hm;
                            s_em[li][wi] = em;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (wave == 0) {
            const int v = lane < kTileWords ? s_wcnt[lane] : 0;
            int incl = v;
#pragma unroll
            for (int dd = 1; dd < kTileWords; dd <<= 1) {
                const int u = __shfl_up(incl, dd, kWave);
                if (lane >= dd) incl += u;
            }
            if (lane < kTileWords && t * kTile + lane * 64 < n) stg(wpre, t * kTileWords + lane, (uint16_t)(incl - v));
            if (lane == kTileWords - 1) stg(tilecnt, t, (int32_t)incl);
        } else if (summ) {
            // waves 1-3: one level each (wave 1 also the fourth)
            for (int li = wave - 1; li < nlev; li += 3) {
                const uint64_t a = tile_summary<kTileWords>(s_hm[li], s_em[li], s_am, s_wcnt);
                if (lane == 0) stg(agg, t * kAggLevels + li, a);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// kb_scan: one workgroup per vector.  Scans the tile counts, checks the
// child buffer and resets the merge ticket.
template <int T>
__global__ void __launch_bounds__(1024) kb_scan(const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                const StepDev* __restrict__ steps, int64_t step_index) {
    constexpr int kTile = T;
    __shared__ int64_t s_w[16];
    __shared__ int64_t s_carry;
    const int b = blockIdx.x;
    SegState* s = S + b;
    const StepDev* st = steps + b;
    const SegDesc& d = D[b];
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    // per-vector values in one round trip (see kb_flags)
    const int active = st->active, stalled = s->stalled;
    const int64_t n = s->nstep;  // written by kb_flags
    const int32_t* __restrict__ tilecnt = d.tilecnt;
    int64_t* __restrict__ tileoff = d.tileoff;
    const int64_t capC = d.capC;
    asm volatile("" ::"s"(active), "s"(stalled), "s"(n), "s"(tilecnt), "s"(tileoff), "s"(capC));
    const bool go = active && !stalled;
    if (!go) return;
    if (tid == 0) {
        s_carry = 0;
        s->ticket = 0u;
    }
    __syncthreads();
    const int64_t ntiles = (n + kTile - 1) / kTile;
    for (int64_t base = 0; base < ntiles; base += 1024) {
        const int64_t t = base + tid;
        const int64_t v = t < ntiles ? ldg(tilecnt, t) : 0;
        int64_t incl = v;
#pragma unroll
        for (int dd = 1; dd < kWave; dd <<= 1) {
            const int64_t u = __shfl_up(incl, dd, kWave);
            if (lane >= dd) incl += u;
        }
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        int64_t wpref = 0;
        for (int w = 0; w < wave; ++w) wpref += s_w[w];
        const int64_t carry = s_carry;
        if (t < ntiles) stg(tileoff, t, carry + wpref + incl - v);
        __syncthreads();
        if (tid == 1023) s_carry = carry + wpref + incl;
        __syncthreads();
    }
    if (tid == 0) {
        const int64_t m = s_carry;
        stg(tileoff, ntiles, m);
        s->m = m;
        if (m > capC) {
            s->stalled = 1;
            s->stall_step = step_index;
            s->need = m;
        }
    }
}

// ---------------------------------------------------------------------------
// kb_bounds: for every (tile, level, side) the A-counts of trie segments that
// cross the tile edge.  Layout bnd[(t * maxlev + li) * 4 + q]:
//   q=0 pre_same : A in [segstart_b(ts-1), ts)
//   q=1 pre_sib  : A in the segment just before segstart_b(ts-1) when it is
//                  that segment's bit-0 sibling and K_{ts-1} has bit b = 1
//   q=2 post_same: A in [te, segend_b(te))
//   q=3 post_sib : A in the segment just after segend_b(te) when it is the
//                  bit-1 sibling and K_te has bit b = 0
template <int W, int T>
__global__ void __launch_bounds__(kBlock) kb_bounds(BlockMap bm, const SegDesc* __restrict__ D,
                                                    const SegState* __restrict__ S, const StepDev* __restrict__ steps) {
    constexpr int kTile = T;
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    const StepDev* st = steps + b;
    const SegState* s = S + b;
    const SegDesc& d = D[b];
    // per-vector values in one round trip (see kb_flags)
    const int active = st->active, stalled = s->stalled, cur = s->cur ^ s->flip, nlev = st->g.nlev,
              maxlev = d.maxlev;
    const int64_t n0 = s->n[0], n1 = s->n[1];
    const uint64_t *Lz0 = d.Lz[0], *Lz1 = d.Lz[1], *Lx0 = d.Lx[0], *Lx1 = d.Lx[1];
    const uint64_t* __restrict__ flags = d.flags;
    const uint16_t* __restrict__ wpre = d.wpre;
    const int64_t* __restrict__ tileoff = d.tileoff;
    int64_t* __restrict__ bnd = d.bnd;
    const uint32_t lev4 = *reinterpret_cast<const uint32_t*>(st->g.lev);  // the first four levels
    asm volatile("" ::"s"(active), "s"(stalled), "s"(cur), "s"(nlev), "s"(maxlev), "s"(n0), "s"(n1), "s"(Lz0),
                 "s"(Lz1), "s"(Lx0), "s"(Lx1), "s"(flags), "s"(wpre), "s"(tileoff), "s"(bnd), "s"(lev4));
    if (!active || stalled) return;
    const int64_t n = cur ? n1 : n0;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const uint64_t* __restrict__ Lz = cur ? Lz1 : Lz0;
    const uint64_t* __restrict__ Lx = cur ? Lx1 : Lx0;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t total = ntiles * 2;
    // one wave per (tile, side); the searches are wave-cooperative
    for (int64_t idx = lb * (kBlock / kWave) + wave; idx < total; idx += nb * (kBlock / kWave)) {
        const int side = (int)(idx & 1);
        const int64_t t = idx >> 1;
        const int64_t ts = t * kTile;
        const int64_t te = ts + kTile < n ? ts + kTile : n;
        Key<W> ka, kb2;
        int h = kHdbNone;
        if (side == 0 && ts > 0) {
            ka = load_key<W>(Lz, Lx, ts - 1);
            h = hdb<W>(ka, load_key<W>(Lz, Lx, ts));
        } else if (side == 1 && te < n) {
            ka = load_key<W>(Lz, Lx, te);
            h = hdb<W>(load_key<W>(Lz, Lx, te - 1), ka);
        }
        for (int li = 0; li < nlev; ++li) {
            const int lvl = li < 4 ? (int)((lev4 >> (8 * li)) & 255u) : (int)st->g.lev[li];
            int64_t same = 0, sib = 0;
            if (h <= lvl) {
                if (side == 0) {
                    const int64_t q = seg_start_wave<W>(Lz, Lx, ts - 1, lvl, ka);
                    same = a_prefix<T>(ts, n, ntiles, flags, wpre, tileoff) -
                           a_prefix<T>(q, n, ntiles, flags, wpre, tileoff);
                    if (h < lvl && key_bit<W>(ka, lvl) == 1 && q > 0) {
                        kb2 = load_key<W>(Lz, Lx, q - 1);
                        if (hdb<W>(kb2, load_key<W>(Lz, Lx, q)) == lvl) {
                            const int64_t q2 = seg_start_wave<W>(Lz, Lx, q - 1, lvl, kb2);
                            sib = a_prefix<T>(q, n, ntiles, flags, wpre, tileoff) -
                                  a_prefix<T>(q2, n, ntiles, flags, wpre, tileoff);
                        }
                    }
                } else {
                    const int64_t r = seg_end_wave<W>(Lz, Lx, te, n, lvl, ka);
                    same = a_prefix<T>(r, n, ntiles, flags, wpre, tileoff) -
                           a_prefix<T>(te, n, ntiles, flags, wpre, tileoff);
                    if (h < lvl && key_bit<W>(ka, lvl) == 0 && r < n) {
                        kb2 = load_key<W>(Lz, Lx, r);
                        if (hdb<W>(load_key<W>(Lz, Lx, r - 1), kb2) == lvl) {
                            const int64_t r2 = seg_end_wave<W>(Lz, Lx, r, n, lvl, kb2);
                            sib = a_prefix<T>(r2, n, ntiles, flags, wpre, tileoff) -
                                  a_prefix<T>(r, n, ntiles, flags, wpre, tileoff);
                        }
                    }
                }
            }
            if (lane == 0) {
                stg(bnd, (t * maxlev + li) * 4 + 2 * side + 0, same);
                stg(bnd, (t * maxlev + li) * 4 + 2 * side + 1, sib);
            }
        }
    }
}

// Run heads of one level in a tile as bit masks (word c: positions
// 64c..64c+63; position 0 is always a head) plus a mask of the words that
// hold a head: the first head after p (Teff if none) and the last head at or
// before p, each in O(1).
__device__ __forceinline__ int next_head(const uint64_t* hm, uint64_t words, int p, int Teff) {
    const int c = p >> 6;
    const uint64_t m = (p & 63) == 63 ? 0ull : hm[c] & (~0ull << ((p & 63) + 1));
    if (m) return c * 64 + __ffsll((unsigned long long)m) - 1;
    const uint64_t wm = c == 63 ? 0ull : words & (~0ull << (c + 1));
    if (!wm) return Teff;
    const int c2 = __ffsll((unsigned long long)wm) - 1;
    return c2 * 64 + __ffsll((unsigned long long)hm[c2]) - 1;
}
__device__ __forceinline__ int last_head(const uint64_t* hm, uint64_t words, int p) {
    const int c = p >> 6;
    const uint64_t m = hm[c] & ((p & 63) == 63 ? ~0ull : ((2ull << (p & 63)) - 1));
    if (m) return c * 64 + 63 - __clzll((long long)m);
    const uint64_t wm = words & ((1ull << c) - 1);  // c >= 1 here: word 0 holds position 0
    const int c2 = 63 - __clzll((long long)wm);
    return c2 * 64 + 63 - __clzll((long long)hm[c2]);
}

// ---------------------------------------------------------------------------
// kb_emit: children written at their rank in key order.
// Layout inside a tile: wave w owns elements [w T/4, (w+1) T/4), round r
// covers 64 consecutive elements (coalesced loads), lane = element % 64.
template <int W, int T, bool TABLE, typename CT>
__global__ void __launch_bounds__(kBlock) kb_emit(BlockMap bm, const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                  const StepDev* __restrict__ steps, int self_off,
                                                  int64_t step_index) {
    constexpr int kTile = T;
    constexpr int R = kTile / kBlock;  // rounds of 64 consecutive terms per wave
    __shared__ int16_t s_hdb[kTile + 1];
    __shared__ int16_t s_aloc[kTile + 1];
    __shared__ int16_t s_st[kTile];
    __shared__ int16_t s_e[kTile];
    __shared__ int s_wred[4];
    __shared__ int64_t s_bq[kAggLevels][4];
    __shared__ uint64_t s_lhm[kAggLevels][kTile / 64];  // run heads per level (summary path)
    __shared__ unsigned long long s_lwords[2][kAggLevels];  // words holding a head, by tile parity
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    const StepDev* st = steps + b;
    SegState* s = S + b;
    const SegDesc& d = D[b];
    // per-vector values in one round trip (see kb_flags); m = tileoff[ntiles]
    const int active = st->active, stalled = s->stalled, cur = s->cur ^ s->flip, nlev = st->g.nlev,
              maxlev = d.maxlev;
    const int64_t n0 = s->n[0], n1 = s->n[1], m_scan = s->m, capC = d.capC;
    const uint64_t *Lz0 = d.Lz[0], *Lz1 = d.Lz[1], *Lx0 = d.Lx[0], *Lx1 = d.Lx[1];
    const CT *Lc0 = reinterpret_cast<const CT*>(d.Lc[0]), *Lc1 = reinterpret_cast<const CT*>(d.Lc[1]);
    const uint64_t* __restrict__ flags = d.flags;
    const uint16_t* __restrict__ wpre = d.wpre;
    const int64_t* __restrict__ tileoff = d.tileoff;
    const int64_t* __restrict__ bnd = d.bnd;
    const uint64_t* __restrict__ agg = d.agg;
    const int32_t* __restrict__ tilecnt = d.tilecnt;
    uint64_t* __restrict__ Cz = d.Cz;
    uint64_t* __restrict__ Cx = d.Cx;
    CT* __restrict__ Cc = reinterpret_cast<CT*>(d.Cc);
    GenReg<W> G;
    G.load(&st->g);
    const uint32_t lev4 = *reinterpret_cast<const uint32_t*>(st->g.lev);  // the first four levels
    asm volatile("" ::"s"(active), "s"(stalled), "s"(cur), "s"(nlev), "s"(maxlev), "s"(n0), "s"(n1), "s"(m_scan),
                 "s"(Lz0), "s"(Lz1), "s"(Lx0), "s"(Lx1), "s"(Lc0), "s"(Lc1), "s"(flags), "s"(wpre), "s"(tileoff),
                 "s"(bnd), "s"(Cz), "s"(Cx), "s"(Cc), "s"(G.gz[0]), "s"(G.gx[0]), "s"(G.gz[W - 1]),
                 "s"(G.gx[W - 1]), "s"(G.sinv), "s"(G.ng), "s"(lev4), "s"(agg), "s"(tilecnt));
    if (!active || stalled) return;
    const int64_t n = cur ? n1 : n0;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const int lane0 = threadIdx.x % kWave;
    // Child offsets: from kb_scan, or (self_off) summed here from the tile
    // counts, with kb_scan's bookkeeping (m, the child-buffer check, the merge
    // ticket).  The counts (at most kEmitSelfTiles) are loaded now and summed
    // after the first tile's own loads are issued: both share one memory
    // round trip.  Every wave sums for itself.
    constexpr int kCW = TABLE ? 1 : kEmitSelfTiles / kWave;  // the table variant runs after kb_scan
    int32_t cpre[kCW];
#pragma unroll
    for (int k = 0; k < kCW; ++k) {
        const int64_t u = lane0 + (int64_t)k * kWave;
        cpre[k] = (!TABLE && self_off && u < ntiles) ? ldg(tilecnt, u) : 0;
    }
    if (self_off && ntiles == 0) {
        if (lb == 0 && threadIdx.x == 0) {
            s->m = 0;
            s->ticket = 0u;
        }
        return;
    }
    const uint64_t* __restrict__ Lz = cur ? Lz1 : Lz0;
    const uint64_t* __restrict__ Lx = cur ? Lx1 : Lx0;
    const CT* __restrict__ Lc = cur ? Lc1 : Lc0;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    // cross-edge segment counts from kb_flags' summaries (1..kAggLevels
    // levels) or from kb_bounds' table
    // TABLE: the cross-edge counts come from kb_bounds' table (any number of
    // levels); otherwise from kb_flags' summaries (at most kAggLevels levels).
    // Two instantiations keep each variant's register use low.
    constexpr bool summ = !TABLE;
    // the head-mask level loop serves every mask of at most kAggLevels levels
    // (with the table's counts in the table variant); more levels, and the
    // table variant's 2048-term tiles (whose registers the loop would push
    // to two waves per SIMD), use scans
    const bool maskp = summ || (kTile <= 512 && nlev <= kAggLevels);
    if (!TABLE && nlev > kAggLevels) {  // the host never launches this: report, do not write
        if (lb == 0 && threadIdx.x == 0) atomicOr(&s->err, 8);
        return;
    }
    if (threadIdx.x < 2 * kAggLevels) s_lwords[threadIdx.x / kAggLevels][threadIdx.x % kAggLevels] = 0ull;
    __syncthreads();
    int64_t mtot = m_scan;
    int par = 0;  // tile parity: which s_lwords row this tile uses
    for (int64_t t = lb; t < ntiles; t += nb, par ^= 1) {
        const int64_t ts = t * kTile;
        const int Teff = (int)((n - ts) < kTile ? (n - ts) : kTile);
        Key<W> K[R];
        CT C[R];
        int64_t rank[R];
        bool fl[R];
        const int64_t toff_scan = self_off ? 0 : ldg(tileoff, t);
        // Large tiles of the table variant (long vectors, bandwidth-bound):
        // every load of the tile is issued before the tests.  Elsewhere the
        // rounds load in turn (fewer registers; the lattice-sized 512-term
        // tiles and the summary variant measured faster that way).
        if constexpr (TABLE && T >= 2048) {
            // every load of the tile first (one memory round trip): keys, flag
            // words, in-tile prefixes and coefficients (each coefficient line is
            // fetched for the flagged terms anyway), and the key before the
            // wave's first term (later rounds take theirs from lane 63 of the
            // round before)
            uint64_t wd[R];
            int wp[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                const int64_t i = ts + j;
                C[r] = coef_zero<CT>();
                wd[r] = 0ull;
                wp[r] = 0;
#pragma unroll
                for (int w = 0; w < W; ++w) K[r].z[w] = K[r].x[w] = 0;
                if (j < Teff) {
                    K[r] = load_key<W>(Lz, Lx, i);
                    wd[r] = ldg(flags, i >> 6);
                    wp[r] = (int)ldg(wpre, i >> 6);
                    C[r] = ldg(Lc, i);
                }
            }
            Key<W> kp0;
#pragma unroll
            for (int w = 0; w < W; ++w) kp0.z[w] = kp0.x[w] = 0;
            {
                const int64_t i0 = ts + wave * (kTile / 4);
                if (lane == 0 && i0 > 0 && wave * (kTile / 4) < Teff) kp0 = load_key<W>(Lz, Lx, i0 - 1);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                const int64_t i = ts + j;
                // the key at i - 1: from the lane below, lane 0 as above
                Key<W> kp;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    kp.z[w] = __shfl_up(K[r].z[w], 1, kWave);
                    kp.x[w] = __shfl_up(K[r].x[w], 1, kWave);
                }
                if (r == 0) {
                    if (lane == 0) kp = kp0;
                } else {
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        const uint64_t pz = __shfl(K[r - 1].z[w], kWave - 1, kWave);
                        const uint64_t px = __shfl(K[r - 1].x[w], kWave - 1, kWave);
                        if (lane == 0) {
                            kp.z[w] = pz;
                            kp.x[w] = px;
                        }
                    }
                }
                fl[r] = false;
                rank[r] = 0;
                if (j < Teff) {
                    const uint64_t word = wd[r];
                    fl[r] = (word >> lane) & 1ull;
                    const uint64_t lowmask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
                    const int aloc = wp[r] + popc64(word & lowmask);
                    s_aloc[j] = (int16_t)aloc;
                    s_hdb[j] = (int16_t)(i == 0 ? kHdbNone : hdb<W>(kp, K[r]));
                    rank[r] = aloc;
                    if (!fl[r]) C[r] = coef_zero<CT>();
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                const int64_t i = ts + j;
                fl[r] = false;
                rank[r] = 0;
                C[r] = coef_zero<CT>();
                if (j < Teff) {
                    K[r] = load_key<W>(Lz, Lx, i);
                    const uint64_t word = ldg(flags, i >> 6);
                    fl[r] = (word >> lane) & 1ull;
                    const uint64_t lowmask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
                    const int aloc = (int)ldg(wpre, i >> 6) + popc64(word & lowmask);
                    s_aloc[j] = (int16_t)aloc;
                    s_hdb[j] = (int16_t)(i == 0 ? kHdbNone : hdb<W>(load_key<W>(Lz, Lx, i - 1), K[r]));
                    rank[r] = aloc;
                    if (fl[r]) C[r] = ldg(Lc, i);
                } else {
#pragma unroll
                    for (int w = 0; w < W; ++w) K[r].z[w] = K[r].x[w] = 0;
                }
            }
        }
        if (maskp) {
            // head masks per level (a tile position 0 is a head of the tile's view)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
#pragma unroll
                for (int li = 0; li < kAggLevels; ++li) {
                    if (li < nlev) {
                        const int lvl = (int)((lev4 >> (8 * li)) & 255u);
                        const uint64_t hm = __ballot(j < Teff && (j == 0 || s_hdb[j] >= lvl));
                        if (lane == 0) {
                            s_lhm[li][wave * R + r] = hm;
                            if (hm) atomicOr(&s_lwords[par][li], 1ull << (wave * R + r));
                        }
                    }
                }
            }
        }
        int64_t toff = toff_scan;
        if (!TABLE && self_off) {
            int64_t all = 0, before = 0;
#pragma unroll
            for (int k = 0; k < kCW; ++k) {
                all += cpre[k];
                before += lane + (int64_t)k * kWave < t ? cpre[k] : 0;
            }
            mtot = wave_sum64(all);
            toff = wave_sum64(before);
            if (t == 0 && threadIdx.x == 0) {
                s->m = mtot;
                s->ticket = 0u;
                if (mtot > capC) {
                    s->stalled = 1;
                    s->stall_step = step_index;
                    s->need = mtot;
                }
            }
            if (mtot > capC) return;  // the same for every workgroup of the vector: nothing is written
        }
#pragma unroll
        for (int r = 0; r < R; ++r) rank[r] += toff;
        if (summ) {
            // wave w: the tile's start edge (w even) or end edge (w odd) at
            // levels w / 2 and w / 2 + 2
            const int side = wave & 1;
            for (int li = wave >> 1; li < nlev; li += 2) {
                const int lvl = (int)((lev4 >> (8 * li)) & 255u);
                int64_t same, sib;
                int h_edge;
                edge_counts<W, (kTile <= 512)>(side, li, lvl, t, ntiles, n, ts, ts + Teff, Lz, Lx, agg, tilecnt,
                                                 same, sib, h_edge);
                if (lane == 0) {
                    s_bq[li][2 * side] = same;
                    s_bq[li][2 * side + 1] = sib;
                    if (wave == 1 && li == 0) s_hdb[Teff] = (int16_t)h_edge;  // the end edge, read only by the level loop
                }
            }
            if (self_off && wave == 0) {
                // the tile's A-count, from the prefetched tile counts
                int32_t c = 0;
#pragma unroll
                for (int k = 0; k < kCW; ++k) c = (int64_t)k == (t >> 6) ? cpre[k] : c;
                c = __shfl(c, (int)(t & 63), kWave);
                if (lane == 0) s_aloc[Teff] = (int16_t)c;
            }
        }
        if (TABLE && maskp && threadIdx.x >= kWave - 4 * kAggLevels && threadIdx.x < kWave) {
            // the table's cross-edge counts for the level loop
            const int q = threadIdx.x - (kWave - 4 * kAggLevels), li = q >> 2;
            if (li < nlev) s_bq[li][q & 3] = ldg(bnd, (t * maxlev + li) * 4 + (q & 3));
        }
        if (threadIdx.x == 0 && !(summ && self_off)) {
            s_aloc[Teff] = (int16_t)(self_off ? ldg(tilecnt, t) : ldg(tileoff, t + 1) - toff_scan);
            if (!summ)
                s_hdb[Teff] = (int16_t)(ts + Teff < n ? hdb<W>(load_key<W>(Lz, Lx, ts + Teff - 1),
                                                               load_key<W>(Lz, Lx, ts + Teff))
                                                      : kHdbNone);
        }
        __syncthreads();
        if (maskp) {
            // the level loop on the head masks: no scans, no barriers
            for (int li = 0; li < nlev; ++li) {
                const int lvl = (int)((lev4 >> (8 * li)) & 255u);
                const uint64_t* hm = s_lhm[li];
                const uint64_t words = s_lwords[par][li];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (!fl[r]) continue;
                    const int j = wave * (kTile / 4) + r * kWave + lane;
                    int64_t cnt = 0;
                    if (key_bit<W>(K[r], lvl) == 0) {
                        const int ee = next_head(hm, words, j, Teff);
                        if (ee < Teff) {
                            if (s_hdb[ee] == lvl) {
                                const int f = next_head(hm, words, ee, Teff);
                                cnt = (int64_t)s_aloc[f] - s_aloc[ee];
                                if (f == Teff && s_hdb[Teff] < lvl) cnt += s_bq[li][2];
                            }
                        } else {
                            const int h = s_hdb[Teff];
                            cnt = h == lvl ? s_bq[li][2] : (h < lvl ? s_bq[li][3] : 0);
                        }
                        rank[r] += cnt;
                    } else {
                        const int sv = last_head(hm, words, j);
                        if (sv > 0) {
                            if (s_hdb[sv] == lvl) {
                                const int s2 = last_head(hm, words, sv - 1);
                                cnt = (int64_t)s_aloc[sv] - s_aloc[s2];
                                if (s2 == 0 && s_hdb[0] < lvl) cnt += s_bq[li][0];
                            }
                        } else {
                            const int h = s_hdb[0];
                            cnt = h == lvl ? s_bq[li][0] : (h < lvl ? s_bq[li][1] : 0);
                        }
                        rank[r] -= cnt;
                    }
                }
            }
        }
        for (int li = 0; li < (maskp ? 0 : nlev); ++li) {
            const int lvl = li < 4 ? (int)((lev4 >> (8 * li)) & 255u) : (int)st->g.lev[li];
            int64_t bq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) bq[q] = ldg(bnd, (t * maxlev + li) * 4 + q);
            // forward: st_j = last head <= j (j = 0 is always a head)
            int stv[R];
            int carry = -1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = wave * (kTile / 4) + r * kWave + lane;
                int v = (j < Teff && (j == 0 || s_hdb[j] >= lvl)) ? j : -1;
#pragma unroll
                for (int dd = 1; dd < kWave; dd <<= 1) {
                    const int u = __shfl_up(v, dd, kWave);
                    if (lane >= dd) v = v > u ? v : u;
                }
                v = v > carry ? v : carry;
                stv[r] = v;
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
                int v = (j < Teff && (j == 0 || s_hdb[j] >= lvl)) ? j : Teff;
#pragma unroll
                for (int dd = 1; dd < kWave; dd <<= 1) {
                    const int u = __shfl_down(v, dd, kWave);
                    if (lane + dd < kWave) v = v < u ? v : u;
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
                stv[r] = stv[r] > pre ? stv[r] : pre;
                e[r] = e[r] < post ? e[r] : post;
                if (j < Teff) {
                    s_st[j] = (int16_t)stv[r];
                    s_e[j] = (int16_t)e[r];
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!fl[r]) continue;
                int64_t cnt = 0;
                if (key_bit<W>(K[r], lvl) == 0) {
                    const int ee = e[r];
                    if (ee < Teff) {
                        if (s_hdb[ee] == lvl) {
                            const int f = s_e[ee];
                            cnt = (int64_t)s_aloc[f] - s_aloc[ee];
                            if (f == Teff && s_hdb[Teff] < lvl) cnt += bq[2];
                        }
                    } else {
                        const int h = s_hdb[Teff];
                        cnt = h == lvl ? bq[2] : (h < lvl ? bq[3] : 0);
                    }
                    rank[r] += cnt;
                } else {
                    const int sv = stv[r];
                    if (sv > 0) {
                        if (s_hdb[sv] == lvl) {
                            const int s2 = s_st[sv - 1];
                            cnt = (int64_t)s_aloc[sv] - s_aloc[s2];
                            if (s2 == 0 && s_hdb[0] < lvl) cnt += bq[0];
                        }
                    } else {
                        const int h = s_hdb[0];
                        cnt = h == lvl ? bq[0] : (h < lvl ? bq[1] : 0);
                    }
                    rank[r] -= cnt;
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!fl[r]) continue;
            if (rank[r] < 0 || rank[r] >= mtot) {  // never expected: report, do not write
                atomicOr(&s->err, 1);
                continue;
            }
            Key<W> ch;
            const CT cn = G.template child<CT>(K[r], C[r], ch);
            store_key<W>(Cz, Cx, rank[r], ch);
            stg(Cc, rank[r], cn);
        }
        if (threadIdx.x < kAggLevels) s_lwords[par ^ 1][threadIdx.x] = 0ull;  // for the next tile
        __syncthreads();
    }
}

// Threads per kb_merge workgroup: E = TM / threads outputs each.
constexpr int merge_threads(int TM) { return TM >= 2048 ? 2 * kBlock : kBlock; }
// Waves per SIMD the register budget must allow: three 512-thread
// workgroups per CU where their LDS fits (48 KiB each at W = 1 with FP64
// coefficients; two with complex ones, 64 KiB).
constexpr int merge_waves(int TM, int coef_bytes) { return TM >= 2048 ? (coef_bytes == 8 ? 6 : 4) : 1; }

// ---------------------------------------------------------------------------
// kb_merge (see the file header).  mode 0: rotation (live + children);
// mode 1: early-merge compaction (live only, no rotation) for the vectors
// kb_precheck flagged.  A tile only ever waits in the look-back on tiles of
// workgroups that have started: with a workgroup per tile, tile = local
// workgroup index; with fewer workgroups, tiles are claimed by ticket.
template <int W, int TM, typename CT>
__global__ void __launch_bounds__(merge_threads(TM), merge_waves(TM, (int)sizeof(CT))) kb_merge(BlockMap bm, const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                   const StepDev* __restrict__ steps, int mode,
                                                   int64_t step_index, const uint32_t* __restrict__ d_epoch,
                                                   uint32_t ep_off) {
    constexpr int NT = merge_threads(TM), NW = NT / kWave;
    constexpr int E = TM / NT;
    __shared__ Key<W> s_key[TM];
    __shared__ CT s_c[TM];
    __shared__ int64_t s_tile;
    __shared__ int64_t s_base;
    __shared__ int64_t s_split[2];
    __shared__ Key<W> s_edge[2];   // C[j1] (first child after the tile), L[i0 - 1] (last live term before it)
    __shared__ CT s_edge_c;    // c of C[j1]
    __shared__ int s_wsum[NW];
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    const StepDev* st = steps + b;
    SegState* s = S + b;
    const SegDesc& d = D[b];
    const int rot = mode == 0;
    // per-vector values in one round trip (see kb_flags)
    const int active = st->active, stalled = s->stalled, compact = s->compact, cur = s->cur ^ s->flip;
    const uint32_t ep0 = *d_epoch;
    const int64_t nstep = s->nstep, m0 = s->m, capOut = d.cap;
    uint64_t *Lz0 = d.Lz[0], *Lz1 = d.Lz[1], *Lx0 = d.Lx[0], *Lx1 = d.Lx[1];
    CT *Lc0 = reinterpret_cast<CT*>(d.Lc[0]), *Lc1 = reinterpret_cast<CT*>(d.Lc[1]);
    const uint64_t* __restrict__ Cz = d.Cz;
    const uint64_t* __restrict__ Cx = d.Cx;
    const CT* __restrict__ Cc = reinterpret_cast<const CT*>(d.Cc);
    uint64_t* __restrict__ status = d.status;
    FilterReg fs;
    fs.load(d.fs);
    GenReg<W> G;
    G.load(&st->g);
    asm volatile("" ::"s"(active), "s"(stalled), "s"(compact), "s"(cur), "s"(ep0), "s"(nstep), "s"(m0),
                 "s"(capOut), "s"(Lz0), "s"(Lz1), "s"(Lx0), "s"(Lx1), "s"(Lc0), "s"(Lc1), "s"(Cz), "s"(Cx), "s"(Cc),
                 "s"(status), "s"(G.gz[0]), "s"(G.gx[0]), "s"(G.gz[W - 1]), "s"(G.gx[W - 1]), "s"(G.cosv),
                 "s"(fs.wmax), "s"(fs.xwmax), "s"(fs.mwmax), "s"(fs.thresh), "s"(fs.dthresh_w), "s"(fs.dthresh_xw));
    if (stalled || !(rot ? active : compact)) return;
    // launch tag of the look-back words: the sequence's base (device memory,
    // advanced after every sequence) plus this launch's offset in it
    const uint64_t ep = ((uint64_t)(ep0 + ep_off) & kEpochMask) << kEpochShift;
    // n from the step's snapshot, not n[cur]: the last tile toggles the
    // parity, and a workgroup without a tile may read it afterwards; its tile
    // count must not change (the parity is used only by workgroups with a
    // tile, which read it before the last tile can finish)
    const int64_t n = nstep;
    const int64_t m = rot ? m0 : 0;
    const int64_t tot = n + m;
    const int64_t ntm = (tot + TM - 1) / TM;
    const uint64_t* __restrict__ Lz = cur ? Lz1 : Lz0;
    const uint64_t* __restrict__ Lx = cur ? Lx1 : Lx0;
    const CT* __restrict__ Lc = cur ? Lc1 : Lc0;
    uint64_t* __restrict__ Oz = cur ? Lz0 : Lz1;
    uint64_t* __restrict__ Ox = cur ? Lx0 : Lx1;
    CT* __restrict__ Oc = cur ? Lc0 : Lc1;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    if (ntm == 0) {  // nothing to merge: the output is empty
        if (lb == 0 && tid == 0) {
            s->n[cur ^ 1] = 0;
            if (rot) s->rotations += 1;
            else s->compact = 0;
            s->flip ^= 1;
        }
        return;
    }
    // Tiles: when the vector has a workgroup per tile, workgroup lb takes
    // tile lb (workgroups start in index order, so every tile a look-back
    // waits on belongs to a started workgroup); otherwise workgroups take
    // tiles by ticket.
    const bool by_ticket = nb < ntm;
    for (int64_t it = 0;; ++it) {
        int64_t t;
        if (by_ticket) {
            if (tid == 0) s_tile = (int64_t)atomicAdd(&s->ticket, 1u);
            __syncthreads();
            t = s_tile;
        } else {
            t = it == 0 ? lb : ntm;
        }
        if (t >= ntm) break;
        const int64_t d0 = t * TM;
        const int64_t d1 = d0 + TM < tot ? d0 + TM : tot;
        // merge-path splits of this tile's window: wave 0 the start, wave 1 the end
        if (wave < 2) {
            const int64_t sp = merge_split_wave<W, true>(Lz, Lx, n, Cz, Cx, m, wave == 0 ? d0 : d1);  // z words first
            if (lane == 0) s_split[wave] = sp;
        }
        __syncthreads();
        int64_t i0 = s_split[0], i1 = s_split[1];
        int64_t j0 = d0 - i0, j1 = d1 - i1;
        if (i0 < 0 || i1 < i0 || i1 > n || j0 < 0 || j1 < j0 || j1 > m) {  // never expected
            if (tid == 0) atomicOr(&s->err, 2);
            i1 = i0 = 0;
            j1 = j0 = 0;
        }
        const int nL = (int)(i1 - i0), nC = (int)(j1 - j0);
        {
            // every load of the tile first (one memory round trip), then LDS
            Key<W> lk[E];
            CT lcv[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = tid + e * NT;
#pragma unroll
                for (int w = 0; w < W; ++w) lk[e].z[w] = lk[e].x[w] = 0;
                lcv[e] = coef_zero<CT>();
                if constexpr (TM >= 2048) {  // large vectors: read once, past the caches
                    if (k < nL) {
                        lk[e] = load_key_nt<W>(Lz, Lx, i0 + k);
                        lcv[e] = ldg_nt(Lc, i0 + k);
                    } else if (k < nL + nC) {
                        lk[e] = load_key_nt<W>(Cz, Cx, j0 + (k - nL));
                        lcv[e] = ldg_nt(Cc, j0 + (k - nL));
                    }
                } else {
                    if (k < nL) {
                        lk[e] = load_key<W>(Lz, Lx, i0 + k);
                        lcv[e] = ldg(Lc, i0 + k);
                    } else if (k < nL + nC) {
                        lk[e] = load_key<W>(Cz, Cx, j0 + (k - nL));
                        lcv[e] = ldg(Cc, j0 + (k - nL));
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = tid + e * NT;
                if (k < nL + nC) {
                    s_key[k] = lk[e];
                    s_c[k] = lcv[e];
                }
            }
        }
        // the neighbours across the tile's edges, loaded with the tile
        if (tid == NT - 1 && j1 < m) {
            s_edge[0] = load_key<W>(Cz, Cx, j1);
            s_edge_c = ldg(Cc, j1);
        }
        if (tid == NT - 2 && i0 > 0) s_edge[1] = load_key<W>(Lz, Lx, i0 - 1);
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
                    if (key_eq<W>(s_edge[0], K)) acc = acc + s_edge_c;
                }
            } else {
                K = sC[jb];
                acc = sCc[jb];
                ++jb;
                // absorbed by an equal live key, which precedes it directly
                if (ia > 0) {
                    if (key_eq<W>(sL[ia - 1], K)) emit = false;
                } else if (i0 > 0) {
                    if (key_eq<W>(s_edge[1], K)) emit = false;
                }
            }
            outK[e] = K;
            outC[e] = acc;
            outE[e] = emit && !should_drop<W>(fs, K, coef_abs(acc));
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
        for (int w = 0; w < NW; ++w) {
            if (w < wave) wpref += s_wsum[w];
            total += s_wsum[w];
        }
        const int myoff = wpref + incl - cnt;
        // kept outputs staged in LDS at their tile offsets (every LDS read of
        // the merge ended at the barrier above); waves 1-3 stage while wave 0
        // runs the look-back, wave 0 after it
        auto stage = [&]() {
            int k = myoff;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (outE[e]) {
                    s_key[k] = outK[e];
                    s_c[k] = outC[e];
                    ++k;
                }
            }
        };
        if (wave != 0) stage();
        // Decoupled look-back for the tile's base, one wave scanning 64
        // predecessors per step.  Each status word carries its own payload
        // (flag | count), so relaxed atomics suffice: no other memory is
        // published through it, and no cache write-back/invalidate is needed.
        if (wave == 0) {
            if (t == 0) {
                if (lane == 0) {
                    __hip_atomic_store((PP_GLOBAL uint64_t*)&status[0], kFlagPre | ep | (uint64_t)total, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    s_base = 0;
                }
            } else {
                if (lane == 0)
                    __hip_atomic_store((PP_GLOBAL uint64_t*)&status[t], kFlagAgg | ep | (uint64_t)total, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                int64_t excl = 0;
                int64_t top = t - 1;
                int64_t spins = 0;
                while (true) {
                    const int64_t q = top - lane;
                    const uint64_t sv = q >= 0 ? __hip_atomic_load((PP_GLOBAL uint64_t*)&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                               : (kFlagPre | ep);  // before tile 0: prefix 0
                    const bool fresh = (sv & (kEpochMask << kEpochShift)) == ep;
                    const uint64_t flag = fresh ? (sv & (3ull << 62)) : 0ull;
                    const uint64_t pre = __ballot(flag == kFlagPre);
                    const uint64_t notready = __ballot(flag == 0);
                    const int first = pre ? __ffsll((unsigned long long)pre) - 1 : kWave;
                    const uint64_t need = first >= kWave - 1 ? ~0ull : ((1ull << (first + 1)) - 1);
                    if (notready & need) {
                        if (++spins > (1ll << 26)) {  // bounded wait: report instead of hanging
                            if (lane == 0) atomicOr(&s->err, 4);
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
                    __hip_atomic_store((PP_GLOBAL uint64_t*)&status[t], kFlagPre | ep | (uint64_t)(excl + total), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    s_base = excl;
                }
            }
            if (lane == 0 && t == ntm - 1) {
                // last tile: publish the vector's new count (or freeze it)
                const int64_t n_out = s_base + total;
                if (n_out > capOut) {
                    s->stalled = 2;
                    s->stall_step = step_index;
                    s->need = n_out;
                } else {
                    s->n[cur ^ 1] = n_out;
                    if (rot) {
                        s->rotations += 1;
                        s->swept += n;
                        s->created += m;
                        s->merge_in += n + m;
                        s->merge_out += n_out;
                        // _merge_spv! growth of the reference's live buffers (spv_kernels.jl:377-383)
                        if (n_out > s->live_cap_v)
                            s->live_cap_v = 2 * s->live_cap_v > n + m ? 2 * s->live_cap_v : n + m;
                    } else {
                        // early merge (spv_evolve.jl:356-362): grow the append length if still short
                        s->merge_out += n_out;
                        if (n_out > s->app_cap_v) s->app_cap_v = 2 * s->app_cap_v > n_out ? 2 * s->app_cap_v : n_out;
                        s->compact = 0;
                    }
                    s->flip ^= 1;
                }
            }
        }
        if (wave == 0) stage();
        __syncthreads();
        const int64_t base = s_base;
        if (base + total <= capOut) {
            for (int k = tid; k < total; k += NT) {
                if constexpr (TM >= 2048) {  // large vectors: stream past the caches
                    store_key_nt<W>(Oz, Ox, base + k, s_key[k]);
                    stg_nt(Oc, base + k, s_c[k]);
                } else {
                    store_key<W>(Oz, Ox, base + k, s_key[k]);
                    stg(Oc, base + k, s_c[k]);
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// kb_precheck: the capacity precheck before the first rotation of a sequence
// (spv_evolve.jl:356-362): when n exceeds the reference's append length the
// reference merges early, which applies the strict filter to every live
// term; flag those vectors for a compaction.  One thread per vector.
__global__ void kb_precheck(SegState* __restrict__ S, const int32_t* __restrict__ en, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    SegState* s = S + b;
    if (s->flip) {
        s->cur ^= 1;
        s->flip = 0;
    }
    s->compact = 0;
    s->ticket = 0u;
    if (!en[b] || s->stalled) return;
    const int64_t n = s->n[s->cur];
    s->nstep = n;
    if (n > s->app_cap_v) {
        s->compact = 1;
        s->early += 1;
        s->merge_in += n;
    }
}

// Advance the look-back tag base past the `count` merge launches of the
// sequence just enqueued.
__global__ void kb_epoch_advance(uint32_t* __restrict__ d_epoch, uint32_t count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d_epoch += count;
}

// Resolve pending parity flips (before host reads and between sequences).
__global__ void kb_resolve(SegState* __restrict__ S, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    SegState* s = S + b;
    if (s->flip) {
        s->cur ^= 1;
        s->flip = 0;
    }
}

// ---------------------------------------------------------------------------
// Channels (channels.jl:203-265): c *= table[popcount] in place.
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) kb_channel(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                     const ChanDev* __restrict__ chans, const int32_t* __restrict__ en,
                                                     BlockMap bm) {
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    if (!en[b]) return;
    const SegState* s = S + b;
    if (s->stalled) return;
    const ChanDev ch = chans[b];
    if (ch.kind < 0) return;
    const int cur = s->cur ^ s->flip;
    const int64_t n = s->n[cur];
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    CT* __restrict__ Lc = reinterpret_cast<CT*>(d.Lc[cur]);
    for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += nb * kBlock) {
        const Key<W> k = load_key<W>(Lz, Lx, i);
        Lc[i] = Lc[i] * chan_factor<W>(ch, k);
    }
}

// ---------------------------------------------------------------------------
// <k|O|k> (spv_kernels.jl:441-452) with a double-double accumulator so the
// parallel sum is accurate to ~2^-106 of sum |c| before its one rounding.
// Partials per block, reduced per vector in block order (deterministic).
// ComplexF64 coefficients: one launch per component (comp 0 real, 1 imaginary;
// the sum of (±1) c is componentwise).
template <int W, typename CT>
__global__ void __launch_bounds__(kBlock) kb_expect(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                    const uint64_t* __restrict__ kets, BlockMap bm,
                                                    double2* __restrict__ partial, int comp) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    const SegState* s = S + b;
    const int cur = s->cur ^ s->flip;
    const int64_t n = s->n[cur];
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    const CT* __restrict__ Lc = reinterpret_cast<const CT*>(d.Lc[cur]);
    uint64_t kw[W];
#pragma unroll
    for (int w = 0; w < W; ++w) kw[w] = kets[(int64_t)b * W + w];
    double h = 0.0, l = 0.0;
    for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += nb * kBlock) {
        const Key<W> k = load_key<W>(Lz, Lx, i);
        bool diag = true;
        int par = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            diag = diag && k.x[w] == 0;
            par += popc64(k.z[w] & kw[w]);
        }
        if (diag) {
            const CT c = Lc[i];
            const double cv = comp ? coef_im(c) : coef_re(c);
            const double v = (par & 1) ? -cv : cv;
            dd_add(h, l, v, 0.0);
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
    if (threadIdx.x == 0) partial[bm.pre[b] + lb] = make_double2(s_h[0], s_l[0]);
}

__global__ void __launch_bounds__(kBlock) kb_expect_final(const double2* __restrict__ partial, BlockMap bm,
                                                          double* __restrict__ out, int stride, int comp) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    const int b = blockIdx.x;
    const int64_t p0 = bm.pre[b], p1 = bm.pre[b + 1];
    double h = 0.0, l = 0.0;
    for (int64_t i = p0 + threadIdx.x; i < p1; i += kBlock) dd_add(h, l, partial[i].x, partial[i].y);
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
    if (threadIdx.x == 0) out[(int64_t)b * stride + comp] = s_h[0] + s_l[0];
}

// check_spv key order (type_SparsePauliVector.jl:273-276): bad[b] = 1 if any
// adjacent pair of vector b is not strictly increasing.
template <int W>
__global__ void __launch_bounds__(kBlock) kb_check(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                   BlockMap bm, int32_t* __restrict__ bad) {
    int b;
    int64_t lb, nb;
    if (!map_block(bm, b, lb, nb)) return;
    const SegState* s = S + b;
    const int cur = s->cur ^ s->flip;
    const int64_t n = s->n[cur];
    const SegDesc& d = D[b];
    for (int64_t i = lb * kBlock + threadIdx.x + 1; i < n; i += nb * kBlock) {
        if (!key_lt<W>(load_key<W>(d.Lz[cur], d.Lx[cur], i - 1), load_key<W>(d.Lz[cur], d.Lx[cur], i))) bad[b] = 1;
    }
}

}  // namespace pp
