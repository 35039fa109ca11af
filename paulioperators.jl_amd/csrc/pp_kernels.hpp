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
//   kb_scan      exclusive scan of tile counts -> child offsets, m; applies
//                the reference's append-capacity precheck bookkeeping
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

// Grid mapping shared by every kernel: blocks [pre[b], pre[b+1]) serve vector b.
struct BlockMap {
    const int64_t* pre;
    int B;
};

__device__ __forceinline__ void map_block(const BlockMap& bm, int& b, int64_t& lb, int64_t& nb) {
    b = find_seg(bm.pre, bm.B, blockIdx.x);
    lb = (int64_t)blockIdx.x - bm.pre[b];
    nb = bm.pre[b + 1] - bm.pre[b];
}

// ---------------------------------------------------------------------------
template <int W, int T>
__global__ void __launch_bounds__(kBlock) kb_flags(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                   const StepDev* __restrict__ steps, BlockMap bm) {
    constexpr int kTile = T;
    constexpr int kTileWords = T / 64;
    __shared__ int s_wcnt[kTileWords];
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
    const StepDev* st = steps + b;
    if (!st->active) return;
    const SegState* s = S + b;
    if (s->stalled) return;
    const int cur = s->cur ^ s->flip;
    const int64_t n = s->n[cur];
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    const double* __restrict__ Lc = d.Lc[cur];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    GenReg<W> G;
    G.load(&st->g);
    const int local_active = d.local_active;
    bool need_z = false, need_x = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        need_z |= G.gx[w] != 0;
        need_x |= G.gz[w] != 0;
    }
    if (local_active) need_z = need_x = true;
    for (int64_t t = lb; t < ntiles; t += nb) {
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
                        const ulonglong2 c2 = reinterpret_cast<const ulonglong2*>(Lx)[i];
                        k.x[0] = c2.x; k.x[1] = c2.y;
                    } else {
                        k.x[0] = k.x[1] = 0;
                    }
                }
                flag = G.anti(k);
                if (flag && local_active) {
                    Key<W> ch;
                    const double cn = G.child(k, Lc[i], ch);
                    flag = !should_drop<W>(d.fl, ch, fabs(cn));
                }
            }
            const uint64_t mask = __ballot(flag);
            if (lane == 0) {
                if (t * kTile + wi * 64 < n) d.flags[t * kTileWords + wi] = mask;
                s_wcnt[wi] = popc64(mask);
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
            if (lane < kTileWords && t * kTile + lane * 64 < n) d.wpre[t * kTileWords + lane] = (uint16_t)(incl - v);
            if (lane == kTileWords - 1) d.tilecnt[t] = incl;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// kb_scan: one workgroup per vector.  Resolves the previous merge's parity
// flip, runs the capacity precheck bookkeeping of spv_evolve.jl:356-362 for
// rotations after the first of a sequence (where the early merge cannot drop
// anything: every live term passed the strict filter at the last merge),
// scans the tile counts and checks the child buffer.
template <int T>
__global__ void __launch_bounds__(1024) kb_scan(const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                const StepDev* __restrict__ steps, int64_t step_index) {
    constexpr int kTile = T;
    __shared__ int64_t s_w[16];
    __shared__ int64_t s_carry;
    __shared__ int64_t s_n;
    __shared__ int s_go;
    const int b = blockIdx.x;
    SegState* s = S + b;
    const StepDev* st = steps + b;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    if (tid == 0) {
        if (s->flip) {
            s->cur ^= 1;
            s->flip = 0;
        }
        s_go = st->active && !s->stalled;
        s_n = s->n[s->cur];
        s_carry = 0;
        s->ticket = 0u;
        if (s_go && st->precheck && s_n > s->app_cap_v) {
            s->early += 1;
            s->merge_in += s_n;
            s->merge_out += s_n;
            s->app_cap_v = 2 * s->app_cap_v > s_n ? 2 * s->app_cap_v : s_n;
        }
    }
    __syncthreads();
    if (!s_go) return;
    const SegDesc& d = D[b];
    const int64_t n = s_n;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    for (int64_t base = 0; base < ntiles; base += 1024) {
        const int64_t t = base + tid;
        const int64_t v = t < ntiles ? d.tilecnt[t] : 0;
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
        if (t < ntiles) d.tileoff[t] = carry + wpref + incl - v;
        __syncthreads();
        if (tid == 1023) s_carry = carry + wpref + incl;
        __syncthreads();
    }
    if (tid == 0) {
        const int64_t m = s_carry;
        d.tileoff[ntiles] = m;
        s->m = m;
        if (m > d.capC) {
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
__global__ void __launch_bounds__(kBlock) kb_bounds(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                    const StepDev* __restrict__ steps, BlockMap bm) {
    constexpr int kTile = T;
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
    const StepDev* st = steps + b;
    if (!st->active) return;
    const SegState* s = S + b;
    if (s->stalled) return;
    const int cur = s->cur;
    const int64_t n = s->n[cur];
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    const int nlev = st->g.nlev;
    const int maxlev = d.maxlev;
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
            const int lvl = st->g.lev[li];
            int64_t same = 0, sib = 0;
            if (h <= lvl) {
                if (side == 0) {
                    const int64_t q = seg_start_wave<W>(Lz, Lx, ts - 1, lvl, ka);
                    same = a_prefix<T>(ts, n, ntiles, d.flags, d.wpre, d.tileoff) -
                           a_prefix<T>(q, n, ntiles, d.flags, d.wpre, d.tileoff);
                    if (h < lvl && key_bit<W>(ka, lvl) == 1 && q > 0) {
                        kb2 = load_key<W>(Lz, Lx, q - 1);
                        if (hdb<W>(kb2, load_key<W>(Lz, Lx, q)) == lvl) {
                            const int64_t q2 = seg_start_wave<W>(Lz, Lx, q - 1, lvl, kb2);
                            sib = a_prefix<T>(q, n, ntiles, d.flags, d.wpre, d.tileoff) -
                                  a_prefix<T>(q2, n, ntiles, d.flags, d.wpre, d.tileoff);
                        }
                    }
                } else {
                    const int64_t r = seg_end_wave<W>(Lz, Lx, te, n, lvl, ka);
                    same = a_prefix<T>(r, n, ntiles, d.flags, d.wpre, d.tileoff) -
                           a_prefix<T>(te, n, ntiles, d.flags, d.wpre, d.tileoff);
                    if (h < lvl && key_bit<W>(ka, lvl) == 0 && r < n) {
                        kb2 = load_key<W>(Lz, Lx, r);
                        if (hdb<W>(load_key<W>(Lz, Lx, r - 1), kb2) == lvl) {
                            const int64_t r2 = seg_end_wave<W>(Lz, Lx, r, n, lvl, kb2);
                            sib = a_prefix<T>(r2, n, ntiles, d.flags, d.wpre, d.tileoff) -
                                  a_prefix<T>(r, n, ntiles, d.flags, d.wpre, d.tileoff);
                        }
                    }
                }
            }
            if (lane == 0) {
                d.bnd[(t * maxlev + li) * 4 + 2 * side + 0] = same;
                d.bnd[(t * maxlev + li) * 4 + 2 * side + 1] = sib;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// kb_emit: children written at their rank in key order.
// Layout inside a tile: wave w owns elements [w T/4, (w+1) T/4), round r
// covers 64 consecutive elements (coalesced loads), lane = element % 64.
template <int W, int T>
__global__ void __launch_bounds__(kBlock) kb_emit(const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                  const StepDev* __restrict__ steps, BlockMap bm) {
    constexpr int kTile = T;
    constexpr int R = kTile / kBlock;  // rounds of 64 consecutive terms per wave
    __shared__ int16_t s_hdb[kTile + 1];
    __shared__ int16_t s_aloc[kTile + 1];
    __shared__ int16_t s_st[kTile];
    __shared__ int16_t s_e[kTile];
    __shared__ int s_wred[4];
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
    const StepDev* st = steps + b;
    if (!st->active) return;
    SegState* s = S + b;
    if (s->stalled) return;
    const int cur = s->cur;
    const int64_t n = s->n[cur];
    const int64_t ntiles = (n + kTile - 1) / kTile;
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    const double* __restrict__ Lc = d.Lc[cur];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    GenReg<W> G;
    G.load(&st->g);
    const int nlev = st->g.nlev;
    const int maxlev = d.maxlev;
    const int64_t mtot = d.tileoff[ntiles];
    for (int64_t t = lb; t < ntiles; t += nb) {
        const int64_t ts = t * kTile;
        const int Teff = (int)((n - ts) < kTile ? (n - ts) : kTile);
        Key<W> K[R];
        double C[R];
        int64_t rank[R];
        bool fl[R];
        const int64_t toff = d.tileoff[t];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = wave * (kTile / 4) + r * kWave + lane;
            const int64_t i = ts + j;
            fl[r] = false;
            rank[r] = 0;
            C[r] = 0.0;
            if (j < Teff) {
                K[r] = load_key<W>(Lz, Lx, i);
                const uint64_t word = d.flags[i >> 6];
                fl[r] = (word >> lane) & 1ull;
                const uint64_t lowmask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
                const int aloc = (int)d.wpre[i >> 6] + popc64(word & lowmask);
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
            s_aloc[Teff] = (int16_t)(d.tileoff[t + 1] - toff);
            s_hdb[Teff] = (int16_t)(ts + Teff < n ? hdb<W>(load_key<W>(Lz, Lx, ts + Teff - 1), load_key<W>(Lz, Lx, ts + Teff))
                                                  : kHdbNone);
        }
        __syncthreads();
        for (int li = 0; li < nlev; ++li) {
            const int lvl = st->g.lev[li];
            const int64_t* bq = d.bnd + (t * maxlev + li) * 4;
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
            const double cn = G.child(K[r], C[r], ch);
            store_key<W>(d.Cz, d.Cx, rank[r], ch);
            d.Cc[rank[r]] = cn;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// mode 0: rotation (live + children); mode 1: early-merge compaction (live
// only, no rotation) for the vectors kb_precheck flagged.
__device__ __forceinline__ bool merge_go(const StepDev* st, const SegState* s, int mode) {
    if (s->stalled) return false;
    return mode == 0 ? st->active != 0 : s->compact != 0;
}

// ---------------------------------------------------------------------------
// kb_merge (see the file header).  Each workgroup takes tiles of its vector
// by ticket, so a tile only ever waits in the look-back on tiles that running
// workgroups have already claimed.
template <int W, int TM>
__global__ void __launch_bounds__(kBlock) kb_merge(const SegDesc* __restrict__ D, SegState* __restrict__ S,
                                                   const StepDev* __restrict__ steps, BlockMap bm, int mode,
                                                   int64_t step_index, const uint32_t* __restrict__ d_epoch,
                                                   uint32_t ep_off) {
    constexpr int E = TM / kBlock;
    __shared__ Key<W> s_key[TM];
    __shared__ double s_c[TM];
    __shared__ int64_t s_tile;
    __shared__ int64_t s_base;
    __shared__ int64_t s_split[2];
    __shared__ int s_wsum[4];
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
    const StepDev* st = steps + b;
    SegState* s = S + b;
    if (!merge_go(st, s, mode)) return;
    const int rot = mode == 0;
    // launch tag of the look-back words: the sequence's base (device memory,
    // advanced after every sequence) plus this launch's offset in it
    const uint64_t ep = ((uint64_t)(*d_epoch + ep_off) & kEpochMask) << kEpochShift;
    const int cur = s->cur;
    const int64_t n = s->n[cur];
    const int64_t m = rot ? s->m : 0;
    const int64_t tot = n + m;
    const int64_t ntm = (tot + TM - 1) / TM;
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    const double* __restrict__ Lc = d.Lc[cur];
    uint64_t* __restrict__ Oz = d.Lz[cur ^ 1];
    uint64_t* __restrict__ Ox = d.Lx[cur ^ 1];
    double* __restrict__ Oc = d.Lc[cur ^ 1];
    const uint64_t* __restrict__ Cz = d.Cz;
    const uint64_t* __restrict__ Cx = d.Cx;
    const double* __restrict__ Cc = d.Cc;
    const int64_t capOut = d.cap;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    GenReg<W> G;
    if (rot) G.load(&st->g);
    if (ntm == 0) {  // nothing to merge: the output is empty
        if (lb == 0 && tid == 0) {
            s->n[cur ^ 1] = 0;
            if (rot) s->rotations += 1;
            else s->compact = 0;
            s->flip = 1;
        }
        return;
    }
    while (true) {
        if (tid == 0) s_tile = (int64_t)atomicAdd(&s->ticket, 1u);
        __syncthreads();
        const int64_t t = s_tile;
        if (t >= ntm) break;
        const int64_t d0 = t * TM;
        const int64_t d1 = d0 + TM < tot ? d0 + TM : tot;
        // merge-path splits of this tile's window: wave 0 the start, wave 1 the end
        if (wave < 2) {
            const int64_t sp = merge_split_wave<W>(Lz, Lx, n, Cz, Cx, m, wave == 0 ? d0 : d1);
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
            if (emit && !should_drop<W>(d.fs, K, fabs(acc))) {
                outK[cnt] = K;
                outC[cnt] = acc;
                ++cnt;
            }
        }
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
        // Decoupled look-back for the tile's base, one wave scanning 64
        // predecessors per step.  Each status word carries its own payload
        // (flag | count), so relaxed atomics suffice: no other memory is
        // published through it, and no cache write-back/invalidate is needed.
        uint64_t* __restrict__ status = d.status;
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
                int64_t excl = 0;
                int64_t top = t - 1;
                int64_t spins = 0;
                while (true) {
                    const int64_t q = top - lane;
                    const uint64_t sv = q >= 0 ? __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
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
                    __hip_atomic_store(&status[t], kFlagPre | ep | (uint64_t)(excl + total), __ATOMIC_RELAXED,
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
                    s->flip = 1;
                }
            }
        }
        __syncthreads();
        const int64_t base = s_base;
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
template <int W>
__global__ void __launch_bounds__(kBlock) kb_channel(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                     const ChanDev* __restrict__ chans, const int32_t* __restrict__ en,
                                                     BlockMap bm) {
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
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
    double* __restrict__ Lc = d.Lc[cur];
    for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += nb * kBlock) {
        const Key<W> k = load_key<W>(Lz, Lx, i);
        Lc[i] = Lc[i] * chan_factor<W>(ch, k);
    }
}

// ---------------------------------------------------------------------------
// <k|O|k> (spv_kernels.jl:441-452) with a double-double accumulator so the
// parallel sum is accurate to ~2^-106 of sum |c| before its one rounding.
// Partials per block, reduced per vector in block order (deterministic).
template <int W>
__global__ void __launch_bounds__(kBlock) kb_expect(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                    const uint64_t* __restrict__ kets, BlockMap bm,
                                                    double2* __restrict__ partial) {
    __shared__ double s_h[kBlock], s_l[kBlock];
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
    const SegState* s = S + b;
    const int cur = s->cur ^ s->flip;
    const int64_t n = s->n[cur];
    const SegDesc& d = D[b];
    const uint64_t* __restrict__ Lz = d.Lz[cur];
    const uint64_t* __restrict__ Lx = d.Lx[cur];
    const double* __restrict__ Lc = d.Lc[cur];
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
            const double v = (par & 1) ? -Lc[i] : Lc[i];
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
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s_h[0], s_l[0]);
}

__global__ void __launch_bounds__(kBlock) kb_expect_final(const double2* __restrict__ partial, BlockMap bm,
                                                          double* __restrict__ out) {
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
    if (threadIdx.x == 0) out[b] = s_h[0] + s_l[0];
}

// check_spv key order (type_SparsePauliVector.jl:273-276): bad[b] = 1 if any
// adjacent pair of vector b is not strictly increasing.
template <int W>
__global__ void __launch_bounds__(kBlock) kb_check(const SegDesc* __restrict__ D, const SegState* __restrict__ S,
                                                   BlockMap bm, int32_t* __restrict__ bad) {
    int b;
    int64_t lb, nb;
    map_block(bm, b, lb, nb);
    const SegState* s = S + b;
    const int cur = s->cur ^ s->flip;
    const int64_t n = s->n[cur];
    const SegDesc& d = D[b];
    for (int64_t i = lb * kBlock + threadIdx.x + 1; i < n; i += nb * kBlock) {
        if (!key_lt<W>(load_key<W>(d.Lz[cur], d.Lx[cur], i - 1), load_key<W>(d.Lz[cur], d.Lx[cur], i))) bad[b] = 1;
    }
}

}  // namespace pp
