// Resident layer kernel: one launch runs a whole rotation sequence (one
// layer call of evolve!, spv_evolve.jl:312-387 at window = 1) for every
// vector of a batch, with the vector's terms resident in LDS.
//
// Each vector is served by a cluster of K <= 32 workgroups of 512 threads on
// one XCD; workgroup k of the cluster holds a contiguous, sorted chunk of the
// live set in its LDS.  A rotation is three phases separated by two cluster
// barriers:
//   1. local: anticommutation flags (and the local filter) for the chunk,
//      per-level run heads of the child-rank computation, the chunk's
//      summary record (size, A-count, first/last keys, per-level head
//      summaries) published to global memory;
//   2. after barrier 1: every member reads all records, resolves the trie
//      segments that cross chunk edges, and writes each child
//      (G xor P, ((k-2) sinθ) c) at its global rank in key order (the same
//      rank formula as kb_emit, with chunks in place of tiles) into the
//      vector's child buffer, counting children per owning chunk (the chunk
//      whose key range holds the child);
//   3. after barrier 2: every member sums the whole per-owner table, so all
//      of them know whether any chunk would outgrow its LDS before one of
//      them merges; then each loads the children of its key range and merges
//      them into its chunk in place (live term first on ties, cos-scaling of
//      anticommuting live terms, the strict filter on every output,
//      spv_kernels.jl:326-386): every thread takes its merge-path share into
//      registers, and writes it back after the barrier that ends the reads.
// The reference's capacity bookkeeping (precheck, early merges, append and
// live lengths, WindowCounters) is replayed from the records exactly as the
// per-step kernels do it.
//
// Coherence.  All members of a cluster run on one XCD (roles are taken from
// the XCD id register, not assumed from the dispatch order), so they share
// one L2: a member's stores are complete in L2 after `s_waitcnt vmcnt(0)`,
// which precedes its barrier arrival, and other members read that data with
// agent-scope atomic loads, which bypass their CU's L1.  Barrier counters use
// agent-scope atomics.  No cache is written back or invalidated.
//
// Failure handling.  Every barrier wait is bounded; a timeout sets the abort
// bit of the vector's counter and the member exits.  Results are staged in
// the vector's control block and applied by kr_commit only when every member
// finished without an abort, so an aborted launch leaves the vector exactly as
// it was (the input buffer is never written) and the host replays the call
// with the per-step kernels.  A chunk that would outgrow its LDS (or a
// child/live buffer that is too small) stops the vector before that rotation,
// with its state unchanged; the host then finishes the call from that
// rotation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_common.hpp"
#include "pp_device.hpp"

namespace pp {

constexpr int kXcdCount = 8;  // XCDs of an MI355X
// 512 threads: two waves per SIMD, so each lane may hold 256 registers (the
// merge keeps E terms per thread in registers; spills would go to L2)
constexpr int kResThreads = 512;
constexpr int kResWaves = kResThreads / kWave;  // 8
constexpr int kResMaxK = 32;                    // workgroups per vector (one XCD)
constexpr int kResLevels = 4;                   // mask levels (more: the per-step kernels)
constexpr int kResWords = 64;                   // chunk positions in 64-term words
// Chunk capacity in terms: one LDS buffer of 16 W + sizeof(CT) bytes per term
// holding the chunk with its incoming children (the merge writes it in
// place), and the positions each thread covers (capacity <= E * kResThreads,
// and at most 64 positions per flag word).  ComplexF64 coefficients take 16
// bytes, so at W = 2 the chunk holds 3072 terms (6 per thread).
template <int W, typename CT = double>
struct ResGeom {
    static constexpr bool kCplx = sizeof(CT) == 16;
    static constexpr int cap = W == 1 ? 4096 : (kCplx ? 3072 : 3712);
    static constexpr int E = (cap + kResThreads - 1) / kResThreads;
};
template <int W, typename CT>
constexpr bool res_geom_ok() {
    return ResGeom<W, CT>::cap <= ResGeom<W, CT>::E * kResThreads && ResGeom<W, CT>::cap <= 64 * kResWords;
}
static_assert(res_geom_ok<1, double>() && res_geom_ok<2, double>() && res_geom_ok<1, pp_cplx>() &&
                  res_geom_ok<2, pp_cplx>(),
              "chunk geometry");
constexpr uint32_t kResAbort = 1u << 30;  // barrier counter bit: a member gave up
constexpr uint32_t kResSpin = 1u << 21;   // polls before giving up (~seconds)
constexpr int kResNone = 1 << 20;         // hdb above the vector's first term
constexpr uint32_t kResIdle = ~0u;        // role of a workgroup with no vector

// Chunk record (one per member and rotation parity).  lv[li] packs the
// internal run heads (positions >= 1) of level li:
//   bits 0-12 A before the first head, 13-25 A from the last head,
//   26-38 A between the first two heads, 39-51 A between the last two,
//   52-53 heads (0, 1, 2 = two or more), 54 first head's hdb == level,
//   55 last head's hdb == level.
struct alignas(16) ResRec {
    int32_t n, a, pad0, pad1;
    uint64_t fz[2], fx[2], lz[2], lx[2];
    uint64_t lv[kResLevels];
};
static_assert(sizeof(ResRec) == 112, "record layout");

// Per-vector control block.  The staged results are applied by kr_commit.
struct ResCtl {
    uint32_t bar;        // barrier arrivals (+ kResAbort)
    int32_t done;        // member 0 staged its results
    int32_t wrote;       // the output buffer holds the new live set
    int32_t stall;       // SegState::stalled code (0, 1, 2, 3 = a chunk would outgrow its LDS)
    int32_t err;
    int32_t pad;
    int64_t n_final, stall_step, need;
    int64_t rotations, swept, created, merge_in, merge_out, early, live_cap_v, app_cap_v;
    int64_t rebalances;  // chunk rebalances of the launch
    ResRec rec[2][kResMaxK];
    int32_t hpre[2][kResMaxK][kResMaxK + 1];  // per member: exclusive prefix of its children per owner
};

struct ResArgs {
    const uint32_t* roles;  // [kXcdCount * kx]: b | k << 20 | K << 26, or kResIdle
    uint32_t* tickets;      // [kXcdCount] workgroups seen per XCD
    int kx;                 // roles per XCD
    int B;
    const SegDesc* D;
    SegState* S;
    const StepDev* steps;   // step s of vector b: steps[(s - first) * B + b]
    int64_t first, nsteps;
    ResCtl* ctl;            // [B]
    int cap;                // chunk capacity used (<= ResGeom::cap; smaller in tests)
};

// The step of a rotation, prefetched into LDS during the previous merge.
struct ResStep {
    uint64_t gz[2], gx[2];
    double cosv, sinv;
    int32_t ng, nlev;
    uint32_t lev4;  // the first four mask levels
    int32_t precheck;
    int64_t s;      // step index (>= nsteps: past the last active step)
};

__device__ __forceinline__ uint64_t res_low(int j) { return j ? (~0ull >> (64 - j)) : 0ull; }       // bits < j
__device__ __forceinline__ uint64_t res_le(int j) { return j >= 63 ? ~0ull : ((1ull << (j + 1)) - 1); }  // bits <= j

// Wave64 scans with DPP row shifts and row broadcasts (inactive source lanes
// contribute the identity 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_src(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
__device__ __forceinline__ int wave_incl(int v) {  // inclusive prefix sum
    v += dpp_src<0x111, 0xf>(v);
    v += dpp_src<0x112, 0xf>(v);
    v += dpp_src<0x114, 0xf>(v);
    v += dpp_src<0x118, 0xf>(v);
    v += dpp_src<0x142, 0xa>(v);
    v += dpp_src<0x143, 0xc>(v);
    return v;
}
__device__ __forceinline__ int wave_incl_max(int v) {  // inclusive prefix maximum of values >= 0
    int u;
    u = dpp_src<0x111, 0xf>(v); v = u > v ? u : v;
    u = dpp_src<0x112, 0xf>(v); v = u > v ? u : v;
    u = dpp_src<0x114, 0xf>(v); v = u > v ? u : v;
    u = dpp_src<0x118, 0xf>(v); v = u > v ? u : v;
    u = dpp_src<0x142, 0xa>(v); v = u > v ? u : v;
    u = dpp_src<0x143, 0xc>(v); v = u > v ? u : v;
    return v;
}
__device__ __forceinline__ int wave_shr1(int v) { return dpp_src<0x138, 0xf>(v); }  // lane l gets lane l - 1 (0 at lane 0)
__device__ __forceinline__ int lane_of(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ int wave_total(int v) { return lane_of(wave_incl(v), kWave - 1); }
__device__ __forceinline__ int wave_max(int v) { return lane_of(wave_incl_max(v), kWave - 1); }  // values >= 0

// Workgroup-uniform values read from LDS, moved to scalar registers (the
// vector registers go to the per-thread terms).  Uniform control flow only.
__device__ __forceinline__ int res_uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t res_uni(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t res_uni(int64_t v) { return (int64_t)res_uni((uint64_t)v); }
__device__ __forceinline__ double res_uni(double v) {
    return __longlong_as_double((long long)res_uni((uint64_t)__double_as_longlong(v)));
}
__device__ __forceinline__ FilterReg res_uni(const FilterReg& f) {
    FilterReg r;
    r.wmax = res_uni(f.wmax);
    r.xwmax = res_uni(f.xwmax);
    r.mwmax = res_uni(f.mwmax);
    r.thresh = res_uni(f.thresh);
    r.dthresh_w = res_uni(f.dthresh_w);
    r.dthresh_xw = res_uni(f.dthresh_xw);
    r.exptab_w = reinterpret_cast<const double*>(res_uni((uint64_t)f.exptab_w));
    r.exptab_xw = reinterpret_cast<const double*>(res_uni((uint64_t)f.exptab_xw));
    return r;
}

// Data other members wrote during the launch (records, per-owner counts,
// children) is read with agent-scope relaxed atomic loads: they miss in the
// CU's L1, which the launch never invalidates, and read the XCD's L2, where a
// member's stores are complete once it has arrived at the barrier.
__device__ __forceinline__ uint64_t ld_l2(const uint64_t* p) {
    return __hip_atomic_load((PP_GLOBAL uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld_l2(const int32_t* p) {
    return __hip_atomic_load((PP_GLOBAL int32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_l2(const double* p) {
    return __longlong_as_double((long long)ld_l2(reinterpret_cast<const uint64_t*>(p)));
}
__device__ __forceinline__ pp_cplx ld_l2(const pp_cplx* p) {
    const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
    pp_cplx r;
    r.x = __longlong_as_double((long long)ld_l2(q));
    r.y = __longlong_as_double((long long)ld_l2(q + 1));
    return r;
}
template <int W>
__device__ __forceinline__ Key<W> load_key_l2(const uint64_t* Z, const uint64_t* X, int64_t i) {
    Key<W> k;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        k.z[w] = ld_l2(Z + i * W + w);
        k.x[w] = ld_l2(X + i * W + w);
    }
    return k;
}

// Cluster barrier number `target / K`.  false: the cluster gave up.
__device__ __forceinline__ bool res_barrier(uint32_t* bar, uint32_t target, int* s_ok) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's stores are in L2
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        for (uint32_t spins = 0;; ++spins) {
            const uint32_t v = __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v & kResAbort) { ok = 0; break; }
            if (v >= target) break;
            if (spins > kResSpin) {
                __hip_atomic_fetch_or(bar, kResAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// One wave: packed internal head summary of one level (lane w = word w) and
// the per-word tables: last head in words < w (-1: none), first head in words
// > w (0x7fff: none).
__device__ uint64_t res_level_summary(const uint64_t* hm_w, const uint64_t* em_w, const uint64_t* am_w, int16_t* lb,
                                      int16_t* fa) {
    const int lane = threadIdx.x % kWave;
    const uint64_t hm = hm_w[lane], em = em_w[lane], am = am_w[lane];
    const int cnt = popc64(am);
    const int incl = wave_incl(cnt);
    const int pre = incl - cnt;
    const int total = lane_of(incl, kWave - 1);
    auto apos = [&](int p) {  // A in [0, p), wave-uniform p
        const int w = p >> 6;
        const uint64_t aw = ((uint64_t)(uint32_t)lane_of((int)(uint32_t)(am >> 32), w) << 32) |
                            (uint32_t)lane_of((int)(uint32_t)am, w);
        return lane_of(pre, w) + popc64(aw & res_low(p & 63));
    };
    auto word_of = [&](uint64_t v, int w) {
        return ((uint64_t)(uint32_t)lane_of((int)(uint32_t)(v >> 32), w) << 32) | (uint32_t)lane_of((int)(uint32_t)v, w);
    };
    const uint64_t hw = __ballot(hm != 0ull);
    int f1 = -1, f2 = -1, l1 = -1, l2 = -1;
    if (hw) {
        const int wf = __ffsll((unsigned long long)hw) - 1;
        const uint64_t m0 = word_of(hm, wf);
        f1 = wf * 64 + __ffsll((unsigned long long)m0) - 1;
        const uint64_t m1 = m0 & (m0 - 1);
        if (m1) {
            f2 = wf * 64 + __ffsll((unsigned long long)m1) - 1;
        } else {
            const uint64_t hw2 = hw & (hw - 1);
            if (hw2) {
                const int w2 = __ffsll((unsigned long long)hw2) - 1;
                f2 = w2 * 64 + __ffsll((unsigned long long)word_of(hm, w2)) - 1;
            }
        }
        const int wl = 63 - __clzll((long long)hw);
        const uint64_t n0 = word_of(hm, wl);
        const int q = 63 - __clzll((long long)n0);
        l1 = wl * 64 + q;
        const uint64_t n1 = n0 & ~(1ull << q);
        if (n1) {
            l2 = wl * 64 + 63 - __clzll((long long)n1);
        } else {
            const uint64_t hw3 = hw & ~(1ull << wl);
            if (hw3) {
                const int w3 = 63 - __clzll((long long)hw3);
                l2 = w3 * 64 + 63 - __clzll((long long)word_of(hm, w3));
            }
        }
    }
    const int af = f1 < 0 ? total : apos(f1);
    const int al = l1 < 0 ? 0 : total - apos(l1);
    const int a12 = f2 < 0 ? 0 : apos(f2) - af;
    const int al2 = l2 < 0 ? 0 : apos(l1) - apos(l2);
    const uint64_t heads = f1 < 0 ? 0u : (f2 < 0 ? 1u : 2u);
    const uint64_t eqf = f1 >= 0 ? ((word_of(em, f1 >> 6) >> (f1 & 63)) & 1u) : 0u;
    const uint64_t eql = l1 >= 0 ? ((word_of(em, l1 >> 6) >> (l1 & 63)) & 1u) : 0u;
    // last head in words < w: exclusive prefix maximum (positions + 1)
    const int lastw = hm ? lane * 64 + 64 - __clzll((long long)hm) : 0;
    lb[lane] = (int16_t)(wave_shr1(wave_incl_max(lastw)) - 1);
    // first head in words > w: on reversed lanes, a prefix maximum of 0x7fff - position
    const int firstw = hm ? lane * 64 + __ffsll((unsigned long long)hm) - 1 : 0x7fff;
    const int rev = __shfl(0x7fff - firstw, kWave - 1 - lane, kWave);
    const int sfx = wave_shr1(wave_incl_max(rev));
    fa[lane] = (int16_t)(0x7fff - __shfl(sfx, kWave - 1 - lane, kWave));
    return (uint64_t)af | ((uint64_t)al << 13) | ((uint64_t)a12 << 26) | ((uint64_t)al2 << 39) | (heads << 52) |
           (eqf << 54) | (eql << 55);
}

// Cross-edge values of chunk j at one level (out layout): ts_head, ts_eq,
// pre_same, pre_sib, te_head, te_eq, post_same, post_sib.  One wave; lane i
// reads record i.
//   pre_same  A of the run that contains the chunk's previous term, from its
//             head to the chunk start
//   pre_sib   A of the run before that run's head when that head's hdb equals
//             the level (else 0)
//   post_same te a head: A of the run starting at te; else A of the run
//             containing te from te to its end
//   post_sib  te not a head: A of the run after the run containing te when its
//             head's hdb equals the level (else 0)
template <int W>
__device__ void res_edges(const ResRec* rec, int K, int j, int li, int lvl, int32_t* out) {
    const int lane = threadIdx.x % kWave;
    const bool valid = lane < K;
    const int n = valid ? rec[lane].n : 0;
    const bool ne = n > 0;
    const int Ai = valid ? rec[lane].a : 0;
    const uint64_t nemask = __ballot(ne);
    const uint64_t below = nemask & res_low(lane);
    int h0 = -1;
    if (ne) {
        if (!below) {
            h0 = kResNone;
        } else {
            const int pv = 63 - __clzll((long long)below);
            Key<W> a, b;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                a.z[w] = rec[pv].lz[w];
                a.x[w] = rec[pv].lx[w];
                b.z[w] = rec[lane].fz[w];
                b.x[w] = rec[lane].fx[w];
            }
            h0 = hdb<W>(a, b);
        }
    }
    const uint64_t lv = valid ? rec[lane].lv[li] : 0ull;
    int af = (int)(lv & 8191u), al = (int)((lv >> 13) & 8191u), a12 = (int)((lv >> 26) & 8191u),
        al2 = (int)((lv >> 39) & 8191u), nint = (int)((lv >> 52) & 3u);
    int eqf = (int)((lv >> 54) & 1u), eql = (int)((lv >> 55) & 1u);
    int heads = nint;
    if (ne && h0 >= lvl) {  // position 0 is a head too
        heads = nint + 1 > 2 ? 2 : nint + 1;
        const int af_int = af;
        af = 0;
        eqf = h0 == lvl;
        a12 = af_int;
        if (nint >= 1) {
            al2 = nint >= 2 ? al2 : af_int;
        } else {
            al = Ai;
            eql = h0 == lvl;
            al2 = 0;
        }
    }
    const uint64_t hmask = __ballot(ne && heads > 0);
    auto sumA = [&](int lo, int hi) {  // A over chunks lo < i < hi
        return wave_total((lane > lo && lane < hi) ? Ai : 0);
    };
    int ts_head, ts_eq, pre_same = 0, pre_sib = 0, te_head, te_eq, post_same = 0, post_sib = 0;
    const uint64_t prevm = nemask & res_low(j);
    if (!prevm) {
        ts_head = 1;
        ts_eq = 0;
    } else {
        const int h0j = lane_of(h0, j);
        ts_head = h0j >= lvl;
        ts_eq = h0j == lvl;
        const uint64_t hb = hmask & res_low(j);
        const int i1 = hb ? 63 - __clzll((long long)hb) : 0;
        pre_same = lane_of(al, i1) + sumA(i1, j);
        if (lane_of(eql, i1)) {
            if (lane_of(heads, i1) >= 2) {
                pre_sib = lane_of(al2, i1);
            } else {
                const uint64_t hb2 = hmask & res_low(i1);
                const int i2 = hb2 ? 63 - __clzll((long long)hb2) : -1;
                const int last2 = lane_of(al, i2 < 0 ? 0 : i2);
                pre_sib = lane_of(af, i1) + sumA(i2, i1) + (i2 >= 0 ? last2 : 0);
            }
        }
    }
    const uint64_t nxm = nemask & ~res_le(j);
    if (!nxm) {
        te_head = 1;
        te_eq = 0;
    } else {
        const int nx = __ffsll((unsigned long long)nxm) - 1;
        const int h0n = lane_of(h0, nx);
        te_head = h0n >= lvl;
        te_eq = h0n == lvl;
        auto run_from = [&](int k) {  // A of the run starting at chunk k's first head
            const uint64_t ha = hmask & ~res_le(k);
            const int k2 = ha ? __ffsll((unsigned long long)ha) - 1 : kWave;
            const int first2 = lane_of(af, k2 < kWave ? k2 : 0);
            const int r2 = lane_of(al, k) + sumA(k, k2) + (k2 < kWave ? first2 : 0);
            return lane_of(heads, k) >= 2 ? lane_of(a12, k) : r2;
        };
        if (te_head) {
            post_same = run_from(nx);
        } else {
            const uint64_t ha = hmask & ~res_le(j);
            const int k1 = ha ? __ffsll((unsigned long long)ha) - 1 : kWave;
            const int kk = k1 < kWave ? k1 : 0;
            const int first1 = lane_of(af, kk);
            post_same = sumA(j, k1) + (k1 < kWave ? first1 : 0);
            const int eq1 = lane_of(eqf, kk);
            const int rf = run_from(kk);
            post_sib = (k1 < kWave && eq1) ? rf : 0;
        }
    }
    if (lane == 0) {
        out[0] = ts_head;
        out[1] = ts_eq;
        out[2] = pre_same;
        out[3] = pre_sib;
        out[4] = te_head;
        out[5] = te_eq;
        out[6] = post_same;
        out[7] = post_sib;
    }
}

// The step of vector b at index s as loaded from the step table (loads issued
// early, stored to LDS after other work: res_step_commit).
struct ResStepLoad {
    const StepDev* sd;
    int64_t s;
    int32_t active, precheck, ng, nlev;
    uint64_t gz0, gz1, gx0, gx1;
    double cosv, sinv;
    uint32_t lev4;
};
__device__ __forceinline__ ResStepLoad res_step_load(const ResArgs& A, int b, int64_t s) {
    ResStepLoad t;
    t.s = s;
    t.sd = nullptr;
    t.active = 0;
    if (s < A.nsteps) {
        const StepDev* sd = A.steps + (s - A.first) * A.B + b;
        t.sd = sd;
        t.active = sd->active;
        t.precheck = sd->precheck;
        t.gz0 = sd->g.gz[0];
        t.gz1 = sd->g.gz[1];
        t.gx0 = sd->g.gx[0];
        t.gx1 = sd->g.gx[1];
        t.cosv = sd->g.cosv;
        t.sinv = sd->g.sinv;
        t.ng = sd->g.ng;
        t.nlev = sd->g.nlev;
        t.lev4 = *reinterpret_cast<const uint32_t*>(sd->g.lev);
    }
    return t;
}
// The first active step >= t.s into *out (further steps loaded here when t.s
// is inactive).
__device__ __forceinline__ void res_step_commit(const ResArgs& A, int b, ResStepLoad t, ResStep* out) {
    while (t.s < A.nsteps && !t.active) t = res_step_load(A, b, t.s + 1);
    ResStep r;
    r.s = t.s;
    if (t.s < A.nsteps) {
        r.gz[0] = t.gz0;
        r.gz[1] = t.gz1;
        r.gx[0] = t.gx0;
        r.gx[1] = t.gx1;
        r.cosv = t.cosv;
        r.sinv = t.sinv;
        r.ng = t.ng;
        r.nlev = t.nlev;
        r.lev4 = t.lev4;
        r.precheck = t.precheck;
    } else {
        r.gz[0] = r.gz[1] = r.gx[0] = r.gx[1] = 0;
        r.cosv = r.sinv = 0.0;
        r.ng = r.nlev = 0;
        r.lev4 = 0;
        r.precheck = 0;
    }
    *out = r;
}

// Chunk edges moved to the first position of the highest differing bit
// within kAlign terms: for the edge at e of a sorted sequence of n keys, the
// smallest i in [e, e + kAlign) ∩ [1, n) that maximises hdb(key[i-1], key[i])
// (e itself when e is 0 or n).  The same function of the same position in
// both members that share an edge, and monotone in it.  Waves [0, kAlignWaves)
// take e_lo, the next kAlignWaves waves e_hi; the whole workgroup calls it,
// and the results are in out[0], out[1] after its closing barrier.
constexpr int kAlignWaves = 1, kAlign = kAlignWaves * kWave;
static_assert(2 * kAlignWaves <= kResWaves, "one wave per 64 positions of each edge");
template <int W, bool L2>
__device__ __forceinline__ void align_edges(const uint64_t* Z, const uint64_t* X, int64_t n, int64_t e_lo,
                                            int64_t e_hi, int32_t (*s_v)[kAlignWaves], int64_t* out) {
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    if (wave < 2 * kAlignWaves) {
        const int side = wave / kAlignWaves, part = wave % kAlignWaves;
        const int64_t e0 = side ? e_hi : e_lo;
        int v = 0;  // (level + 1) << 12 | (4095 - offset); 0: none
        if (e0 > 0 && e0 < n) {
            const int64_t i = e0 + part * kWave + lane;
            int hd = -1;
            if (i < n) {
                const Key<W> a = L2 ? load_key_l2<W>(Z, X, i - 1) : load_key<W>(Z, X, i - 1);
                const Key<W> c = L2 ? load_key_l2<W>(Z, X, i) : load_key<W>(Z, X, i);
                hd = hdb<W>(a, c);
            }
            const int mx = wave_max(hd + 1);  // every lane active
            const uint64_t bm = __ballot(hd >= 0 && hd + 1 == mx);
            if (bm) v = (mx << 12) | (4095 - (part * kWave + __ffsll((unsigned long long)bm) - 1));
        }
        if (lane == 0) s_v[side][part] = v;
    }
    __syncthreads();
    if (tid < 2) {
        const int64_t e0 = tid ? e_hi : e_lo;
        int best = 0;
        for (int q = 0; q < kAlignWaves; ++q) best = s_v[tid][q] > best ? s_v[tid][q] : best;
        out[tid] = best ? e0 + (4095 - (best & 4095)) : e0;
    }
    __syncthreads();
}

// Bookkeeping slots (thread 0 of each member; member 0 stages them).
constexpr int kBkRot = 0, kBkSwept = 1, kBkCreated = 2, kBkMin = 3, kBkMout = 4, kBkEarly = 5, kBkLive = 6,
              kBkApp = 7, kBkPn = 8, kBkPm = 9, kBk = 10;

template <int W, typename CT>
__global__ void __launch_bounds__(kResThreads) kr_layer(ResArgs A) {
    constexpr int Cap = ResGeom<W, CT>::cap;
    constexpr int E = ResGeom<W, CT>::E;
    __shared__ Key<W> s_k[Cap];
    __shared__ CT s_c[Cap];
    __shared__ uint64_t s_am[kResWords];
    __shared__ int32_t s_apre[kResWords];
    __shared__ uint64_t s_hm[kResLevels][kResWords];
    __shared__ uint64_t s_em[kResLevels][kResWords];
    __shared__ int16_t s_lb[kResLevels][kResWords];
    __shared__ int16_t s_fa[kResLevels][kResWords];
    __shared__ ResRec s_rec[kResMaxK];
    __shared__ int32_t s_mk[kResMaxK + 1];  // children per member; [K]: this member's first rank
    __shared__ int32_t s_rbw[kResMaxK], s_rbn[kResMaxK], s_rbp[kResMaxK + 1];  // rebalance prefixes, bounds
    __shared__ int32_t s_edge[kResLevels][8];
    __shared__ int32_t s_hist[kResMaxK + 1];
    __shared__ Key<W> s_ownk[kResMaxK];
    __shared__ int32_t s_owni[kResMaxK];
    __shared__ int32_t s_wsum[kResWaves];
    __shared__ int32_t s_int[8];   // role, slot, error bits, barrier result, owners, overflow
    __shared__ int32_t s_tot[3];   // n_tot, m_tot, A before this chunk
    __shared__ int64_t s_al[2];    // a chunk's edges, moved by align_edges
    __shared__ int32_t s_alv[2][kAlignWaves];
    __shared__ int64_t s_bk[kBk];
    __shared__ FilterReg s_fs, s_fl;
    __shared__ ResStep s_step[2];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;

    // role from the XCD this workgroup runs on
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    xcc &= 15u;
    if (tid == 0) {
        uint32_t role = kResIdle, slot = 0;
        if (xcc < (uint32_t)kXcdCount) {
            const uint32_t t = __hip_atomic_fetch_add(A.tickets + xcc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t < (uint32_t)A.kx) {
                slot = xcc * (uint32_t)A.kx + t;
                role = A.roles[slot];
            }
        }
        s_int[0] = (int32_t)role;
        s_int[1] = (int32_t)slot;
        s_int[2] = 0;
    }
    __syncthreads();
    const uint32_t role = (uint32_t)s_int[0];
    if (role == kResIdle) return;
    const int b = (int)(role & 0xFFFFFu), k = (int)((role >> 20) & 63u), K = (int)((role >> 26) & 63u);
    const SegDesc* __restrict__ dp = A.D + b;
    SegState* st = A.S + b;
    ResCtl* ctl = A.ctl + b;
    const int cap = A.cap;
    const int stalled0 = st->stalled;
    const int cur = st->cur ^ st->flip;
    const int64_t n0 = st->n[cur];
    const int local_active = dp->local_active;
    if (tid == 0) {
        for (int q = 0; q < kBk; ++q) s_bk[q] = 0;
        s_bk[kBkLive] = st->live_cap_v;
        s_bk[kBkApp] = st->app_cap_v;
        s_fs.load(dp->fs);
        if (local_active) s_fl.load(dp->fl);
        res_step_commit(A, b, res_step_load(A, b, A.first), &s_step[0]);
    }
    int32_t* s_err = &s_int[2];
    if (tid < kResWords) {  // every flag word starts at zero
        s_am[tid] = 0ull;
        for (int li = 0; li < kResLevels; ++li) s_hm[li][tid] = s_em[li][tid] = 0ull;
    }

    if (stalled0) {  // frozen by an earlier call: nothing to do
        if (k == 0 && tid == 0) {
            ctl->wrote = 0;
            ctl->stall = 0;
            ctl->rotations = ctl->swept = ctl->created = ctl->merge_in = ctl->merge_out = ctl->early = 0;
            ctl->rebalances = 0;
            ctl->live_cap_v = st->live_cap_v;
            ctl->app_cap_v = st->app_cap_v;
            ctl->n_final = n0;
            __hip_atomic_store(&ctl->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }

    // initial chunk: [k n0 / K, (k + 1) n0 / K)
    int nj = 0;  // chunk size
    int stall = 0;
    int64_t stall_step = A.first, need = 0;
    {
        int64_t lo = (n0 * k) / K, hi = (n0 * (k + 1)) / K;
        if ((n0 + K - 1) / K > cap) {  // the largest chunk (the same test in every member)
            stall = 3;
        } else {
            const uint64_t* Lz = cur ? dp->Lz[1] : dp->Lz[0];
            const uint64_t* Lx = cur ? dp->Lx[1] : dp->Lx[0];
            const CT* Lc = reinterpret_cast<const CT*>(cur ? dp->Lc[1] : dp->Lc[0]);
            // Edges at the highest differing bit within reach, when the
            // larger chunks still fit: a rotation whose mask lies below every
            // edge's level then keeps every child in its parent's chunk (see
            // the local rotations below).
            if (K > 1 && (n0 + K - 1) / K + kAlign <= cap) {
                align_edges<W, false>(Lz, Lx, n0, lo, hi, s_alv, s_al);
                lo = s_al[0];
                hi = s_al[1];
            }
            nj = (int)(hi - lo);
            for (int p = tid; p < nj; p += kResThreads) {
                s_k[p] = load_key<W>(Lz, Lx, lo + p);
                s_c[p] = ldg(Lc, lo + p);
            }
        }
    }
    bool pend = false;
    bool aborted = false;
    int64_t rb_step = -1;  // step of the last rebalance
    int nrb = 0;           // rebalances
    int e_dirty = 0;  // flag words at e * kResWaves + wave, e < e_dirty, may be nonzero
    uint32_t nbar = 0;
    int par = 0, cs = 0;  // record parity, step slot

    while (stall == 0) {
        __syncthreads();  // the chunk and the prefetched step are complete
        const ResStep& S = s_step[cs];
        const int64_t s = res_uni(S.s);
        const bool last = s >= A.nsteps;
        GenReg<W> G;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            G.gz[w] = res_uni(S.gz[w]);
            G.gx[w] = res_uni(S.gx[w]);
        }
        G.cosv = res_uni(S.cosv);
        G.sinv = res_uni(S.sinv);
        G.ng = res_uni(S.ng);
        const uint32_t lev4 = (uint32_t)res_uni((int)S.lev4);
        const int precheck = res_uni(S.precheck);
        int nlev = last ? 0 : res_uni(S.nlev);
        if (nlev > kResLevels) {  // the host never sends these
            if (tid == 0) atomicOr(s_err, 8);
            nlev = kResLevels;
        }
        Key<W>* sk = s_k;
        CT* sc = s_c;
        // ---- phase 1: flags, run heads, summaries, record
        const int e_need = (nj + kResThreads - 1) / kResThreads;  // position rows in use
        if (!last) {
            if (e_dirty > e_need) {  // clear the rows the chunk no longer reaches
                const int w0 = e_need * kResWaves, nw = (e_dirty - e_need) * kResWaves;
                for (int q = tid; q < nw * (1 + 2 * kResLevels); q += kResThreads) {
                    const int arr = q / nw, w = w0 + q % nw;
                    if (arr == 0) s_am[w] = 0ull;
                    else if (arr <= kResLevels) s_hm[arr - 1][w] = 0ull;
                    else s_em[arr - 1 - kResLevels][w] = 0ull;
                }
            }
            e_dirty = e_need;
#pragma unroll 1
            for (int e = 0; e < e_need; ++e) {
                const int p = e * kResThreads + tid;
                const int wi = e * kResWaves + wave;
                const bool valid = p < nj;
                bool flag = false;
                int hd = -1;
                if (valid) {
                    const Key<W> Kp = sk[p];
                    flag = G.anti(Kp);
                    if (flag && local_active) {
                        const FilterReg fl = s_fl;  // (divergent here: stays in vector registers)
                        Key<W> ch;
                        const CT cn = G.child(Kp, sc[p], ch);
                        flag = !should_drop<W>(fl, ch, coef_abs(cn));
                    }
                    if (p > 0) hd = hdb<W>(sk[p - 1], Kp);
                }
                const uint64_t am = __ballot(flag);
                if (lane == 0) s_am[wi] = am;
#pragma unroll
                for (int li = 0; li < kResLevels; ++li) {
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
            __syncthreads();
        }
        ResRec* grec = &ctl->rec[par][k];
        ResRec* myrec = &s_rec[0];
        if (wave == 0) {
            int a = 0;
            if (!last) {
                const int cnt = popc64(s_am[lane]);
                const int incl = wave_incl(cnt);
                s_apre[lane] = incl - cnt;
                a = lane_of(incl, kWave - 1);
            }
            if (lane == 0) {
                if (K == 1) {
                    myrec->n = nj;
                    myrec->a = a;
                } else {
                    stg(&grec->n, 0, nj);
                    stg(&grec->a, 0, a);
                }
            }
            if (lane < 2 * W && nj > 0) {  // first and last keys
                const int w = lane % W;
                const bool lastk = lane >= W;
                const int q = lastk ? nj - 1 : 0;
                const uint64_t zv = sk[q].z[w], xv = sk[q].x[w];
                if (K == 1) {
                    (lastk ? myrec->lz : myrec->fz)[w] = zv;
                    (lastk ? myrec->lx : myrec->fx)[w] = xv;
                } else {
                    stg(lastk ? grec->lz : grec->fz, w, zv);
                    stg(lastk ? grec->lx : grec->fx, w, xv);
                }
            }
        } else if (wave <= nlev) {
            const int li = wave - 1;
            const uint64_t lv = res_level_summary(s_hm[li], s_em[li], s_am, s_lb[li], s_fa[li]);
            if (lane == 0) {
                if (K == 1) myrec->lv[li] = lv;
                else stg(grec->lv, li, lv);
            }
        }
        // ---- barrier 1, then every record of the cluster in LDS
        if (K > 1) {
            nbar += 1;
            if (!res_barrier(&ctl->bar, nbar * (uint32_t)K, &s_int[3])) {
                aborted = true;
                break;
            }
            if (tid < K * 7) {  // 7 x 16 bytes per record
                const int r = tid / 7, q = tid % 7;
                const uint64_t* src = reinterpret_cast<const uint64_t*>(&ctl->rec[par][r]) + 2 * q;
                uint64_t* dst = reinterpret_cast<uint64_t*>(&s_rec[r]) + 2 * q;
                const uint64_t v0 = ld_l2(src), v1 = ld_l2(src + 1);
                dst[0] = v0;
                dst[1] = v1;
            }
        }
        __syncthreads();
        // cluster totals (wave 0), cross-edge values (one wave per level) and
        // the owner table (last wave)
        if (wave == 0) {
            const bool in = lane < K;
            const int ni = in ? s_rec[lane].n : 0, ai = in ? s_rec[lane].a : 0;
            const int nt = wave_total(ni), mt = wave_total(ai), ab = wave_total(lane < k ? ai : 0);
            // Local rotation: every pair of neighbouring non-empty chunks
            // differs above the mask's top bit, so a child (which differs from
            // its parent only at the mask's bits) stays between the chunk's
            // neighbours, in its parent's chunk, at rank - (A before the
            // chunk) among the chunk's children; and every chunk keeps room for
            // its own children.  The same records give every member the same
            // answer.
            const int topm = nlev > 0 ? (int)((lev4 >> (8 * (nlev - 1))) & 255u) : -1;
            const bool ne = in && ni > 0;
            const uint64_t nem = __ballot(ne);
            bool ok = true;
            if (ne) {
                ok = ni + ai <= cap;
                const uint64_t below = nem & res_low(lane);
                if (below) {
                    const int pv = 63 - __clzll((long long)below);
                    Key<W> a, c;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        a.z[w] = s_rec[pv].lz[w];
                        a.x[w] = s_rec[pv].lx[w];
                        c.z[w] = s_rec[lane].fz[w];
                        c.x[w] = s_rec[lane].fx[w];
                    }
                    ok = ok && hdb<W>(a, c) > topm;
                }
            }
            const bool loc = K > 1 && !last && __ballot(!ok) == 0ull;
            if (lane == 0) {
                s_tot[0] = nt;
                s_tot[1] = mt;
                s_tot[2] = ab;
                s_int[7] = loc ? 1 : 0;
            }
        } else if (!last && wave <= nlev) {
            const int li = wave - 1;
            const int lvl = (int)((lev4 >> (8 * li)) & 255u);
            res_edges<W>(s_rec, K, k, li, lvl, s_edge[li]);
        } else if (!last && wave == kResWaves - 1) {
            const bool ne = lane < K && s_rec[lane < K ? lane : 0].n > 0;
            const uint64_t nem = __ballot(ne);
            if (ne) {
                const int pos = popc64(nem & res_low(lane));
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    s_ownk[pos].z[w] = s_rec[lane].fz[w];
                    s_ownk[pos].x[w] = s_rec[lane].fx[w];
                }
                s_owni[pos] = lane;
            }
            if (lane == 0) s_int[4] = popc64(nem);
            if (lane <= kResMaxK) s_hist[lane] = s_mk[lane] = 0;
        }
        __syncthreads();
        const int64_t n_tot = s_tot[0], m_tot = s_tot[1], abase = s_tot[2];
        if (pend) {
            if (tid == 0) {  // confirm the pending rotation (kb_merge's bookkeeping)
                const int64_t pn = s_bk[kBkPn], pm = s_bk[kBkPm];
                s_bk[kBkRot] += 1;
                s_bk[kBkSwept] += pn;
                s_bk[kBkCreated] += pm;
                s_bk[kBkMin] += pn + pm;
                s_bk[kBkMout] += n_tot;
                const int64_t lc = s_bk[kBkLive];
                if (n_tot > lc) s_bk[kBkLive] = 2 * lc > pn + pm ? 2 * lc : pn + pm;
            }
            pend = false;
        }
        if (last) break;
        if (tid == 0 && precheck) {  // spv_evolve.jl:356-362 (the early merge drops nothing here)
            const int64_t ac = s_bk[kBkApp];
            if (n_tot > ac) {
                s_bk[kBkEarly] += 1;
                s_bk[kBkMin] += n_tot;
                s_bk[kBkMout] += n_tot;
                s_bk[kBkApp] = 2 * ac > n_tot ? 2 * ac : n_tot;
            }
        }
        if (m_tot > dp->capC) {
            stall = 1;
            need = m_tot;
            stall_step = s;
            break;
        }
        if (n_tot + m_tot > dp->cap) {
            stall = 2;
            need = n_tot + m_tot;
            stall_step = s;
            break;
        }
        if (K == 1 && nj + m_tot > cap) {
            stall = 3;
            stall_step = s;
            break;
        }
        Key<W>* skw = s_k;  // children are staged after the chunk
        CT* scw = s_c;
        // ---- phase 2: children at their ranks (and the next rotation's step)
        ResStepLoad nxt;
        if (tid == kResThreads - 1) nxt = res_step_load(A, b, s + 1);
        const bool loc = s_int[7] != 0;
        {
            const int nown = s_int[4];
            const int Aj = s_rec[k].a;
            uint64_t* Cz = dp->Cz;
            uint64_t* Cx = dp->Cx;
            CT* Cc = reinterpret_cast<CT*>(dp->Cc);
#pragma unroll 1
            for (int e = 0; e < e_need; ++e) {
                const int p = e * kResThreads + tid;
                const int wi = e * kResWaves + wave;
                if (p >= nj || !((s_am[wi] >> lane) & 1ull)) continue;
                const Key<W> Kp = sk[p];
                int rank = (int)abase + s_apre[wi] + popc64(s_am[wi] & res_low(lane));
#pragma unroll 1
                for (int li = 0; li < nlev; ++li) {
                    const int lvl = (int)((lev4 >> (8 * li)) & 255u);
                    const int32_t* ed = s_edge[li];
                    const bool ts_head = ed[0] != 0, ts_eq = ed[1] != 0, te_head = ed[4] != 0, te_eq = ed[5] != 0;
                    auto apre = [&](int q) -> int {
                        return q >= nj ? Aj : s_apre[q >> 6] + popc64(s_am[q >> 6] & res_low(q & 63));
                    };
                    auto last_le = [&](int q) -> int {  // last head <= q
                        const uint64_t m = s_hm[li][q >> 6] & res_le(q & 63);
                        int r = m ? ((q & ~63) + 63 - __clzll((long long)m)) : (int)s_lb[li][q >> 6];
                        if (r < 0 && ts_head) r = 0;
                        return r;
                    };
                    auto first_gt = [&](int q) -> int {  // first internal head > q (nj: none)
                        const int j = q & 63;
                        const uint64_t m = j == 63 ? 0ull : (s_hm[li][q >> 6] & (~0ull << (j + 1)));
                        const int r = m ? ((q & ~63) + __ffsll((unsigned long long)m) - 1) : (int)s_fa[li][q >> 6];
                        return r < nj ? r : nj;
                    };
                    auto eq_at = [&](int q) -> bool {
                        return q == 0 ? ts_eq : (((s_em[li][q >> 6] >> (q & 63)) & 1ull) != 0);
                    };
                    int delta = 0;
                    if (key_bit<W>(Kp, lvl)) {
                        const int s0 = last_le(p);
                        if (s0 < 0) {
                            delta = -ed[3];
                        } else if (eq_at(s0)) {
                            const int s2 = s0 > 0 ? last_le(s0 - 1) : -1;
                            delta = s2 >= 0 ? -(apre(s0) - apre(s2)) : -(apre(s0) + ed[2]);
                        }
                    } else {
                        const int e0 = first_gt(p);
                        if (e0 < nj) {
                            if (eq_at(e0)) {
                                const int e2 = first_gt(e0);
                                delta = e2 < nj ? apre(e2) - apre(e0) : (Aj - apre(e0)) + (te_head ? 0 : ed[6]);
                            }
                        } else {
                            delta = te_head ? (te_eq ? ed[6] : 0) : ed[7];
                        }
                    }
                    rank += delta;
                }
                Key<W> ch;
                const CT cn = G.child(Kp, sc[p], ch);
                if (rank < 0 || rank >= m_tot) {  // never expected
                    atomicOr(s_err, 1);
                    continue;
                }
                if (K == 1) {
                    skw[nj + rank] = ch;
                    scw[nj + rank] = cn;
                } else if (loc) {
                    const int r = rank - (int)abase;
                    if (r < 0 || r >= Aj) {  // never expected
                        atomicOr(s_err, 1);
                        continue;
                    }
                    skw[nj + r] = ch;
                    scw[nj + r] = cn;
                } else {
                    store_key<W>(Cz, Cx, rank, ch);
                    stg(Cc, rank, cn);
                    int lo = 0, hi = nown;  // owner: last non-empty chunk whose first key <= ch
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (key_le<W>(s_ownk[mid], ch)) lo = mid;
                        else hi = mid;
                    }
                    // ... or the next one when the child agrees with that
                    // chunk's first key above the level of the edge between
                    // them (the highest bit where the owner's last key and
                    // that first key differ): the child then lies above every
                    // key of the owner, owners stay monotone in the key, and
                    // no edge's level falls, which keeps the rotations whose
                    // masks lie below every edge local
                    if (lo + 1 < nown) {
                        const Key<W>& nf = s_ownk[lo + 1];
                        const ResRec& ro = s_rec[s_owni[lo]];
                        Key<W> la;
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            la.z[w] = ro.lz[w];
                            la.x[w] = ro.lx[w];
                        }
                        if (hdb<W>(ch, nf) < hdb<W>(la, nf)) lo += 1;
                    }
                    atomicAdd(&s_hist[s_owni[lo]], 1);
                }
            }
        }
        // the next rotation's step (its loads were issued at the start of
        // phase 2); the barrier at the top of the loop publishes it
        if (tid == kResThreads - 1) res_step_commit(A, b, nxt, &s_step[cs ^ 1]);
        int mk = (int)m_tot;
        if (K > 1 && loc) {
            mk = s_rec[k].a;  // the chunk's own children, staged in LDS above
        } else if (K > 1) {
            __syncthreads();
            if (wave == 0) {  // this member's children per owner, as an exclusive prefix row
                const int v = lane < K ? s_hist[lane] : 0;
                const int incl = wave_incl(v);
                if (lane <= K) stg(&ctl->hpre[par][k][0], lane, incl - v);
            }
            // ---- barrier 2
            nbar += 1;
            if (!res_barrier(&ctl->bar, nbar * (uint32_t)K, &s_int[3])) {
                aborted = true;
                break;
            }
            // every member's incoming children from the whole per-owner table,
            // so that all members see the same overflow decision before any
            // of them merges
            for (int q = tid; q < K * K; q += kResThreads) {
                const int i = q / K, j = q % K;
                const int v0 = ld_l2(&ctl->hpre[par][i][j]), v1 = ld_l2(&ctl->hpre[par][i][j + 1]);
                atomicAdd(&s_mk[j], v1 - v0);
                if (j == k) atomicAdd(&s_mk[kResMaxK], v0);
            }
            __syncthreads();
            if (wave == 0) {
                const int mj = lane < K ? s_mk[lane] : 0;
                const bool over = lane < K && (mj < 0 || s_rec[lane].n + mj > cap);
                const uint64_t ob = __ballot(over);
                const int msum = wave_total(mj);
                if (lane == 0) {
                    const int R0 = s_mk[kResMaxK];
                    int bad = ob != 0ull ? 1 : 0;
                    if (msum != (int)m_tot || R0 < 0 || R0 + s_mk[k] > (int)m_tot) {  // never expected
                        atomicOr(s_err, 2);
                        bad = 2;
                    }
                    s_int[5] = bad;
                }
            }
            __syncthreads();
            if (s_int[5]) {  // a chunk would outgrow its LDS
                // once per step, when the cluster has room: new chunk bounds
                // that share live terms plus incoming children evenly (the
                // children taken as spread evenly over each old chunk), the
                // chunks restaged through the spare live buffer, and rotation
                // s redone
                bool rebalanced = false;
                if (s_int[5] == 1 && rb_step != s && (n_tot + m_tot) * 16 <= (int64_t)K * cap * 15) {
                    if (wave == 0) {
                        const bool in = lane < K;
                        const int nl = in ? s_rec[in ? lane : 0].n : 0, wl = in ? nl + s_mk[lane] : 0;
                        const int pw = wave_incl(wl) - wl, pn = wave_incl(nl) - nl;
                        if (in) {
                            s_rbw[lane] = pw;
                            s_rbn[lane] = pn;
                        }
                    }
                    __syncthreads();
                    if (wave == 0) {
                        int P = 0;
                        if (lane <= K) {
                            const int64_t Wt = n_tot + m_tot;
                            if (lane == K) {
                                P = (int)n_tot;
                            } else if (lane > 0) {
                                const int64_t T = (int64_t)lane * Wt / K;
                                int j = -1;
                                for (int i = 0; i < K; ++i)
                                    if (s_rec[i].n > 0 && s_rbw[i] <= T) j = i;
                                if (j >= 0) {
                                    const int64_t nj_ = s_rec[j].n, wj = nj_ + s_mk[j];
                                    int64_t u = ((T - s_rbw[j]) * nj_ + wj - 1) / wj;
                                    if (u > nj_) u = nj_;
                                    P = s_rbn[j] + (int)u;
                                }
                            }
                            s_rbp[lane] = P;
                        }
                        const int nxtP = __shfl_down(P, 1, kWave);
                        const bool fitsk = lane >= K || (nxtP - P >= 0 && nxtP - P <= cap);
                        const bool roomk = lane >= K || nxtP - P + kAlign <= cap;
                        // 2: every chunk also has room for its edges to move (align_edges)
                        const uint64_t bad = __ballot(!fitsk), tight = __ballot(!roomk);
                        if (lane == 0) s_int[6] = bad != 0ull ? 0 : (tight == 0ull ? 2 : 1);
                    }
                    __syncthreads();
                    if (s_int[6]) {
                        uint64_t* Oz = cur ? dp->Lz[0] : dp->Lz[1];
                        uint64_t* Ox = cur ? dp->Lx[0] : dp->Lx[1];
                        CT* Oc = reinterpret_cast<CT*>(cur ? dp->Lc[0] : dp->Lc[1]);
                        const int off = s_rbn[k];
                        for (int p = tid; p < nj; p += kResThreads) {
                            store_key<W>(Oz, Ox, off + p, sk[p]);
                            stg(Oc, off + p, sc[p]);
                        }
                        nbar += 1;
                        if (!res_barrier(&ctl->bar, nbar * (uint32_t)K, &s_int[3])) {
                            aborted = true;
                            break;
                        }
                        int P0 = s_rbp[k], P1 = s_rbp[k + 1];
                        if (s_int[6] == 2) {  // as for the initial chunks
                            align_edges<W, true>(Oz, Ox, n_tot, P0, P1, s_alv, s_al);
                            P0 = (int)s_al[0];
                            P1 = (int)s_al[1];
                        }
                        for (int p = tid; p < P1 - P0; p += kResThreads) {
                            sk[p] = load_key_l2<W>(Oz, Ox, P0 + p);
                            sc[p] = ld_l2(Oc + P0 + p);
                        }
                        nj = P1 - P0;
                        rb_step = s;
                        nrb += 1;
                        par ^= 1;
                        rebalanced = true;
                    }
                }
                if (rebalanced) continue;
                stall = 3;
                stall_step = s;
                break;
            }
            mk = s_mk[k];
            {
                const int64_t R0 = s_mk[kResMaxK];
                const uint64_t* Cz = dp->Cz;
                const uint64_t* Cx = dp->Cx;
                const CT* Cc = reinterpret_cast<const CT*>(dp->Cc);
                for (int p = tid; p < mk; p += kResThreads) {
                    skw[nj + p] = load_key_l2<W>(Cz, Cx, R0 + p);
                    scw[nj + p] = ld_l2(Cc + R0 + p);
                }
            }
        }
        __syncthreads();
        // ---- phase 3: merge chunk [0, nj) with children [nj, nj + mk) in
        // place: each thread holds its outputs in registers across the barrier
        // that ends every read of the buffer
        {
            const FilterReg fs = res_uni(s_fs);
            const int tot = nj + mk;
            const Key<W>* sL = sk;
            const Key<W>* sC = sk + nj;
            // outputs per thread: the chunk spread over every wave
            const int Ee = (tot + kResThreads - 1) / kResThreads;
            const int dk = tid * Ee < tot ? tid * Ee : tot;
            int lo = dk - mk > 0 ? dk - mk : 0;
            int hi = dk < nj ? dk : nj;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (key_le<W>(sL[mid], sC[dk - 1 - mid])) lo = mid + 1;
                else hi = mid;
            }
            int ia = lo, jb = dk - lo;
            const int stop = dk + Ee < tot ? dk + Ee : tot;
            // a child equal to the live term before this share was absorbed
            // by it (in the previous share); inside a share an equal child is
            // consumed with its live term
            bool skip_first = false;
            if (dk < stop && ia > 0 && jb < mk) skip_first = key_eq<W>(sL[ia - 1], sC[jb]);
            Key<W> outK[E];
            CT outC[E];
            uint32_t keep = 0;
            int cnt = 0;
            const int lmax = Cap - 1;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                outC[e] = coef_zero<CT>();
                if (ia + jb >= stop) break;
                const bool haveL = ia < nj, haveC = jb < mk;
                const int il = ia < lmax ? ia : lmax;
                const int ic = nj + jb < lmax ? nj + jb : lmax;
                const Key<W> a = sk[il], c = sk[ic];
                const CT ca = sc[il], cc = sc[ic];
                const bool takeL = haveL & (!haveC | key_le<W>(a, c));
                const bool eq = takeL & haveC & key_eq<W>(a, c);
                Key<W> kv;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    kv.z[w] = takeL ? a.z[w] : c.z[w];
                    kv.x[w] = takeL ? a.x[w] : c.x[w];
                }
                CT acc = ca;
                if (G.anti(a)) acc = acc * G.cosv;
                if (eq) acc = acc + cc;  // the equal-key child follows its live term
                if (!takeL) acc = cc;
                const bool emit = !(e == 0 && skip_first && !takeL);
                ia += takeL ? 1 : 0;
                jb += (!takeL | eq) ? 1 : 0;
                outK[e] = kv;
                if (emit && !should_drop<W>(fs, kv, coef_abs(acc))) {
                    keep |= 1u << e;
                    outC[e] = acc;
                    ++cnt;
                }
            }
            const int incl = wave_incl(cnt);
            if (lane == kWave - 1) s_wsum[wave] = incl;
            __syncthreads();
            const int wv = lane < kResWaves ? s_wsum[lane] : 0;
            const int wincl = wave_incl(wv);
            const int wpref = lane_of(wincl - wv, wave);
            const int total = lane_of(wincl, kResWaves - 1);
            int o = wpref + incl - cnt;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if ((keep >> e) & 1u) {
                    sk[o] = outK[e];
                    sc[o] = outC[e];
                    ++o;
                }
            }
            nj = total;
        }
        if (tid == 0) {
            s_bk[kBkPn] = n_tot;
            s_bk[kBkPm] = m_tot;
        }
        pend = true;
        par ^= 1;
        cs ^= 1;
    }

    if (aborted) return;  // kr_commit sees the abort bit and discards the launch
    // ---- write-out of the final live set (or of the state before the stall)
    __syncthreads();
    const bool write = s_bk[kBkRot] > 0;
    if (write) {
        int64_t off = 0;
        for (int i = 0; i < k; ++i) off += s_rec[i].n;
        const int mine = nj;
        uint64_t* Oz = cur ? dp->Lz[0] : dp->Lz[1];
        uint64_t* Ox = cur ? dp->Lx[0] : dp->Lx[1];
        CT* Oc = reinterpret_cast<CT*>(cur ? dp->Lc[0] : dp->Lc[1]);
        for (int p = tid; p < mine; p += kResThreads) {
            store_key<W>(Oz, Ox, off + p, s_k[p]);
            stg(Oc, off + p, s_c[p]);
        }
    }
    if (tid == 0) {
        const int e = *s_err;
        if (e) atomicOr(&ctl->err, e);
    }
    if (k == 0 && tid == 0) {
        int64_t n_final = 0;
        for (int i = 0; i < K; ++i) n_final += s_rec[i].n;
        ctl->wrote = write ? 1 : 0;
        ctl->stall = stall;
        ctl->stall_step = stall_step;
        ctl->need = need;
        ctl->n_final = write ? n_final : n0;
        ctl->rotations = s_bk[kBkRot];
        ctl->swept = s_bk[kBkSwept];
        ctl->created = s_bk[kBkCreated];
        ctl->merge_in = s_bk[kBkMin];
        ctl->merge_out = s_bk[kBkMout];
        ctl->early = s_bk[kBkEarly];
        ctl->live_cap_v = s_bk[kBkLive];
        ctl->app_cap_v = s_bk[kBkApp];
        ctl->rebalances = nrb;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __hip_atomic_store(&ctl->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Before a resident launch: counters, tickets and staged results cleared.
__global__ void kr_reset(ResCtl* __restrict__ ctl, uint32_t* __restrict__ tickets, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < kXcdCount) tickets[b] = 0u;
    if (b >= B) return;
    ResCtl* c = ctl + b;
    c->bar = 0u;
    c->done = 0;
    c->err = 0;
}

// After a resident launch: apply the staged results of every vector whose
// cluster finished; a vector whose launch was aborted (or never ran) is left
// as it was and marked for replay from the call's first step (code 4).
// work[0..3] accumulate the rotations, swept terms and merge outputs of the
// committed launches (the resident kernel's share of the counters) and their
// chunk rebalances.
__global__ void kr_commit(const ResCtl* __restrict__ ctl, const int32_t* __restrict__ part, SegState* __restrict__ S,
                          int B, int64_t first, unsigned long long* __restrict__ work) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || !part[b]) return;
    const ResCtl* c = ctl + b;
    SegState* s = S + b;
    if (s->stalled) return;
    if ((c->bar & kResAbort) || c->done != 1) {
        s->stalled = 4;
        s->stall_step = first;
        s->need = 0;
        return;
    }
    if (c->err) s->err |= c->err << 8;
    if (c->wrote) {
        const int cur = s->cur ^ s->flip;
        s->n[cur ^ 1] = c->n_final;
        s->flip ^= 1;
    }
    s->rotations += c->rotations;
    s->swept += c->swept;
    s->created += c->created;
    s->merge_in += c->merge_in;
    s->merge_out += c->merge_out;
    s->early += c->early;
    atomicAdd(work + 0, (unsigned long long)c->rotations);
    atomicAdd(work + 1, (unsigned long long)c->swept);
    atomicAdd(work + 2, (unsigned long long)c->merge_out);
    atomicAdd(work + 3, (unsigned long long)c->rebalances);
    s->live_cap_v = c->live_cap_v;
    s->app_cap_v = c->app_cap_v;
    if (c->stall) {
        s->stalled = c->stall;
        s->stall_step = c->stall_step;
        s->need = c->need;
    }
}

}  // namespace pp
