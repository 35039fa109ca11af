// libpauliprop: host engine and C ABI of the MI355X SparsePauliVector path.
//
// The driver keeps the reference's semantics (spv_evolve.jl:312-387 at
// window = 1): before every rotation the append-capacity precheck
// `n > length(az)` may force an early merge (a strict-filter compaction) and
// grows the virtual append capacity; each rotation is sweep + merge under
// the strict filter.  Device buffers are sized independently of those
// virtual capacities and grow on demand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pauliprop.h"
#include "pp_common.hpp"
#include "pp_kernels.hpp"

using namespace pp;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw PPError(_e == hipErrorOutOfMemory ? PP_ENOMEM : PP_EHIP,                        \
                          std::string(#expr) + ": " + hipGetErrorString(_e));                     \
    } while (0)

struct PPError {
    int code;
    std::string msg;
    PPError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

template <typename T>
T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) throw PPError(PP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return static_cast<T*>(p);
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

struct DevSet {
    uint64_t* z = nullptr;
    uint64_t* x = nullptr;
    double* c = nullptr;
    int64_t cap = 0;
};

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Julia's Float64^Int: base/math.jl pow_body (compensated power by squaring),
// with muladd taken as fused.
double julia_powi(double x, int64_t n) {
    if (n == 0) return 1.0;
    if (n < -4096 || n > 3 * 8192) return std::pow(x, (double)n);
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    if (n == 3) return x * x * x;
    if (n < 0) {
        const double rx = 1.0 / x;
        if (n == -2) return rx * rx;
        if (std::isfinite(x)) xnlo = -std::fma(x, rx, -1.0) * rx;
        x = rx;
        n = -n;
    }
    while (n > 1) {
        if (n & 1) {
            const double err = std::fma(y, xnlo, x * ynlo);
            const double p = x * y;
            const double plo = std::fma(x, y, -p);
            y = p;
            ynlo = plo + err;
        }
        const double err = x * 2 * xnlo;
        const double p = x * x;
        const double plo = std::fma(x, x, -p);
        x = p;
        xnlo = plo + err;
        n >>= 1;
    }
    const double err = std::fma(y, xnlo, x * ynlo);
    return (std::isfinite(x) && std::isfinite(err)) ? std::fma(x, y, err) : x * y;
}

FilterDev make_filter_dev(const pp_filter* f) {
    FilterDev d;
    if (f) {
        d.wmax = f->wmax; d.xwmax = f->xwmax; d.mwmax = f->mwmax;
        d.thresh = f->thresh; d.alpha_w = f->alpha_w; d.dthresh_w = f->dthresh_w;
        d.alpha_xw = f->alpha_xw; d.dthresh_xw = f->dthresh_xw;
    } else {  // NOFILTER, spv_kernels.jl:44-45
        d.wmax = d.xwmax = d.mwmax = INT64_MAX;
        d.thresh = -1.0; d.alpha_w = 0.0; d.dthresh_w = -1.0; d.alpha_xw = 0.0; d.dthresh_xw = -1.0;
    }
    for (int w = 0; w < kExpTab; ++w) {
        d.exptab_w[w] = std::exp(-d.alpha_w * (double)w);
        d.exptab_xw[w] = std::exp(-d.alpha_xw * (double)w);
    }
    return d;
}

// A filter that can drop a freshly created child (anything but NOFILTER-like).
bool filter_active(const pp_filter* f) {
    if (!f) return false;
    return f->thresh >= 0.0 || f->wmax != INT64_MAX || f->xwmax != INT64_MAX || f->mwmax != INT64_MAX ||
           f->dthresh_w >= 0.0 || f->dthresh_xw >= 0.0;
}

}  // namespace

struct pp_handle {
    int N = 0, W = 1;
    int ncu = 256;
    hipStream_t stream = nullptr;
    DevSet L[2];
    int cur = 0;
    DevSet C;
    uint64_t* flags = nullptr;
    uint16_t* wpre = nullptr;
    int64_t cap_words = 0;
    int32_t* tilecnt = nullptr;
    int64_t* tileoff = nullptr;
    int64_t cap_tiles = 0;
    int64_t* bnd = nullptr;
    int64_t cap_bnd = 0;
    int64_t* split = nullptr;
    uint64_t* status = nullptr;
    int64_t cap_split = 0;
    uint32_t* ticket = nullptr;
    int64_t* d_cnt = nullptr;   // [0], [1] live counts of L[0], L[1]; [2] m; [3] constant 0
    int32_t* d_ovf = nullptr;   // [0] overflow, [1] check flag, [2] internal error bits
    GenDev* d_gens = nullptr;
    int64_t cap_gens = 0;
    FilterDev* d_filt = nullptr;  // [0] strict, [1] local, [2] NOFILTER
    double* d_tabs = nullptr;     // 3 x 257 channel tables
    double2* d_partial = nullptr;
    double* d_out = nullptr;
    int64_t* h_pin = nullptr;     // pinned scalars
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // reference capacity bookkeeping (type_SparsePauliVector.jl:69-126)
    int64_t live_cap_v = 16, app_cap_v = 16;
    int64_t n = 0;  // live count (host copy)
    double m_ratio = 1.0;  // children per live term in the last rotation

    ~pp_handle() {
        for (int s = 0; s < 2; ++s) { dfree(L[s].z); dfree(L[s].x); dfree(L[s].c); }
        dfree(C.z); dfree(C.x); dfree(C.c);
        dfree(flags); dfree(wpre); dfree(tilecnt); dfree(tileoff); dfree(bnd); dfree(split); dfree(status);
        dfree(ticket); dfree(d_cnt); dfree(d_ovf); dfree(d_gens); dfree(d_filt); dfree(d_tabs);
        dfree(d_partial); dfree(d_out);
        if (h_pin) (void)hipHostFree(h_pin);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    size_t key_bytes() const { return sizeof(uint64_t) * (size_t)W; }

    void alloc_set(DevSet& s, int64_t cap) {
        s.z = dalloc<uint64_t>((size_t)cap * W);
        s.x = dalloc<uint64_t>((size_t)cap * W);
        s.c = dalloc<double>((size_t)cap);
        s.cap = cap;
    }
    void free_set(DevSet& s) {
        dfree(s.z); dfree(s.x); dfree(s.c);
        s = DevSet();
    }

    // Grow one buffer set to hold `need` terms (+25% headroom).  keep > 0
    // copies that many leading terms across; otherwise the old contents are
    // dropped first, which keeps the peak footprint at one copy.
    void grow_set(DevSet& s, int64_t need, int64_t keep) {
        if (s.cap >= need) return;
        const int64_t cap = std::max<int64_t>(need + need / 4, 1024);
        if (keep <= 0) {
            free_set(s);
            alloc_set(s, cap);
            return;
        }
        DevSet nw;
        alloc_set(nw, cap);
        HIPCHK(hipMemcpyAsync(nw.z, s.z, (size_t)keep * key_bytes(), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(nw.x, s.x, (size_t)keep * key_bytes(), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(nw.c, s.c, (size_t)keep * sizeof(double), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        free_set(s);
        s = nw;
    }

    // Current live buffer holds `need` terms (keeping the live set); the
    // merge target is sized separately per call.
    void ensure_live(int64_t need) { grow_set(L[cur], need, n); }
    void ensure_target(int64_t need) { grow_set(L[cur ^ 1], need, 0); }
    void ensure_children(int64_t need) { grow_set(C, need, 0); }

    void ensure_scratch(int64_t nlive, int nlev) {
        const int64_t words = cdiv(std::max<int64_t>(nlive, 1), 64) + 1;
        if (cap_words < words) {
            dfree(flags); dfree(wpre);
            cap_words = std::max(words, 2 * cap_words);
            flags = dalloc<uint64_t>(cap_words);
            wpre = dalloc<uint16_t>(cap_words);
        }
        const int64_t tiles = cdiv(std::max<int64_t>(nlive, 1), kTile) + 2;
        if (cap_tiles < tiles) {
            dfree(tilecnt); dfree(tileoff);
            cap_tiles = std::max(tiles, 2 * cap_tiles);
            tilecnt = dalloc<int32_t>(cap_tiles);
            tileoff = dalloc<int64_t>(cap_tiles);
        }
        const int64_t nb = tiles * std::max(nlev, 1) * 4;
        if (cap_bnd < nb) {
            dfree(bnd);
            cap_bnd = std::max(nb, 2 * cap_bnd);
            bnd = dalloc<int64_t>(cap_bnd);
        }
        // merge tiles: outputs <= 2 n (live + at most one child per live term)
        const int TM = W == 1 ? MergeTile<1>::kSize : MergeTile<2>::kSize;
        const int64_t sp = cdiv(2 * std::max<int64_t>(nlive, 1), TM) + 2;
        if (cap_split < sp) {
            dfree(split); dfree(status);
            cap_split = std::max(sp, 2 * cap_split);
            split = dalloc<int64_t>(cap_split);
            status = dalloc<uint64_t>(cap_split);
        }
    }

    void ensure_gens(int64_t Lcount) {
        if (cap_gens >= Lcount) return;
        dfree(d_gens);
        cap_gens = std::max<int64_t>(Lcount, 2 * cap_gens);
        d_gens = dalloc<GenDev>(cap_gens);
    }

    GenDev make_gen(const uint64_t* gz, const uint64_t* gx, double cosv, double sinv) const {
        GenDev g;
        std::memset(&g, 0, sizeof(g));
        int ng = 0;
        for (int w = 0; w < W; ++w) {
            g.gz[w] = gz[w];
            g.gx[w] = gx[w];
            ng += __builtin_popcountll(gz[w] & gx[w]);
        }
        g.ng = ng;
        g.cosv = cosv;
        g.sinv = sinv;
        int nlev = 0;
        for (int w = 0; w < W; ++w)  // x bits: combined positions [0, 64 W)
            for (int b = 0; b < 64; ++b)
                if ((gx[w] >> b) & 1ull) g.lev[nlev++] = (uint8_t)(64 * w + b);
        for (int w = 0; w < W; ++w)  // z bits above
            for (int b = 0; b < 64; ++b)
                if ((gz[w] >> b) & 1ull) g.lev[nlev++] = (uint8_t)(64 * W + 64 * w + b);
        g.nlev = nlev;
        return g;
    }

    int grid_for(int64_t items, int per_block_items, int max_per_cu = 8) const {
        const int64_t need = cdiv(std::max<int64_t>(items, 1), per_block_items);
        return (int)std::min<int64_t>(need, (int64_t)ncu * max_per_cu);
    }

    // Read back the live count written by the last merge, the child count and
    // the overflow flag (synchronises).  Internal consistency failures throw.
    int64_t sync_count(int idx, int32_t* ovf) {
        HIPCHK(hipMemcpyAsync(h_pin, d_cnt + idx, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(h_pin + 1, d_cnt + 2, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(h_pin + 2, d_ovf, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        const int32_t* flags3 = (const int32_t*)(h_pin + 2);
        if (flags3[2] != 0)
            throw PPError(PP_EHIP, "internal consistency check failed in a kernel (bits " + std::to_string(flags3[2]) + ")");
        *ovf = flags3[0];
        return h_pin[0];
    }

    template <int Wt>
    void launch_merge(const GenDev* g, const FilterDev* f, const DevSet& src, const int64_t* d_n, const DevSet& ch,
                      const int64_t* d_m, int64_t mbound, DevSet& dst, int64_t* d_nout, bool timed = false) {
        constexpr int TM = MergeTile<Wt>::kSize;
        const int64_t ntm_ub = cdiv(n + mbound, TM);
        const int64_t threads = ntm_ub + 1;
        ChanDev nochan{};
        hipLaunchKernelGGL(k_partition<Wt>, dim3((unsigned)cdiv(threads, kBlock)), dim3(kBlock), 0, stream, src.z,
                           src.x, d_n, ch.z, ch.x, d_m, split, status, ticket, d_nout, threads, d_ovf);
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(ntm_ub, (int64_t)ncu * 4));
        if (timed) HIPCHK(hipEventRecord(ev[2], stream));
        hipLaunchKernelGGL(k_merge<Wt>, dim3(grid), dim3(kBlock), 0, stream, src.z, src.x, src.c, d_n, ch.z, ch.x,
                           ch.c, d_m, split, g, f, nochan, 0, dst.z, dst.x, dst.c, dst.cap, status, ticket, d_nout,
                           d_ovf, d_ovf + 2);
        HIPCHK(hipGetLastError());
    }

    // One rotation at window = 1 on the current live set (host count n).
    // Buffers are sized from the previous rotation's child ratio; if the
    // children or the merged output do not fit, the kernels raise the
    // overflow flag and write nothing out of bounds, the live set is
    // untouched, and the rotation is re-run with exact sizes (the reference's
    // grow-and-retry merge, spv_kernels.jl:377-383).
    template <int Wt>
    void rotate(const GenDev* dg, const GenDev& hg, int local_active, bool use_strict, pp_counters* cnt) {
        ensure_scratch(n, hg.nlev);
        const int64_t m_est = std::min<int64_t>(n, (int64_t)(m_ratio * (double)n * 1.15) + 4096);
        ensure_children(m_est + 1);
        ensure_target(n + m_est + 1);
        int64_t n_out = 0, m = 0;
        for (int attempt = 0;; ++attempt) {
            launch_rotation<Wt>(dg, hg, local_active, use_strict, cnt != nullptr);
            int32_t ovf = 0;
            n_out = sync_count(cur ^ 1, &ovf);
            m = h_pin[1];
            if (!ovf) break;
            if (attempt >= 2) throw PPError(PP_EOVERFLOW, "merge overflow persisted after regrowth");
            HIPCHK(hipMemsetAsync(d_ovf, 0, sizeof(int32_t), stream));
            ensure_children(m + 1);
            ensure_target(n + m + 1);
        }
        if (cnt) {
            float a = 0.f, b = 0.f;
            HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
            HIPCHK(hipEventElapsedTime(&b, ev[2], ev[3]));
            cnt->t_rotate += a * 1e-3;
            cnt->t_merge += b * 1e-3;
            cnt->rotations += 1;
            cnt->terms_swept += n;
            cnt->terms_created += m;
            cnt->merge_in += n + m;
            cnt->merge_out += n_out;
        }
        if (n > 0 && hg.nlev > 0) m_ratio = (double)m / (double)n;
        // _merge_spv! growth of the reference's live buffers (spv_kernels.jl:377-383)
        if (n_out > live_cap_v) live_cap_v = std::max(2 * live_cap_v, n + m);
        cur ^= 1;
        n = n_out;
    }

    template <int Wt>
    void launch_rotation(const GenDev* dg, const GenDev& hg, int local_active, bool use_strict, bool timed) {
        DevSet& src = L[cur];
        DevSet& dst = L[cur ^ 1];
        const int64_t* d_n = d_cnt + cur;
        int64_t* d_nout = d_cnt + (cur ^ 1);
        const FilterDev* fstrict = d_filt + (use_strict ? 0 : 2);
        if (timed) HIPCHK(hipEventRecord(ev[0], stream));
        if (hg.nlev > 0 && n > 0) {
            const int64_t ntiles = cdiv(n, kTile);
            const int gt = (int)std::min<int64_t>(ntiles, (int64_t)ncu * 8);
            hipLaunchKernelGGL(k_flags<Wt>, dim3(gt), dim3(kBlock), 0, stream, src.z, src.x, src.c, d_n, dg,
                               d_filt + 1, local_active, flags, wpre, tilecnt);
            hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, tilecnt, d_n, tileoff, d_cnt + 2, C.cap, d_ovf);
            const int64_t nb = ntiles * hg.nlev * 2;
            hipLaunchKernelGGL(k_bounds<Wt>, dim3((unsigned)std::min<int64_t>(cdiv(nb, kBlock), (int64_t)ncu * 16)),
                               dim3(kBlock), 0, stream, src.z, src.x, d_n, dg, flags, wpre, tileoff, bnd);
            hipLaunchKernelGGL(k_emit<Wt>, dim3(gt), dim3(kBlock), 0, stream, src.z, src.x, src.c, d_n, dg, flags,
                               wpre, tileoff, bnd, C.z, C.x, C.c, d_ovf, d_ovf + 2);
            HIPCHK(hipGetLastError());
        } else {
            HIPCHK(hipMemsetAsync(d_cnt + 2, 0, sizeof(int64_t), stream));
        }
        if (timed) HIPCHK(hipEventRecord(ev[1], stream));
        launch_merge<Wt>(dg, fstrict, src, d_n, C, d_cnt + 2, hg.nlev > 0 ? n : 0, dst, d_nout, timed);
        if (timed) HIPCHK(hipEventRecord(ev[3], stream));
    }

    template <int Wt>
    void compact(const FilterDev* f) {
        ensure_scratch(n, 0);
        ensure_target(n + 1);
        DevSet empty;
        empty.z = C.z ? C.z : L[cur].z;
        empty.x = C.x ? C.x : L[cur].x;
        empty.c = C.c ? C.c : L[cur].c;
        launch_merge<Wt>(nullptr, f, L[cur], d_cnt + cur, empty, d_cnt + 3, 0, L[cur ^ 1], d_cnt + (cur ^ 1));
        int32_t ovf = 0;
        const int64_t n_out = sync_count(cur ^ 1, &ovf);
        if (ovf) throw PPError(PP_EOVERFLOW, "compaction overflow");
        cur ^= 1;
        n = n_out;
    }

    void upload_filters(const pp_filter* strict, const pp_filter* local) {
        FilterDev f[3] = {make_filter_dev(strict), make_filter_dev(local), make_filter_dev(nullptr)};
        HIPCHK(hipMemcpyAsync(d_filt, f, sizeof(f), hipMemcpyHostToDevice, stream));
    }

    template <int Wt>
    void evolve(const uint64_t* gz, const uint64_t* gx, const double* cosv, const double* sinv, int64_t Lcount,
                const pp_filter* strict, const pp_filter* local, pp_counters* cnt) {
        std::vector<GenDev> hg((size_t)Lcount);
        for (int64_t i = 0; i < Lcount; ++i) hg[i] = make_gen(gz + i * W, gx + i * W, cosv[i], sinv[i]);
        ensure_gens(Lcount);
        HIPCHK(hipMemcpyAsync(d_gens, hg.data(), (size_t)Lcount * sizeof(GenDev), hipMemcpyHostToDevice, stream));
        upload_filters(strict, local);
        const int local_active = filter_active(local) ? 1 : 0;
        for (int64_t i = 0; i < Lcount; ++i) {
            // capacity precheck, spv_evolve.jl:356-362 (no pending appends at window = 1)
            if (n > app_cap_v) {
                const int64_t before = n;
                compact<Wt>(d_filt + 0);
                if (cnt) {
                    cnt->early_merges += 1;
                    cnt->merge_in += before;
                    cnt->merge_out += n;
                }
                if (n > app_cap_v) app_cap_v = std::max(2 * app_cap_v, n);
            }
            rotate<Wt>(d_gens + i, hg[i], local_active, true, cnt);
        }
    }

    template <int Wt>
    void evolve_single(const uint64_t* gz, const uint64_t* gx, double cosv, double sinv) {
        GenDev g = make_gen(gz, gx, cosv, sinv);
        ensure_gens(1);
        HIPCHK(hipMemcpyAsync(d_gens, &g, sizeof(GenDev), hipMemcpyHostToDevice, stream));
        upload_filters(nullptr, nullptr);
        // spv_evolve.jl:231
        if (n > app_cap_v) app_cap_v = std::max(2 * app_cap_v, n);
        rotate<Wt>(d_gens, g, 0, false, nullptr);
    }
};

// ---------------------------------------------------------------------------

#define API_BEGIN try {
#define API_END                                           \
    }                                                     \
    catch (const PPError& e) {                            \
        return fail(e.code, e.msg);                       \
    }                                                     \
    catch (const std::exception& e) {                     \
        return fail(PP_EHIP, e.what());                   \
    }                                                     \
    g_last_error.clear();                                 \
    return PP_OK;

extern "C" {

const char* pp_last_error(void) { return g_last_error.c_str(); }

const char* pp_version(void) { return "pauliprop 0.1 gfx950"; }

double pp_julia_powi(double x, int64_t n) { return julia_powi(x, n); }

int pp_create(int N, int64_t live_cap, int64_t append_cap, pp_handle** out) {
    if (!out) return fail(PP_EARG, "out is NULL");
    *out = nullptr;
    if (N < 1 || N > 128) return fail(PP_EARG, "N must be in 1:128");
    if (live_cap < 1 || append_cap < 1) return fail(PP_EARG, "capacities must be >= 1");
    pp_handle* h = nullptr;
    API_BEGIN
    h = new pp_handle();
    h->N = N;
    h->W = N <= 64 ? 1 : 2;
    h->live_cap_v = live_cap;
    h->app_cap_v = append_cap;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    h->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (auto& e : h->ev) HIPCHK(hipEventCreate(&e));
    h->ticket = dalloc<uint32_t>(1);
    h->d_cnt = dalloc<int64_t>(4);
    h->d_ovf = dalloc<int32_t>(4);
    h->d_filt = dalloc<FilterDev>(3);
    h->d_tabs = dalloc<double>(3 * kExpTab);
    h->d_partial = dalloc<double2>(4096);
    h->d_out = dalloc<double>(2);
    HIPCHK(hipHostMalloc((void**)&h->h_pin, 8 * sizeof(int64_t), hipHostMallocDefault));
    HIPCHK(hipMemsetAsync(h->d_cnt, 0, 4 * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->d_ovf, 0, 4 * sizeof(int32_t), h->stream));
    h->ensure_live(std::max<int64_t>(live_cap, 16));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = h;
    h = nullptr;
    }
    catch (const PPError& e) {
        delete h;
        return fail(e.code, e.msg);
    }
    g_last_error.clear();
    return PP_OK;
}

int pp_destroy(pp_handle* h) {
    if (!h) return PP_OK;
    (void)hipStreamSynchronize(h->stream);
    delete h;
    return PP_OK;
}

int pp_upload(pp_handle* h, const uint64_t* z, const uint64_t* x, const double* c, int64_t n, int sorted) {
    if (!h) return fail(PP_EARG, "null handle");
    if (n < 0 || (n > 0 && (!z || !x || !c))) return fail(PP_EARG, "bad arrays");
    API_BEGIN
    const int W = h->W;
    std::vector<uint64_t> hz(z, z + n * W), hx(x, x + n * W);
    std::vector<double> hc(c, c + n);
    auto less = [&](int64_t a, int64_t b) {
        for (int w = W - 1; w >= 0; --w)
            if (hz[a * W + w] != hz[b * W + w]) return hz[a * W + w] < hz[b * W + w];
        for (int w = W - 1; w >= 0; --w)
            if (hx[a * W + w] != hx[b * W + w]) return hx[a * W + w] < hx[b * W + w];
        return false;
    };
    if (!sorted) {
        std::vector<int64_t> perm((size_t)n);
        std::iota(perm.begin(), perm.end(), 0);
        std::sort(perm.begin(), perm.end(), less);
        std::vector<uint64_t> sz((size_t)n * W), sx((size_t)n * W);
        std::vector<double> sc((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            for (int w = 0; w < W; ++w) {
                sz[i * W + w] = hz[perm[i] * W + w];
                sx[i * W + w] = hx[perm[i] * W + w];
            }
            sc[i] = hc[perm[i]];
        }
        hz.swap(sz);
        hx.swap(sx);
        hc.swap(sc);
    }
    for (int64_t i = 1; i < n; ++i)
        if (!less(i - 1, i))
            throw PPError(PP_EARG, sorted ? "keys not strictly increasing" : "duplicate keys in upload");
    h->n = 0;
    h->ensure_live(std::max<int64_t>(n, 1));
    h->m_ratio = 1.0;
    DevSet& s = h->L[h->cur];
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(s.z, hz.data(), (size_t)n * W * 8, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(s.x, hx.data(), (size_t)n * W * 8, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(s.c, hc.data(), (size_t)n * 8, hipMemcpyHostToDevice, h->stream));
    }
    h->h_pin[3] = n;
    HIPCHK(hipMemcpyAsync(h->d_cnt + h->cur, &h->h_pin[3], sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->n = n;
    if (n > h->live_cap_v) h->live_cap_v = n;
    API_END
}

int pp_download(pp_handle* h, uint64_t* z, uint64_t* x, double* c, int64_t cap, int64_t* n) {
    if (!h || !n) return fail(PP_EARG, "null argument");
    API_BEGIN
    *n = h->n;
    if (cap < h->n) throw PPError(PP_EARG, "output capacity too small");
    const DevSet& s = h->L[h->cur];
    if (h->n > 0) {
        HIPCHK(hipMemcpyAsync(z, s.z, (size_t)h->n * h->W * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(x, s.x, (size_t)h->n * h->W * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(c, s.c, (size_t)h->n * 8, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    API_END
}

int pp_size(pp_handle* h, int64_t* n) {
    if (!h || !n) return fail(PP_EARG, "null argument");
    *n = h->n;
    return PP_OK;
}

int pp_capacities(pp_handle* h, int64_t* live_cap, int64_t* append_cap) {
    if (!h || !live_cap || !append_cap) return fail(PP_EARG, "null argument");
    *live_cap = h->live_cap_v;
    *append_cap = h->app_cap_v;
    return PP_OK;
}

int pp_evolve(pp_handle* h, const uint64_t* gz, const uint64_t* gx, const double* cosv, const double* sinv,
              int64_t L, int64_t window, const pp_filter* strict, const pp_filter* local, pp_counters* counters) {
    if (!h) return fail(PP_EARG, "null handle");
    if (window < 1) return fail(PP_EARG, "window must be >= 1");
    if (L < 0) return fail(PP_EARG, "L must be >= 0");
    if (L > 0 && (!gz || !gx || !cosv || !sinv)) return fail(PP_EARG, "null generator arrays");
    if (window != 1) return fail(PP_EUNSUPPORTED, "window > 1 is not available in this build");
    if (counters) std::memset(counters, 0, sizeof(*counters));
    if (L == 0) return PP_OK;
    API_BEGIN
    if (h->W == 1) h->evolve<1>(gz, gx, cosv, sinv, L, strict, local, counters);
    else h->evolve<2>(gz, gx, cosv, sinv, L, strict, local, counters);
    API_END
}

int pp_evolve_single(pp_handle* h, const uint64_t* gz, const uint64_t* gx, double cosv, double sinv) {
    if (!h || !gz || !gx) return fail(PP_EARG, "null argument");
    API_BEGIN
    if (h->W == 1) h->evolve_single<1>(gz, gx, cosv, sinv);
    else h->evolve_single<2>(gz, gx, cosv, sinv);
    API_END
}

int pp_compact(pp_handle* h, const pp_filter* f) {
    if (!h || !f) return fail(PP_EARG, "null argument");
    API_BEGIN
    h->upload_filters(f, nullptr);
    if (h->W == 1) h->compact<1>(h->d_filt + 0);
    else h->compact<2>(h->d_filt + 0);
    API_END
}

int pp_scale_by_popcount(pp_handle* h, int kind, const uint64_t* mask, const double* t0, const double* t1,
                         const double* t2) {
    if (!h || !mask || !t0) return fail(PP_EARG, "null argument");
    if (kind < 0 || kind > 4) return fail(PP_EARG, "kind must be in 0:4");
    if (kind == 4 && (!t1 || !t2)) return fail(PP_EARG, "general channel needs three tables");
    API_BEGIN
    const int nt = h->N + 1;
    std::vector<double> tabs(3 * kExpTab, 0.0);
    for (int k = 0; k < nt; ++k) {
        tabs[k] = t0[k];
        if (kind == 4) {
            tabs[kExpTab + k] = t1[k];
            tabs[2 * kExpTab + k] = t2[k];
        }
    }
    HIPCHK(hipMemcpyAsync(h->d_tabs, tabs.data(), tabs.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    ChanDev ch{};
    ch.kind = kind;
    for (int w = 0; w < h->W; ++w) ch.mask[w] = mask[w];
    ch.t0 = h->d_tabs;
    ch.t1 = h->d_tabs + kExpTab;
    ch.t2 = h->d_tabs + 2 * kExpTab;
    DevSet& s = h->L[h->cur];
    if (h->n > 0) {
        const int grid = h->grid_for(h->n, kBlock * 4);
        if (h->W == 1)
            hipLaunchKernelGGL(k_channel<1>, dim3(grid), dim3(kBlock), 0, h->stream, s.z, s.x, s.c, h->d_cnt + h->cur, ch);
        else
            hipLaunchKernelGGL(k_channel<2>, dim3(grid), dim3(kBlock), 0, h->stream, s.z, s.x, s.c, h->d_cnt + h->cur, ch);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    API_END
}

int pp_expectation_ket(pp_handle* h, const uint64_t* ket, double* out) {
    if (!h || !ket || !out) return fail(PP_EARG, "null argument");
    API_BEGIN
    const DevSet& s = h->L[h->cur];
    const int np = (int)std::min<int64_t>(std::max<int64_t>(1, cdiv(h->n, kBlock * 8)), 4096);
    const uint64_t k0 = ket[0], k1 = h->W == 2 ? ket[1] : 0;
    if (h->W == 1)
        hipLaunchKernelGGL(k_expect<1>, dim3(np), dim3(kBlock), 0, h->stream, s.z, s.x, s.c, h->d_cnt + h->cur, k0, k1,
                           h->d_partial);
    else
        hipLaunchKernelGGL(k_expect<2>, dim3(np), dim3(kBlock), 0, h->stream, s.z, s.x, s.c, h->d_cnt + h->cur, k0, k1,
                           h->d_partial);
    hipLaunchKernelGGL(k_expect_final, dim3(1), dim3(kBlock), 0, h->stream, h->d_partial, np, h->d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, h->d_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    API_END
}

int pp_check(pp_handle* h, int* ok) {
    if (!h || !ok) return fail(PP_EARG, "null argument");
    API_BEGIN
    HIPCHK(hipMemsetAsync(h->d_ovf + 1, 0, sizeof(int32_t), h->stream));
    const DevSet& s = h->L[h->cur];
    const int grid = h->grid_for(h->n, kBlock * 4);
    if (h->W == 1)
        hipLaunchKernelGGL(k_check<1>, dim3(grid), dim3(kBlock), 0, h->stream, s.z, s.x, h->d_cnt + h->cur, h->d_ovf + 1);
    else
        hipLaunchKernelGGL(k_check<2>, dim3(grid), dim3(kBlock), 0, h->stream, s.z, s.x, h->d_cnt + h->cur, h->d_ovf + 1);
    HIPCHK(hipGetLastError());
    int32_t bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, h->d_ovf + 1, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *ok = bad ? 0 : 1;
    API_END
}

}  // extern "C"
