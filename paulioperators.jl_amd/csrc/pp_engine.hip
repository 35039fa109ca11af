// libpauliprop: host engine and C ABI of the MI355X SparsePauliVector path.
//
// A Batch holds B independent vectors with the same word count; every
// operation on it is a short, fixed sequence of kernel launches whose grids
// depend only on buffer capacities, so a whole rotation sequence is enqueued
// without waiting on the device.  Live counts, the reference's capacity
// bookkeeping (spv_evolve.jl:312-387 at window = 1: the append-capacity
// precheck that may force an early merge before a rotation) and the
// WindowCounters sums live in device memory (SegState).  A vector whose
// buffers are too small freezes; after the sequence the host regrows its
// buffers and replays it from the step that stalled.  The single-vector API
// is a batch of one.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pauliprop.h"
#include "pp_common.hpp"
#include "pp_kernels.hpp"
#include "pp_window.hpp"
#include "pp_ops.hpp"
#include "pp_algebra.hpp"
#include "pp_resident.hpp"

using namespace pp;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

struct PPError {
    int code;
    std::string msg;
    PPError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw PPError(_e == hipErrorOutOfMemory ? PP_ENOMEM : PP_EHIP,                        \
                          std::string(#expr) + ": " + hipGetErrorString(_e));                     \
    } while (0)

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

// Device allocations freed when the owner leaves scope (also on errors).
// Grow-only device buffers in numbered slots, kept between calls (the
// products reuse them instead of allocating and freeing per call).
struct DevArena {
    static constexpr int kSlots = 14;
    void* p[kSlots] = {};
    size_t bytes[kSlots] = {};
    template <typename T>
    T* get(int slot, size_t n) {
        const size_t need = (n ? n : 1) * sizeof(T);
        if (bytes[slot] < need) {
            dfree(p[slot]);
            p[slot] = nullptr;
            bytes[slot] = 0;
            const size_t grow = need + need / 4;
            p[slot] = dalloc<char>(grow);
            bytes[slot] = grow;
        }
        return static_cast<T*>(p[slot]);
    }
    void release() {
        for (int i = 0; i < kSlots; ++i) dfree(p[i]);
        *this = DevArena();
    }
};

struct DevSet {
    uint64_t* z = nullptr;
    uint64_t* x = nullptr;
    double* c = nullptr;
    int64_t cap = 0;
};

// Per-vector device buffers (owned).
struct InstBuf {
    DevSet L[2];
    DevSet C;
    uint64_t* flags = nullptr;
    uint16_t* wpre = nullptr;
    int32_t* tilecnt = nullptr;
    int64_t* tileoff = nullptr;
    int64_t* bnd = nullptr;
    uint64_t* agg = nullptr;
    uint64_t* status = nullptr;
    int64_t cap_tiles = 0;   // sweep-tile capacity of flags/wpre/tilecnt/tileoff
    int64_t cap_bnd = 0;     // entries of bnd
    int64_t cap_status = 0;   // entries of status (merge tiles + 1)
    int maxlev = 1;
};

// Buffers of the windowed path (one vector): the reference's append buffer,
// the radix-sort ping-pong twin and the scan / selection scratch.
struct WinBuf {
    DevSet A, T;
    int64_t* bcount = nullptr;
    int64_t* boff = nullptr;
    int64_t cap_blocks = 0;
    int64_t* bsum = nullptr;
    int64_t cap_bsum = 0;
    int64_t* hist = nullptr;
    int64_t* hoff = nullptr;
    int64_t cap_hist = 0;
    uint64_t* status = nullptr;
    int64_t cap_status = 0;
    double2* partial = nullptr;
    GenDev* gens = nullptr;
    int64_t cap_gens = 0;
    FilterDev* filt = nullptr;         // [0] strict (or NOFILTER), [1] local, [2] compaction
    int64_t* small = nullptr;          // [0] scan total, [1] merge output count
    int32_t* err = nullptr;
    uint32_t* ticket = nullptr;
    unsigned long long* orand = nullptr;  // [4 W]: OR then AND of the key words
    unsigned long long* orpart = nullptr; // [kOrandBlocks][4 W]: per-block partials
    unsigned long long* qhist = nullptr;  // [256]
    double* out = nullptr;
    DevArena alg;                      // operands, pair coefficients and results of pp_algebra
    DevArena vs;                       // variance: tile totals, head flags, carries
    uint32_t epoch = 0;
    uint64_t rng_calls = 0;
};

}  // namespace

namespace pp {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace pp

// ---------------------------------------------------------------------------
// Scratch of pp_lookup_noalloc: its staged query keys and results live in this
// per-GPU array (held by each GPU the library runs on until the process ends),
// so the call allocates nothing.  Host code uses the address kept per device
// by scratch_setup(), never the array's name (a host-side stand-in).
constexpr int64_t kLookupNoAllocMax = int64_t(1) << 16;  // queries per call
constexpr size_t kScratchBytes = (size_t)kLookupNoAllocMax * (2 * 16 + 16 + 4);  // z, x (2 words each), c (re, im), found
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

// Once per device, with that device current (pp_create): the device address
// of its copy of the scratch array.
int scratch_setup(int dev) {
    if (dev < 0 || dev >= kScratchDevices) return fail(PP_EUNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return PP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(PP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
    return PP_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// A step table (nsteps x B entries, step-major) in host memory.
struct StepSpan {
    StepDev* p = nullptr;
    size_t n = 0;
    StepDev& operator[](size_t i) const { return p[i]; }
    StepDev* data() const { return p; }
    size_t size() const { return n; }
    StepDev* begin() const { return p; }
    StepDev* end() const { return p + n; }
};

// Coefficient arrays are `double*` in the descriptors and buffers; the
// kernels see them as CT (double for Float64, pp_cplx for ComplexF64).
template <typename CT>
inline CT* cptr(double* p) {
    return reinterpret_cast<CT*>(p);
}
template <typename CT>
inline const CT* cptr(const double* p) {
    return reinterpret_cast<const CT*>(p);
}

struct Batch {
    int W = 1, B = 0, ncu = 256, dev = 0;
    int cw = 1;  // doubles per coefficient: 1 Float64, 2 ComplexF64 (re, im interleaved)
    int sweep_fold_above = 32;  // sweep workgroups take several tiles past this many tiles per CU
    hipStream_t stream = nullptr;
    std::vector<int> N;
    std::vector<InstBuf> ib;
    std::vector<int64_t> n_known;  // live count at the last host sync
    SegDesc* d_desc = nullptr;
    SegState* d_state = nullptr;
    int64_t* d_maps = nullptr;     // 4 x (B + 1) block prefixes: tile, bounds, partition, merge
    uint64_t* d_blk = nullptr;     // per-block (vector, local block, blocks) tables of the 4 maps
    std::vector<uint64_t> h_blk;
    int64_t blk_off[4] = {0, 0, 0, 0};
    int64_t blk_cap[4] = {0, 0, 0, 0};
    int64_t h_grid[4] = {1, 1, 1, 1};
    std::vector<int64_t> h_maps;
    FilterDev* d_filt = nullptr;   // [2B]: strict, local
    std::vector<int32_t> local_active;
    ChanDev* d_chan = nullptr;     // [B]
    double* d_tabs = nullptr;      // [B][3][kExpTab]
    int32_t* d_en = nullptr;       // [B]
    StepDev* d_steps = nullptr;
    int64_t cap_steps = 0;
    double2* d_partial = nullptr;
    int64_t cap_partial = 0;
    double* d_out = nullptr;
    uint64_t* d_kets = nullptr;
    int32_t* d_bad = nullptr;
    SegState* h_state = nullptr;   // pinned mirror
    StepDev* h_steps = nullptr;    // pinned step table of the current call
    size_t h_steps_cap = 0;
    // Optional HIP-event timing of every kb_merge launch (for roofline figures).
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    double t_merge = 0.0;
    int64_t merge_launches = 0;
    // Look-back tags: the base lives in device memory (so captured graphs can
    // be replayed) and is advanced by every sequence; epoch_used mirrors it.
    uint32_t* d_epoch = nullptr;
    uint64_t epoch_used = 0;
    // Captured launch sequences (hipGraph), keyed by everything a replay bakes
    // in: step count, tile shapes, grid sizes and the step-table address.
    bool graphs = false;
    struct GraphKey {
        int64_t nsteps, first;
        int tile, mtile, W;
        unsigned g0, g1, g3;
        const void* steps;
        uint64_t bounds_hash;  // which steps launch kb_bounds
        bool operator==(const GraphKey& o) const {
            return nsteps == o.nsteps && first == o.first && tile == o.tile && mtile == o.mtile && W == o.W &&
                   g0 == o.g0 && g1 == o.g1 && g3 == o.g3 && steps == o.steps && bounds_hash == o.bounds_hash;
        }
    };
    // per step of the current sequence: 1 when some vector's mask has more
    // levels than kb_flags summarises, so kb_bounds must run
    std::vector<uint8_t> step_bounds;
    std::vector<std::pair<GraphKey, hipGraphExec_t>> gcache;
    int64_t graph_launches = 0, graph_captures = 0;

    // Resident layer launches (kr_layer): one launch per evolve call with the
    // vectors' terms in LDS.  res_mode 0 off, 1 where the plan fits, 2 for
    // every call the kernel can take (tests); res_cap_limit < chunk capacity
    // forces the overflow paths (tests).
    int res_mode = 1;
    int res_cap_limit = 0;
    bool res_broken = false;        // a launch was aborted: per-step kernels from then on
    int res_kx = 0;                 // roles per XCD (CUs per XCD)
    ResCtl* d_rctl = nullptr;       // [B]
    uint32_t* d_rroles = nullptr;   // [kXcdCount * res_kx]
    uint32_t* d_rtick = nullptr;    // [kXcdCount]
    std::vector<uint32_t> h_roles;
    int64_t res_launches = 0, res_fallbacks = 0, res_aborts = 0;
    std::vector<double> res_boost;   // per vector: divisor of the starting fill after chunk overflows (this call)
    std::vector<char> res_off;       // per vector: per-step kernels for the rest of this call
    int32_t* d_rpart = nullptr;      // [B] vectors in the current resident launch
    unsigned long long* d_rwork = nullptr;  // [4] rotations, swept terms, merge outputs, rebalances of committed launches
    std::vector<int32_t> h_rpart;
    double t_res = 0.0;             // HIP-event time of resident launches (timing mode)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> res_ev;

    WinBuf win;

    ~Batch() {
        for (auto& b : ib) free_inst(b);
        free_win();
        dfree(d_rctl); dfree(d_rroles); dfree(d_rtick); dfree(d_rpart); dfree(d_rwork);
        for (auto& e : res_ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        dfree(d_desc); dfree(d_state); dfree(d_maps); dfree(d_blk); dfree(d_filt); dfree(d_chan); dfree(d_tabs);
        dfree(d_en); dfree(d_steps); dfree(d_partial); dfree(d_out); dfree(d_kets); dfree(d_bad);
        if (h_state) (void)hipHostFree(h_state);
        if (h_steps) (void)hipHostFree(h_steps);
        for (auto e : evpool) (void)hipEventDestroy(e);
        for (auto& g : gcache) (void)hipGraphExecDestroy(g.second);
        dfree(d_epoch);
        if (stream) (void)hipStreamDestroy(stream);
    }

    hipEvent_t next_event() {
        if (evused == evpool.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evpool.push_back(e);
        }
        return evpool[evused++];
    }

    // Sum the recorded kb_merge intervals (call after a synchronisation).
    void harvest_timing() {
        for (size_t i = 0; i + 1 < evused; i += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, evpool[i], evpool[i + 1]));
            t_merge += ms * 1e-3;
            merge_launches += 1;
        }
        evused = 0;
    }

    size_t kb() const { return sizeof(uint64_t) * (size_t)W; }
    size_t cb() const { return sizeof(double) * (size_t)cw; }

    // Whether `terms` more terms of this word size fit in half of the free
    // device memory (growth beyond what is needed is only taken when cheap).
    bool fits_device(int64_t terms) const {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
        return (double)terms * (double)(2 * kb() + cb()) < 0.5 * (double)fr;
    }
    // Tile shapes of the next launches (0 = choose from the live counts).
    int tile_req = 0, mtile_req = 0;
    int tile = 2048, mtile = 1024;
    void choose_tiles() {
        int64_t nmax = 0;
        for (int64_t v : n_known) nmax = std::max(nmax, v);
        const bool small = nmax < (1 << 17);
        tile = tile_req ? tile_req : (small ? 512 : 2048);
        mtile = mtile_req ? mtile_req : (small ? (W == 1 ? 1024 : 512) : (W == 1 ? 2048 : 1024));
    }
    int tm() const { return mtile; }

    void alloc_set(DevSet& s, int64_t cap) {
        s.z = dalloc<uint64_t>((size_t)cap * W);
        s.x = dalloc<uint64_t>((size_t)cap * W);
        s.c = dalloc<double>((size_t)cap * cw);
        s.cap = cap;
    }
    static void free_set(DevSet& s) {
        dfree(s.z); dfree(s.x); dfree(s.c);
        s = DevSet();
    }
    static void free_inst(InstBuf& b) {
        free_set(b.L[0]); free_set(b.L[1]); free_set(b.C);
        dfree(b.flags); dfree(b.wpre); dfree(b.tilecnt); dfree(b.tileoff); dfree(b.bnd); dfree(b.agg); dfree(b.status);
        b = InstBuf();
    }

    void init(int W_, const std::vector<int>& Ns, const std::vector<int64_t>& live_cap_v,
              const std::vector<int64_t>& app_cap_v, int cw_ = 1) {
        W = W_;
        cw = cw_;
        B = (int)Ns.size();
        N = Ns;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        ib.assign(B, InstBuf());
        n_known.assign(B, 0);
        local_active.assign(B, 0);
        d_desc = dalloc<SegDesc>(B);
        d_state = dalloc<SegState>(B);
        d_maps = dalloc<int64_t>(4 * (size_t)(B + 1));
        h_maps.assign(4 * (size_t)(B + 1), 0);
        {
            // block tables start at a few blocks per vector and grow with
            // plan()'s grids (ensure_blk)
            const int64_t first[4] = {16 * (int64_t)B + 8 * ncu, 16 * (int64_t)B + 8 * ncu, 2 * (int64_t)B + 8,
                                      16 * (int64_t)B + 8 * ncu};
            ensure_blk(first);
        }
        d_filt = dalloc<FilterDev>(2 * (size_t)B);
        d_chan = dalloc<ChanDev>(B);
        d_tabs = dalloc<double>((size_t)B * 3 * kExpTab);
        d_en = dalloc<int32_t>(B);
        d_out = dalloc<double>(2 * (size_t)B);
        d_kets = dalloc<uint64_t>((size_t)B * W);
        d_bad = dalloc<int32_t>(B);
        d_epoch = dalloc<uint32_t>(1);
        HIPCHK(hipMemsetAsync(d_epoch, 0, sizeof(uint32_t), stream));
        // resident launches need whole XCDs of equal size, each XCD's workgroups
        // at most kResMaxK per vector
        res_boost.assign((size_t)B, 1.0);
        res_off.assign((size_t)B, 0);
        h_rpart.assign((size_t)B, 0);
        if (ncu % kXcdCount == 0 && ncu / kXcdCount >= 1 && ncu / kXcdCount <= 64) {
            res_kx = ncu / kXcdCount;
            d_rctl = dalloc<ResCtl>((size_t)B);
            d_rroles = dalloc<uint32_t>((size_t)kXcdCount * res_kx);
            d_rtick = dalloc<uint32_t>(kXcdCount);
            d_rpart = dalloc<int32_t>((size_t)B);
            d_rwork = dalloc<unsigned long long>(4);
            HIPCHK(hipMemsetAsync(d_rwork, 0, 4 * sizeof(unsigned long long), stream));
            h_roles.assign((size_t)kXcdCount * res_kx, kResIdle);
        }
        HIPCHK(hipHostMalloc((void**)&h_state, sizeof(SegState) * (size_t)B, hipHostMallocDefault));
        std::memset(h_state, 0, sizeof(SegState) * (size_t)B);
        for (int b = 0; b < B; ++b) {
            h_state[b].app_cap_v = app_cap_v[b];
            h_state[b].live_cap_v = live_cap_v[b];
        }
        HIPCHK(hipMemcpyAsync(d_state, h_state, sizeof(SegState) * (size_t)B, hipMemcpyHostToDevice, stream));
        std::vector<FilterDev> f(2 * (size_t)B, make_filter_dev(nullptr));
        HIPCHK(hipMemcpyAsync(d_filt, f.data(), f.size() * sizeof(FilterDev), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int b = 0; b < B; ++b) ensure_caps(b, 4096, 4096, 2, 0);
        HIPCHK(hipStreamSynchronize(stream));
        upload_desc();
    }

    // Buffers of vector b hold `cap` live terms and `capC` children; keep > 0
    // preserves that many terms of the current live buffer `cur`.
    void ensure_caps(int b, int64_t cap, int64_t capC, int maxlev, int64_t keep, int cur = 0) {
        InstBuf& I = ib[b];
        if (I.L[0].cap < cap || I.L[1].cap < cap) {
            const int64_t nc = std::max<int64_t>(cap, 1024);
            DevSet keepset;
            if (keep > 0) {
                alloc_set(keepset, nc);
                HIPCHK(hipMemcpyAsync(keepset.z, I.L[cur].z, (size_t)keep * kb(), hipMemcpyDeviceToDevice, stream));
                HIPCHK(hipMemcpyAsync(keepset.x, I.L[cur].x, (size_t)keep * kb(), hipMemcpyDeviceToDevice, stream));
                HIPCHK(hipMemcpyAsync(keepset.c, I.L[cur].c, (size_t)keep * cb(), hipMemcpyDeviceToDevice, stream));
                HIPCHK(hipStreamSynchronize(stream));
            }
            free_set(I.L[0]);
            free_set(I.L[1]);
            if (keep > 0) I.L[cur] = keepset;
            else alloc_set(I.L[cur], nc);
            alloc_set(I.L[cur ^ 1], nc);
        }
        if (I.C.cap < capC) {
            free_set(I.C);
            alloc_set(I.C, std::max<int64_t>(capC, 1024));
        }
        const int64_t tiles = cdiv(I.L[0].cap, kTileMin) + 2;
        if (I.cap_tiles < tiles) {
            dfree(I.flags); dfree(I.wpre); dfree(I.tilecnt); dfree(I.tileoff); dfree(I.agg);
            I.cap_tiles = tiles;
            I.agg = dalloc<uint64_t>((size_t)tiles * kAggLevels);
            I.flags = dalloc<uint64_t>((size_t)tiles * (kTileMin / 64));
            I.wpre = dalloc<uint16_t>((size_t)tiles * (kTileMin / 64));
            I.tilecnt = dalloc<int32_t>(tiles);
            I.tileoff = dalloc<int64_t>(tiles + 1);
        }
        const int ml = std::min(std::max(maxlev, I.maxlev), kMaxLevels);
        if (I.cap_bnd < tiles * ml * 4) {
            dfree(I.bnd);
            I.cap_bnd = tiles * ml * 4;
            I.bnd = dalloc<int64_t>(I.cap_bnd);
        }
        I.maxlev = ml;
        const int64_t sp = cdiv(I.L[0].cap + I.C.cap, 512) + 2;
        if (I.cap_status < sp) {
            dfree(I.status);
            I.cap_status = sp;
            I.status = dalloc<uint64_t>(sp);
            // look-back words must start "not ready": freed memory may hold
            // words of another handle whose tags collide with this one's
            HIPCHK(hipMemsetAsync(I.status, 0, (size_t)sp * sizeof(uint64_t), stream));
        }
    }

    void upload_desc() {
        std::vector<SegDesc> h((size_t)B);
        for (int b = 0; b < B; ++b) {
            InstBuf& I = ib[b];
            SegDesc& d = h[b];
            std::memset(&d, 0, sizeof(d));
            for (int p = 0; p < 2; ++p) {
                d.Lz[p] = I.L[p].z;
                d.Lx[p] = I.L[p].x;
                d.Lc[p] = I.L[p].c;
            }
            d.Cz = I.C.z; d.Cx = I.C.x; d.Cc = I.C.c;
            d.flags = I.flags; d.wpre = I.wpre; d.tilecnt = I.tilecnt; d.tileoff = I.tileoff;
            d.bnd = I.bnd; d.agg = I.agg; d.status = I.status;
            d.fs = d_filt + 2 * b;
            d.fl = d_filt + 2 * b + 1;
            d.cap = std::min(I.L[0].cap, I.L[1].cap);
            d.capC = I.C.cap;
            d.maxlev = I.maxlev;
            d.local_active = local_active[b];
        }
        HIPCHK(hipMemcpyAsync(d_desc, h.data(), h.size() * sizeof(SegDesc), hipMemcpyHostToDevice, stream));
    }

    // Block prefixes per kernel family from the expected live count (blocks
    // only set parallelism: every kernel loops over the vector's real tiles).
    void plan() {
        choose_tiles();
        const int T2 = mtile;
        const int kTile = tile;
        int64_t acc[4] = {0, 0, 0, 0};
        // expected live count during the coming sequence: blocks beyond the
        // real tiles only exit (sweeps) or contend for the ticket (merge)
        auto expected = [&](int b) {
            const int64_t cap = std::min(ib[b].L[0].cap, ib[b].L[1].cap);
            return std::min<int64_t>(cap, std::max<int64_t>(n_known[b] + n_known[b] / 2, 1024));
        };
        // with many sweep tiles per CU, each sweep workgroup takes several
        // tiles in turn (4 of 512 terms, 2 of 2048): its prologue's dependent
        // loads are paid once for them
        int64_t all_tiles = 0;
        for (int b = 0; b < B; ++b) all_tiles += cdiv(expected(b), kTile);
        const int64_t fold = all_tiles > (int64_t)sweep_fold_above * ncu ? (kTile <= 512 ? 4 : 2) : 1;
        // the same for the merge grid: workgroups then take merge tiles by
        // ticket, 4 per workgroup on average
        int64_t all_mtiles = 0;
        for (int b = 0; b < B; ++b) all_mtiles += cdiv(expected(b) + expected(b) / 2, T2);
        const int64_t mfold = all_mtiles > (int64_t)sweep_fold_above * ncu ? 4 : 1;
        for (int b = 0; b <= B; ++b) {
            for (int k = 0; k < 4; ++k) h_maps[(size_t)k * (B + 1) + b] = acc[k];
            if (b == B) break;
            const int64_t nest = expected(b);
            const int64_t tiles = cdiv(nest, kTile);
            acc[0] += std::max<int64_t>(1, std::min<int64_t>(cdiv(tiles, fold), (int64_t)ncu * 8));
            acc[1] += std::max<int64_t>(1, std::min<int64_t>(cdiv(2 * tiles, kBlock / kWave), (int64_t)ncu * 4));
            acc[2] += 1;
            acc[3] += std::max<int64_t>(1, std::min<int64_t>(cdiv(nest + nest / 2, T2 * mfold), (int64_t)ncu * 4));
        }
        // XCD placement: vectors by decreasing size, each to the XCD row with
        // the least tile blocks so far; a map's grid is kXcds times its
        // longest row (blocks beyond a row's end have no work)
        std::vector<int> order((size_t)B), row((size_t)B);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return n_known[a] > n_known[c]; });
        int64_t row_max = 0;  // sweep blocks of the fullest XCD row
        {
            int64_t load[kXcds] = {0};
            for (int b : order) {
                int x = 0;
                for (int y = 1; y < kXcds; ++y)
                    if (load[y] < load[x]) x = y;
                row[(size_t)b] = x;
                load[x] += h_maps[(size_t)0 * (B + 1) + b + 1] - h_maps[(size_t)0 * (B + 1) + b];
            }
            for (int x = 0; x < kXcds; ++x) row_max = std::max(row_max, load[x]);
        }
        auto round = [&](int64_t g) {
            // with captured graphs each grid is rounded up to a power of
            // two, so that a graph, whose grids are fixed, is replayed by
            // most calls; the block table itself is data read at run time
            g = std::max<int64_t>(1, g);
            if (!graphs) return g;
            int64_t r = 1;
            while (r < g) r <<= 1;
            return r;
        };
        // rows pay off where kernels are latency-bound (L2 hits between the
        // kernels of a step); a call with more tiles than a few per CU is
        // bandwidth-bound and wants every XCD on every vector.  So does a
        // small batch whose fullest row holds more than a few tiles per CU of
        // its XCD for each of its vectors: one large vector alone would run
        // on one XCD (measured on the 36-qubit chain one vector at a time:
        // 45 evals/s with rows, 59 without)
        const int64_t row_ok = 16 * std::max<int64_t>(1, ncu / kXcds) * (int64_t)B;
        const bool rows = acc[0] <= 16 * (int64_t)ncu && row_max <= row_ok;
        std::vector<int64_t> st((size_t)4 * B);
        int64_t grids[4];
        for (int k = 0; k < 4; ++k) {
            const int64_t* pre = h_maps.data() + (size_t)k * (B + 1);
            int64_t start[kXcds] = {0};
            for (int b : order) {
                st[(size_t)k * B + b] = start[row[(size_t)b]];
                start[row[(size_t)b]] += pre[b + 1] - pre[b];
            }
            int64_t len = 0;
            for (int x = 0; x < kXcds; ++x) len = std::max(len, start[x]);
            grids[k] = rows ? round((int64_t)kXcds * len) : round(pre[B]);  // contiguous layout: pre[B] entries
        }
        ensure_blk(grids);
        for (int k = 0; k < 4; ++k) {
            const int64_t* pre = h_maps.data() + (size_t)k * (B + 1);
            const int64_t grid = grids[k];
            h_grid[k] = grid;
            uint64_t* tab = h_blk.data() + blk_off[k];
            std::fill(tab, tab + grid, (uint64_t)kBlkNone);
            for (int b = 0; b < B; ++b) {
                const int64_t nbk = pre[b + 1] - pre[b];
                for (int64_t lb = 0; lb < nbk; ++lb) {
                    const int64_t g = rows ? kXcds * (st[(size_t)k * B + b] + lb) + row[(size_t)b] : pre[b] + lb;
                    tab[(size_t)g] =
                        (uint64_t)b | ((uint64_t)lb << kBlkVecBits) | ((uint64_t)nbk << (kBlkVecBits + kBlkLocBits));
                }
            }
            HIPCHK(hipMemcpyAsync(d_blk + blk_off[k], tab, (size_t)grid * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
        }
        HIPCHK(hipMemcpyAsync(d_maps, h_maps.data(), h_maps.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
        if (cap_partial < acc[0]) {
            dfree(d_partial);
            cap_partial = acc[0];
            d_partial = dalloc<double2>(cap_partial);
        }
    }

    // Block tables of at least need[k] entries per map.  Growing moves the
    // tables: in-flight kernels finish first, and the captured graphs, which
    // hold the old address, are dropped (they are captured again on use).
    void ensure_blk(const int64_t need[4]) {
        bool fits = d_blk != nullptr;
        for (int k = 0; k < 4; ++k) fits = fits && need[k] <= blk_cap[k];
        if (fits) return;
        auto p2 = [](int64_t v) {
            int64_t r = 1;
            while (r < v) r <<= 1;
            return r;
        };
        int64_t cap[4], off = 0;
        for (int k = 0; k < 4; ++k) {
            cap[k] = std::max(blk_cap[k], 2 * p2(std::max<int64_t>(need[k], 1)));
            off += cap[k];
        }
        if (stream) HIPCHK(hipStreamSynchronize(stream));
        for (auto& g : gcache) (void)hipGraphExecDestroy(g.second);
        gcache.clear();
        dfree(d_blk);
        d_blk = nullptr;
        d_blk = dalloc<uint64_t>((size_t)off);
        h_blk.assign((size_t)off, 0);
        off = 0;
        for (int k = 0; k < 4; ++k) {
            blk_off[k] = off;
            blk_cap[k] = cap[k];
            off += cap[k];
        }
    }

    BlockMap map(int k) const { return BlockMap{d_maps + (size_t)k * (B + 1), d_blk + blk_off[k], B}; }
    unsigned blocks(int k) const { return (unsigned)h_grid[k]; }

    std::vector<pp_filter> last_filters;  // what d_filt holds (strict, local per vector)

    void set_filters(const pp_filter* strict, const pp_filter* local) {
        static const pp_filter kNo = {INT64_MAX, INT64_MAX, INT64_MAX, -1.0, 0.0, -1.0, 0.0, -1.0};
        std::vector<pp_filter> want(2 * (size_t)B);
        for (int b = 0; b < B; ++b) {
            want[2 * b] = strict ? strict[b] : kNo;
            want[2 * b + 1] = local ? local[b] : kNo;
        }
        if (want.size() == last_filters.size() &&
            std::memcmp(want.data(), last_filters.data(), want.size() * sizeof(pp_filter)) == 0)
            return;
        last_filters = want;
        std::vector<FilterDev> f(2 * (size_t)B);
        for (int b = 0; b < B; ++b) {
            f[2 * b] = make_filter_dev(strict ? strict + b : nullptr);
            f[2 * b + 1] = make_filter_dev(local ? local + b : nullptr);
            local_active[b] = filter_active(local ? local + b : nullptr) ? 1 : 0;
        }
        HIPCHK(hipMemcpyAsync(d_filt, f.data(), f.size() * sizeof(FilterDev), hipMemcpyHostToDevice, stream));
    }

    void ensure_steps(int64_t count) {
        if (cap_steps >= count) return;
        dfree(d_steps);
        cap_steps = std::max<int64_t>(count, 2 * cap_steps);
        d_steps = dalloc<StepDev>(cap_steps);
    }

    // Reserve `count` look-back tags for the next sequence.  Before the 22-bit
    // tag would wrap, every look-back array is cleared and the base reset, so
    // no word of an older launch can carry a tag in use.
    void reserve_epochs(uint64_t count) {
        if (epoch_used + count + 1 > (uint64_t)kEpochMask) {
            for (auto& I : ib)
                if (I.status) HIPCHK(hipMemsetAsync(I.status, 0, (size_t)I.cap_status * sizeof(uint64_t), stream));
            static const uint32_t one = 1;
            HIPCHK(hipMemcpyAsync(d_epoch, &one, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
            epoch_used = 1;
        }
        epoch_used += count;
    }

    // kb_scan runs when kb_bounds needs the tile offsets or a vector may
    // hold more than kEmitSelfTiles sweep tiles; otherwise kb_emit sums the
    // tile counts itself (one kernel less per rotation).
    // kb_bounds also runs when a vector may hold more than kSummMaxTiles
    // sweep tiles in this call (those vectors then use its table)
    bool big_call() const {
        for (int b = 0; b < B; ++b)
            if (cdiv(n_known[b] + n_known[b] / 2, (int64_t)tile) > kSummMaxTiles) return true;
        return false;
    }

    bool scan_needed(bool bounds) const {
        if (bounds) return true;
        for (const InstBuf& I : ib)
            if (cdiv(std::min(I.L[0].cap, I.L[1].cap), (int64_t)tile) > kEmitSelfTiles) return true;
        return false;
    }

    template <int Wt, typename CT, int T, int TM>
    void launch_step(const StepDev* dsteps, int64_t step_index, uint32_t ep_off, bool bounds, bool scan) {
        hipLaunchKernelGGL((kb_flags<Wt, T, CT>), dim3(blocks(0)), dim3(kBlock), 0, stream, map(0), d_desc, d_state, dsteps,
                           bounds ? 1 : 0);
        if (scan) hipLaunchKernelGGL(kb_scan<T>, dim3(B), dim3(1024), 0, stream, d_desc, d_state, dsteps, step_index);
        if (bounds)
            hipLaunchKernelGGL((kb_bounds<Wt, T>), dim3(blocks(1)), dim3(kBlock), 0, stream, map(1), d_desc, d_state,
                               dsteps);
        if (bounds)
            hipLaunchKernelGGL((kb_emit<Wt, T, true, CT>), dim3(blocks(0)), dim3(kBlock), 0, stream, map(0), d_desc, d_state,
                               dsteps, 0, step_index);
        else
            hipLaunchKernelGGL((kb_emit<Wt, T, false, CT>), dim3(blocks(0)), dim3(kBlock), 0, stream, map(0), d_desc,
                               d_state, dsteps, scan ? 0 : 1, step_index);
        if (timing) {
            // events recorded by the launch itself at the kernel's start and
            // end (not around it), so the interval is the kernel's duration
            hipEvent_t e0 = next_event();
            hipEvent_t e1 = next_event();
            hipExtLaunchKernelGGL((kb_merge<Wt, TM, CT>), dim3(blocks(3)), dim3(merge_threads(TM)), 0, stream, e0, e1, 0, map(3),
                                  d_desc, d_state, dsteps, 0, step_index, (const uint32_t*)d_epoch, ep_off);
        } else {
            hipLaunchKernelGGL((kb_merge<Wt, TM, CT>), dim3(blocks(3)), dim3(merge_threads(TM)), 0, stream, map(3), d_desc, d_state,
                               dsteps, 0, step_index, (const uint32_t*)d_epoch, ep_off);
        }
    }

    // Compaction launch (the step table is not read in this mode but must be
    // a valid pointer: the caller has sized it to at least B entries).
    template <int Wt, typename CT, int TM>
    void launch_compact(uint32_t ep_off) {
        hipLaunchKernelGGL((kb_merge<Wt, TM, CT>), dim3(blocks(3)), dim3(merge_threads(TM)), 0, stream, map(3), d_desc, d_state, d_steps,
                           1, (int64_t)-1, (const uint32_t*)d_epoch, ep_off);
    }

    // Enqueue a sequence: optional precheck + compaction, steps [first,
    // nsteps) from the table at dsteps, then the tag advance.  In graph mode
    // the launches are captured once per key and replayed.
    template <int Wt, typename CT>
    void enqueue_sequence(bool with_precheck, const StepDev* dsteps, int64_t first, int64_t nsteps, bool use_graph) {
        const uint32_t count = (uint32_t)(nsteps - first + (with_precheck ? 1 : 0));
        reserve_epochs(count);
        auto body = [&]() {
            if (with_precheck)
                hipLaunchKernelGGL(kb_precheck, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, stream, d_state,
                                   (const int32_t*)d_en, B);
            with_tiles<Wt>([&](auto T, auto TM) {
                uint32_t off = 0;
                if (with_precheck) launch_compact<Wt, CT, decltype(TM)::value>(off++);
                const bool big = big_call();
                for (int64_t s = first; s < nsteps; ++s) {
                    const bool bnd = step_bounds[(size_t)s] != 0 || big;
                    launch_step<Wt, CT, decltype(T)::value, decltype(TM)::value>(dsteps + (s - first) * B, s, off++, bnd,
                                                                             scan_needed(bnd));
                }
            });
            hipLaunchKernelGGL(kb_epoch_advance, dim3(1), dim3(1), 0, stream, d_epoch, count);
        };
        if (!use_graph || timing) {
            body();
            HIPCHK(hipGetLastError());
            return;
        }
        uint64_t bh = 1469598103934665603ull;
        for (int64_t s = first; s < nsteps; ++s) bh = (bh ^ step_bounds[(size_t)s]) * 1099511628211ull;
        bh = (bh ^ (scan_needed(false) ? 2u : 0u) ^ (big_call() ? 4u : 0u)) * 1099511628211ull;
        GraphKey key{nsteps, with_precheck ? -1 - first : first, tile, mtile, Wt, blocks(0), blocks(1), blocks(3),
                     (const void*)dsteps, bh};
        launch_captured(key, body);
    }

    // Replay the graph captured for `key`, or capture body() into one.
    template <typename F>
    void launch_captured(const GraphKey& key, F&& body) {
        for (auto& g : gcache)
            if (g.first == key) {
                HIPCHK(hipGraphLaunch(g.second, stream));
                graph_launches += 1;
                return;
            }
        if (gcache.size() >= 16) {
            HIPCHK(hipStreamSynchronize(stream));
            for (auto& g : gcache) (void)hipGraphExecDestroy(g.second);
            gcache.clear();
        }
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        body();
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        HIPCHK(le);
        HIPCHK(ce);
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(ie);
        gcache.emplace_back(key, exec);
        graph_captures += 1;
        HIPCHK(hipGraphLaunch(exec, stream));
        graph_launches += 1;
    }

    // Run f<T, TM>() with the current tile shapes as template arguments.
    template <int Wt, typename F>
    void with_tiles(F&& f) {
        using I = std::integral_constant<int, 512>;
        using J = std::integral_constant<int, 1024>;
        using K = std::integral_constant<int, 2048>;
        if (tile == 512) {
            if (mtile == 512) f(I(), I());
            else if (mtile == 1024) f(I(), J());
            else {
                if constexpr (Wt == 1) f(I(), K());
                else throw PPError(PP_EARG, "merge tile 2048 needs N <= 64");
            }
        } else {
            if (mtile == 512) f(K(), I());
            else if (mtile == 1024) f(K(), J());
            else {
                if constexpr (Wt == 1) f(K(), K());
                else throw PPError(PP_EARG, "merge tile 2048 needs N <= 64");
            }
        }
    }

    // ----------------------------------------------------------------------
    // Resident layer launches (kr_layer).

    int res_cap() const {
        const int full = W == 1 ? ResGeom<1>::cap : (cw == 1 ? ResGeom<2, double>::cap : ResGeom<2, pp_cplx>::cap);
        return res_cap_limit > 0 && res_cap_limit < full ? res_cap_limit : full;
    }

    // Roles of one resident launch: the vectors with active steps go to XCDs
    // by decreasing size, each with K workgroups whose starting chunks fill at
    // most 25% (else 40%, else 60%) of a chunk, divided by the vector's overflow boost
    // (a vector whose chunks overflowed resumes with more workgroups).
    // false: the call runs on the per-step kernels (auto mode also declines
    // calls of few steps and vectors that do not fit one XCD at 90% fill).
    bool res_plan(const StepSpan& steps, int64_t nsteps) {
        if (res_mode == 0 || res_broken || res_kx == 0) return false;
        if (res_mode == 1 && nsteps < 8) return false;
        std::vector<char> part((size_t)B, 0);
        for (int64_t s = 0; s < nsteps; ++s)
            for (int b = 0; b < B; ++b) {
                const StepDev& sd = steps[(size_t)(s * B + b)];
                if (!sd.active) continue;
                if (sd.g.nlev > kResLevels || res_off[(size_t)b]) return false;
                part[(size_t)b] = 1;
            }
        const int cap = res_cap();
        const int kmax = std::min(res_kx, kResMaxK);
        int nparts = 0;
        for (int b = 0; b < B; ++b) nparts += part[(size_t)b];
        if (nparts > kXcdCount * res_kx) return false;
        // a vector held to fewer workgroups than it asked for (the XCD's
        // limit) still runs resident with its chunks up to 90% full: the
        // largest vectors of a sweep grow little within a call
        const int64_t tight = std::max<int64_t>(1, (int64_t)(0.9 * cap));
        std::vector<int> K((size_t)B, 0);
        std::vector<int> xcd((size_t)B, 0);
        auto fits = [&](int b) { return n_known[b] <= (int64_t)K[b] * tight; };
        std::vector<int> order;
        for (int b = 0; b < B; ++b)
            if (part[(size_t)b]) order.push_back(b);
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return n_known[a] > n_known[c]; });
        // one assignment at a starting fill of `frac` of a chunk; false when
        // the XCDs run out of workgroups
        auto assign = [&](double frac) {
            const int64_t base = std::max<int64_t>(1, (int64_t)(frac * cap));
            for (int b = 0; b < B; ++b) {
                K[(size_t)b] = 0;
                if (!part[(size_t)b]) continue;
                const int64_t fill = std::max<int64_t>(1, (int64_t)(base / res_boost[(size_t)b]));
                const int64_t want = cdiv(std::max<int64_t>(n_known[b], 1), fill);
                K[b] = (int)std::min<int64_t>(want, kmax);
                if (res_mode == 1 && !fits(b)) return false;
            }
            int freeslots[kXcdCount];
            int64_t load[kXcdCount];
            for (int x = 0; x < kXcdCount; ++x) {
                freeslots[x] = res_kx;
                load[x] = 0;
            }
            for (int b : order) {
                int x = 0;
                for (int y = 1; y < kXcdCount; ++y)
                    if (freeslots[y] > freeslots[x] || (freeslots[y] == freeslots[x] && load[y] < load[x])) x = y;
                if (freeslots[x] < 1) return false;
                if (K[b] > freeslots[x]) {
                    K[b] = freeslots[x];
                    if (res_mode == 1 && !fits(b)) return false;
                }
                xcd[(size_t)b] = x;
                freeslots[x] -= K[b];
                load[x] += n_known[b];
            }
            return true;
        };
        // Small starting chunks (a quarter full) spread a vector over more
        // CUs and leave room to grow: on the 121-qubit sweep 25-30% starting
        // fill ran 7-11% faster than 60%, with fewer vectors handed back to
        // the per-step kernels.  Where the XCDs cannot hold every vector at
        // that fill, 40% and then 60% are tried before declining (declining
        // sends the whole call to the per-step kernels).
        if (!assign(0.25) && !assign(0.4) && !assign(0.6)) return false;
        std::fill(h_roles.begin(), h_roles.end(), kResIdle);
        int next[kXcdCount] = {0};
        for (int b = 0; b < B; ++b)
            for (int k = 0; k < K[b]; ++k)
                h_roles[(size_t)xcd[b] * res_kx + next[xcd[b]]++] =
                    (uint32_t)b | ((uint32_t)k << 20) | ((uint32_t)K[b] << 26);
        for (int b = 0; b < B; ++b) h_rpart[(size_t)b] = part[(size_t)b];
        HIPCHK(hipMemcpyAsync(d_rroles, h_roles.data(), h_roles.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              stream));
        HIPCHK(hipMemcpyAsync(d_rpart, h_rpart.data(), h_rpart.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              stream));
        return true;
    }

    // (Sequence start: the precheck and compaction, then) one resident launch
    // for steps [first, nsteps) of the planned vectors and the commit of its
    // results.
    template <int Wt, typename CT>
    void enqueue_resident(const StepDev* dsteps, int64_t first, int64_t nsteps, bool with_precheck) {
        if (with_precheck) reserve_epochs(1);
        ResArgs a;
        a.roles = d_rroles;
        a.tickets = d_rtick;
        a.kx = res_kx;
        a.B = B;
        a.D = d_desc;
        a.S = d_state;
        a.steps = dsteps;
        a.first = first;
        a.nsteps = nsteps;
        a.ctl = d_rctl;
        a.cap = res_cap();
        const dim3 grid((unsigned)(kXcdCount * res_kx)), block(kResThreads);
        auto body = [&](bool timed) {
            if (with_precheck) {
                hipLaunchKernelGGL(kb_precheck, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, stream, d_state,
                                   (const int32_t*)d_en, B);
                with_tiles<Wt>([&](auto, auto TM) { launch_compact<Wt, CT, decltype(TM)::value>(0); });
                hipLaunchKernelGGL(kb_epoch_advance, dim3(1), dim3(1), 0, stream, d_epoch, 1u);
            }
            hipLaunchKernelGGL(kr_reset, dim3((unsigned)cdiv(std::max(B, kXcdCount), 256)), dim3(256), 0, stream,
                               d_rctl, d_rtick, B);
            if (timed) {
                hipEvent_t e0, e1;
                HIPCHK(hipEventCreate(&e0));
                HIPCHK(hipEventCreate(&e1));
                res_ev.emplace_back(e0, e1);
                hipExtLaunchKernelGGL((kr_layer<Wt, CT>), grid, block, 0, stream, e0, e1, 0, a);
            } else {
                hipLaunchKernelGGL((kr_layer<Wt, CT>), grid, block, 0, stream, a);
            }
            hipLaunchKernelGGL(kr_commit, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, stream, (const ResCtl*)d_rctl,
                               (const int32_t*)d_rpart, d_state, B, first, d_rwork);
        };
        res_launches += 1;
        if (!graphs || timing) {
            body(timing);
            HIPCHK(hipGetLastError());
            return;
        }
        // the role table is data (uploaded before the launch), so one graph
        // serves every call of the same step count and grids
        GraphKey key{nsteps, -(int64_t(1) << 40) - first, tile, mtile, Wt, blocks(3), (unsigned)res_cap(),
                     (unsigned)res_kx, (const void*)dsteps, 0x7265736964656e74ull ^ (with_precheck ? 1u : 0u)};
        launch_captured(key, [&]() { body(false); });
    }

    void harvest_res_timing() {
        for (auto& e : res_ev) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
            t_res += ms * 1e-3;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        res_ev.clear();
    }

    void sync_state() {
        hipLaunchKernelGGL(kb_resolve, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, stream, d_state, B);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h_state, d_state, sizeof(SegState) * (size_t)B, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (timing) harvest_timing();
        if (!res_ev.empty()) harvest_res_timing();
        for (int b = 0; b < B; ++b) {
            if (h_state[b].err)
                throw PPError(PP_EHIP, "internal consistency check failed in a kernel (vector " + std::to_string(b) +
                                           ", bits " + std::to_string(h_state[b].err) + ")");
            n_known[b] = h_state[b].n[h_state[b].cur];
        }
    }

    void put_state() {
        HIPCHK(hipMemcpyAsync(d_state, h_state, sizeof(SegState) * (size_t)B, hipMemcpyHostToDevice, stream));
    }

    // One sequence: for every vector b, the rotations steps[s * B + b] with
    // active != 0, in step order (one evolve! call per vector), preceded by
    // the reference's capacity precheck.  Stalled vectors are regrown and
    // replayed until every vector has finished the sequence.
    template <int Wt, typename CT>
    void run_sequence(const StepSpan& steps, int64_t nsteps, const std::vector<int32_t>& en) {
        std::fill(res_off.begin(), res_off.end(), 0);
        std::fill(res_boost.begin(), res_boost.end(), 1.0);  // boosts last one call
        res_ovf.assign((size_t)B, 0);
        run_sequence_body<Wt, CT>(steps, nsteps, en);
    }
    std::vector<char> res_ovf;  // per vector: a chunk overflowed in this call

    // The pinned step table for n entries (its previous contents kept while
    // it is large enough; an upload from it has finished by the next call,
    // which starts after the stream was synchronised).
    StepSpan host_steps(size_t n) {
        if (h_steps_cap < n) {
            HIPCHK(hipStreamSynchronize(stream));
            if (h_steps) HIPCHK(hipHostFree(h_steps));
            h_steps = nullptr;
            h_steps_cap = 0;
            const size_t cap = std::max<size_t>(n, 1024);
            HIPCHK(hipHostMalloc((void**)&h_steps, cap * sizeof(StepDev), hipHostMallocDefault));
            std::memset(h_steps, 0, cap * sizeof(StepDev));
            h_steps_cap = cap;
        }
        return StepSpan{h_steps, n};
    }

    template <int Wt, typename CT>
    void run_sequence_body(const StepSpan& steps, int64_t nsteps, const std::vector<int32_t>& en) {
        int maxlev = 1;
        for (const StepDev& s : steps) maxlev = std::max(maxlev, (int)s.g.nlev);
        step_bounds.assign((size_t)nsteps, 0);
        for (int64_t s = 0; s < nsteps; ++s)
            for (int b = 0; b < B; ++b) {
                const StepDev& sd = steps[(size_t)(s * B + b)];
                if (sd.active && sd.g.nlev > kAggLevels) step_bounds[(size_t)s] = 1;
            }
        for (int b = 0; b < B; ++b) {
            const int64_t cap = std::min(ib[b].L[0].cap, ib[b].L[1].cap);
            const int64_t n = n_known[b];
            // headroom for growth within the call: a vector with less than 2x
            // room is grown to 4x (children to 2x) when device memory allows,
            // so that a stall and replay is the exception
            if (en[b] && cap < 2 * n && fits_device(3 * 4 * n)) {
                ensure_caps(b, 4 * n, std::max<int64_t>(ib[b].C.cap, 2 * n), maxlev, n, h_state[b].cur);
            } else if (ib[b].maxlev < maxlev) {
                ensure_caps(b, cap, ib[b].C.cap, maxlev, 0, h_state[b].cur);
            }
        }
        upload_desc();
        plan();
        ensure_steps(nsteps * B);
        HIPCHK(hipMemcpyAsync(d_steps, steps.data(), (size_t)(nsteps * B) * sizeof(StepDev), hipMemcpyHostToDevice,
                              stream));
        HIPCHK(hipMemcpyAsync(d_en, en.data(), (size_t)B * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        if (res_plan(steps, nsteps)) enqueue_resident<Wt, CT>(d_steps, 0, nsteps, true);
        else enqueue_sequence<Wt, CT>(true, d_steps, 0, nsteps, graphs);
        sync_state();
        std::vector<int64_t> last_stall((size_t)B, -1);
        for (int round = 0;; ++round) {
            std::vector<int> stalled;
            for (int b = 0; b < B; ++b)
                if (h_state[b].stalled) stalled.push_back(b);
            if (stalled.empty()) break;
            if (round > 64) throw PPError(PP_EOVERFLOW, "vector kept stalling after regrowth");
            int64_t first = nsteps;
            for (int b : stalled) {
                SegState& st = h_state[b];
                if (st.stalled >= 3) {
                    // a resident launch stopped this vector: a chunk outgrew
                    // its LDS (3; the vector resumes in a resident launch with
                    // twice the workgroups, or on the per-step kernels when it
                    // overflows twice at one step) or the launch was aborted
                    // (4; per-step kernels from then on)
                    res_fallbacks += 1;
                    if (st.stalled == 4) {
                        res_aborts += 1;
                        res_broken = true;
                    } else {
                        if (res_ovf[(size_t)b] && last_stall[(size_t)b] == st.stall_step) res_off[(size_t)b] = 1;
                        res_ovf[(size_t)b] = 1;
                        last_stall[(size_t)b] = st.stall_step;
                        res_boost[(size_t)b] = std::min(32.0, res_boost[(size_t)b] * 2.0);
                    }
                    first = std::min(first, st.stall_step);
                    continue;
                }
                const int64_t cap = std::min(ib[b].L[0].cap, ib[b].L[1].cap);
                const int64_t n = st.n[st.cur];
                // a vector that outgrew its buffers mid-sequence usually keeps
                // growing: regrow 4x (children and live together) so that
                // the replay rarely stalls again
                const int64_t big = 4 * st.need + 1024;
                const int64_t need = fits_device(3 * big) ? big : st.need + st.need / 2 + 1024;
                if (st.stalled == 1) ensure_caps(b, std::max(cap, n + need), need, maxlev, n, st.cur);
                else ensure_caps(b, need, std::max<int64_t>(ib[b].C.cap, need / 2), maxlev, n, st.cur);
                first = std::min(first, st.stall_step);
            }
            // replay table: only the stalled vectors, each from its own stall
            // step, whose capacity precheck already ran
            std::vector<StepDev> rs((size_t)((nsteps - first) * B));
            std::vector<char> is_st(B, 0);
            for (int b : stalled) is_st[b] = 1;
            for (int64_t s = first; s < nsteps; ++s)
                for (int b = 0; b < B; ++b) {
                    StepDev sd = steps[(size_t)(s * B + b)];
                    const bool on = is_st[b] && s >= h_state[b].stall_step && sd.active;
                    sd.active = on ? 1 : 0;
                    if (s == h_state[b].stall_step) sd.precheck = 0;
                    rs[(size_t)((s - first) * B + b)] = sd;
                }
            for (int b : stalled) {
                h_state[b].stalled = 0;
                h_state[b].flip = 0;
            }
            put_state();
            upload_desc();
            plan();
            HIPCHK(hipMemcpyAsync(d_steps, rs.data(), rs.size() * sizeof(StepDev), hipMemcpyHostToDevice, stream));
            if (res_plan(StepSpan{rs.data(), rs.size()}, nsteps - first)) enqueue_resident<Wt, CT>(d_steps, first, nsteps, false);
            else enqueue_sequence<Wt, CT>(false, d_steps, first, nsteps, false);
            sync_state();
        }
    }

    template <int Wt, typename CT>
    void channel(const std::vector<ChanDev>& ch, const std::vector<double>& tabs, const std::vector<int32_t>& en) {
        HIPCHK(hipMemcpyAsync(d_tabs, tabs.data(), tabs.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_chan, ch.data(), ch.size() * sizeof(ChanDev), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_en, en.data(), en.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        plan();
        hipLaunchKernelGGL((kb_channel<Wt, CT>), dim3(blocks(0)), dim3(kBlock), 0, stream, d_desc, d_state, d_chan, d_en, map(0));
        HIPCHK(hipGetLastError());
    }

    // Compaction of the selected vectors under their strict filter (clips).
    template <int Wt, typename CT>
    void compact(const std::vector<int32_t>& sel) {
        sync_state();
        for (int b = 0; b < B; ++b) {
            h_state[b].compact = sel[b] ? 1 : 0;
            h_state[b].ticket = 0u;
            h_state[b].nstep = h_state[b].n[h_state[b].cur];  // the merge's live count
        }
        put_state();
        upload_desc();
        plan();
        ensure_steps(B);
        reserve_epochs(1);
        with_tiles<Wt>([&](auto, auto TM) { launch_compact<Wt, CT, decltype(TM)::value>(0); });
        hipLaunchKernelGGL(kb_epoch_advance, dim3(1), dim3(1), 0, stream, d_epoch, 1u);
        HIPCHK(hipGetLastError());
        sync_state();
    }

    // out[b * cw + comp]: the real part (and for ComplexF64 the imaginary part)
    template <int Wt, typename CT>
    void expectation(const uint64_t* kets, double* out) {
        HIPCHK(hipMemcpyAsync(d_kets, kets, (size_t)B * W * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
        plan();
        for (int comp = 0; comp < cw; ++comp) {
            hipLaunchKernelGGL((kb_expect<Wt, CT>), dim3(blocks(0)), dim3(kBlock), 0, stream, d_desc, d_state, d_kets,
                               map(0), d_partial, comp);
            hipLaunchKernelGGL(kb_expect_final, dim3(B), dim3(kBlock), 0, stream, d_partial, map(0), d_out, cw, comp);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, d_out, (size_t)B * cb(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    template <int Wt, typename CT>
    void check(int32_t* ok) {
        HIPCHK(hipMemsetAsync(d_bad, 0, (size_t)B * sizeof(int32_t), stream));
        plan();
        hipLaunchKernelGGL(kb_check<Wt>, dim3(blocks(0)), dim3(kBlock), 0, stream, d_desc, d_state, map(0), d_bad);
        HIPCHK(hipGetLastError());
        std::vector<int32_t> bad((size_t)B);
        HIPCHK(hipMemcpyAsync(bad.data(), d_bad, (size_t)B * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int b = 0; b < B; ++b) ok[b] = bad[b] ? 0 : 1;
    }

    // ----------------------------------------------------------------------
    // Windowed path (window > 1, non-compilable boundary strategies,
    // EnergyCorrection): one vector, host-driven, spv_evolve.jl:312-387.

    void free_win() {
        WinBuf& w = win;
        free_set(w.A); free_set(w.T);
        dfree(w.bcount); dfree(w.boff); dfree(w.bsum); dfree(w.hist); dfree(w.hoff); dfree(w.status);
        dfree(w.partial); dfree(w.gens); dfree(w.filt); dfree(w.small); dfree(w.err); dfree(w.ticket);
        dfree(w.orand); dfree(w.orpart); dfree(w.qhist); dfree(w.out);
        w.alg.release();
        w.vs.release();
        w = WinBuf();
    }

    void win_init() {
        WinBuf& w = win;
        if (w.filt) return;
        w.filt = dalloc<FilterDev>(3);
        w.small = dalloc<int64_t>(2);
        w.err = dalloc<int32_t>(1);
        w.ticket = dalloc<uint32_t>(1);
        w.orand = dalloc<unsigned long long>(4 * (size_t)W);
        w.orpart = dalloc<unsigned long long>((size_t)kOrandBlocks * 4 * W);
        w.qhist = dalloc<unsigned long long>(kRadix);
        w.out = dalloc<double>(3);
        w.partial = dalloc<double2>(1024);
        HIPCHK(hipMemsetAsync(w.err, 0, sizeof(int32_t), stream));
    }

    // Grow a set to hold `need` terms, keeping the first `keep`.
    void grow_set(DevSet& s, int64_t need, int64_t keep) {
        if (s.cap >= need) return;
        DevSet n2;
        alloc_set(n2, std::max<int64_t>(need + need / 2, 1024));
        if (keep > 0) {
            HIPCHK(hipMemcpyAsync(n2.z, s.z, (size_t)keep * kb(), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(n2.x, s.x, (size_t)keep * kb(), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(n2.c, s.c, (size_t)keep * cb(), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        free_set(s);
        s = n2;
    }

    void win_blocks(int64_t nb) {
        WinBuf& w = win;
        if (w.cap_blocks >= nb) return;
        dfree(w.bcount); dfree(w.boff);
        w.cap_blocks = std::max<int64_t>(nb, 2 * w.cap_blocks);
        w.bcount = dalloc<int64_t>(w.cap_blocks);
        w.boff = dalloc<int64_t>(w.cap_blocks);
    }

    // Exclusive scan of count int64 values; the sum lands in *total (device).
    void scan_i64(const int64_t* in, int64_t count, int64_t* out, int64_t* total) {
        WinBuf& w = win;
        if (count == 0) {
            HIPCHK(hipMemsetAsync(total, 0, sizeof(int64_t), stream));
            return;
        }
        const int64_t nb = cdiv(count, (int64_t)kBlock * 8);
        if (w.cap_bsum < nb) {
            dfree(w.bsum);
            w.cap_bsum = std::max<int64_t>(nb, 2 * w.cap_bsum);
            w.bsum = dalloc<int64_t>(w.cap_bsum);
        }
        hipLaunchKernelGGL(kx_scan_blocks, dim3((unsigned)nb), dim3(kBlock), 0, stream, in, count, out, w.bsum);
        hipLaunchKernelGGL(kx_scan_sums, dim3(1), dim3(1024), 0, stream, w.bsum, nb, total);
        hipLaunchKernelGGL(kx_scan_add, dim3((unsigned)nb), dim3(kBlock), 0, stream, out, count, w.bsum);
    }

    int64_t read_small(int i) {
        int64_t v = 0;
        HIPCHK(hipMemcpyAsync(&v, win.small + i, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return v;
    }

    void check_win_err() {
        int32_t e = 0;
        HIPCHK(hipMemcpyAsync(&e, win.err, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (e) throw PPError(PP_EHIP, "internal consistency check failed in a windowed-path kernel (bits " +
                                          std::to_string(e) + ")");
    }

    // One rotation over [live (n) ; appends (an)]; returns the children appended.
    template <int Wt, typename CT>
    int64_t win_rotate(const GenDev* g, int64_t n, int64_t an, int cur) {
        WinBuf& w = win;
        const int64_t tot = n + an;
        if (tot == 0) return 0;
        grow_set(w.A, an + tot, an);
        const int64_t nb = cdiv(tot, kRB);
        win_blocks(nb);
        const DevSet& Ls = ib[0].L[cur];
        hipLaunchKernelGGL((kw_rotate<Wt, CT>), dim3((unsigned)nb), dim3(kBlock), 0, stream, Ls.z, Ls.x, cptr<CT>(Ls.c), n, w.A.z,
                           w.A.x, cptr<CT>(w.A.c), an, g, w.filt + 1, 1, w.bcount, (const int64_t*)nullptr, w.A.cap);
        scan_i64(w.bcount, nb, w.boff, w.small);
        hipLaunchKernelGGL((kw_rotate<Wt, CT>), dim3((unsigned)nb), dim3(kBlock), 0, stream, Ls.z, Ls.x, cptr<CT>(Ls.c), n, w.A.z,
                           w.A.x, cptr<CT>(w.A.c), an, g, w.filt + 1, 0, w.bcount, (const int64_t*)w.boff, w.A.cap);
        HIPCHK(hipGetLastError());
        return read_small(0);
    }

    // Stable LSD radix sort of the m appends; returns the buffer holding them.
    template <int Wt, typename CT>
    DevSet* win_sort(int64_t m) {
        WinBuf& w = win;
        if (m <= 1) return &w.A;
        grow_set(w.T, m, 0);
        const int onb = (int)std::min<int64_t>(cdiv(m, kBlock), kOrandBlocks);
        hipLaunchKernelGGL(ks_orand<Wt>, dim3((unsigned)onb), dim3(kBlock), 0, stream, w.A.z, w.A.x, m, w.orpart);
        hipLaunchKernelGGL(ks_orand_fin<Wt>, dim3(1), dim3(kBlock), 0, stream, (const unsigned long long*)w.orpart,
                           onb, w.orand);
        unsigned long long oa[4 * 2];
        HIPCHK(hipMemcpyAsync(oa, w.orand, 4 * (size_t)Wt * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        const int64_t nb = cdiv(m, kRB);
        const int64_t nh = (int64_t)kRadix * nb;
        if (w.cap_hist < nh) {
            dfree(w.hist); dfree(w.hoff);
            w.cap_hist = std::max<int64_t>(nh, 2 * w.cap_hist);
            w.hist = dalloc<int64_t>(w.cap_hist);
            w.hoff = dalloc<int64_t>(w.cap_hist);
        }
        DevSet* src = &w.A;
        DevSet* dst = &w.T;
        for (int dd = 0; dd < 16 * Wt; ++dd) {
            const int cw = (dd * 8) >> 6;  // combined word: x words low, z words high
            const uint64_t vary = (uint64_t)(oa[cw] ^ oa[2 * Wt + cw]);
            if (((vary >> ((dd * 8) & 63)) & 0xffull) == 0) continue;
            hipLaunchKernelGGL(ks_hist<Wt>, dim3((unsigned)nb), dim3(kBlock), 0, stream, src->z, src->x, m, dd, w.hist,
                               nb);
            scan_i64(w.hist, nh, w.hoff, w.small);
            hipLaunchKernelGGL((ks_scatter<Wt, CT>), dim3((unsigned)nb), dim3(kBlock), 0, stream, src->z, src->x, cptr<CT>(src->c), m,
                               dd, w.hoff, nb, dst->z, dst->x, cptr<CT>(dst->c));
            std::swap(src, dst);
        }
        HIPCHK(hipGetLastError());
        return src;
    }

    // Merge live (n, buffer cur) with the sorted appends under filter slot
    // fslot; the result lands in buffer cur ^ 1.  Returns the output count.
    template <int Wt, typename CT>
    int64_t win_merge(int64_t n, int cur, const DevSet* S, int64_t m, int fslot) {
        constexpr int TM = 1024;
        WinBuf& w = win;
        ensure_caps(0, n + m, ib[0].C.cap, ib[0].maxlev, n, cur);
        const int64_t tiles = cdiv(n + m, TM);
        if (w.cap_status < tiles + 1) {
            dfree(w.status);
            w.cap_status = std::max<int64_t>(tiles + 1, 2 * w.cap_status);
            w.status = dalloc<uint64_t>(w.cap_status);
            HIPCHK(hipMemsetAsync(w.status, 0, (size_t)w.cap_status * sizeof(uint64_t), stream));
        }
        w.epoch = (w.epoch + 1) & (uint32_t)kEpochMask;
        if (w.epoch == 0) {
            HIPCHK(hipMemsetAsync(w.status, 0, (size_t)w.cap_status * sizeof(uint64_t), stream));
            w.epoch = 1;
        }
        HIPCHK(hipMemsetAsync(w.ticket, 0, sizeof(uint32_t), stream));
        const DevSet& Li = ib[0].L[cur];
        const DevSet& Lo = ib[0].L[cur ^ 1];
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)ncu * 4));
        hipLaunchKernelGGL((kw_merge<Wt, TM, CT>), dim3(grid), dim3(kBlock), 0, stream, Li.z, Li.x, cptr<CT>(Li.c), n, S->z, S->x,
                           cptr<CT>(S->c), m, w.filt + fslot, Lo.z, Lo.x, cptr<CT>(Lo.c), Lo.cap, w.status, w.ticket, w.epoch,
                           w.small + 1, w.err);
        HIPCHK(hipGetLastError());
        const int64_t out = read_small(1);
        check_win_err();
        return out;
    }

    // Order-preserving compaction of the live set (mode 0: filter slot 2;
    // mode 1: stochastic clip).  Returns the new count; flips cur.
    template <int Wt, typename CT>
    int64_t win_compact(int64_t n, int& cur, int mode, double eps, uint64_t seed, SampleSel ss = SampleSel{}) {
        WinBuf& w = win;
        if (n == 0) return 0;
        ensure_caps(0, n, ib[0].C.cap, ib[0].maxlev, n, cur);
        const int64_t nb = cdiv(n, kRB);
        win_blocks(nb);
        const DevSet& Li = ib[0].L[cur];
        const DevSet& Lo = ib[0].L[cur ^ 1];
        hipLaunchKernelGGL((kw_compact<Wt, CT>), dim3((unsigned)nb), dim3(kBlock), 0, stream, Li.z, Li.x, cptr<CT>(Li.c), n,
                           w.filt + 2, mode, eps, seed, 1, w.bcount, (const int64_t*)nullptr, Lo.z, Lo.x, cptr<CT>(Lo.c), ss);
        scan_i64(w.bcount, nb, w.boff, w.small);
        hipLaunchKernelGGL((kw_compact<Wt, CT>), dim3((unsigned)nb), dim3(kBlock), 0, stream, Li.z, Li.x, cptr<CT>(Li.c), n,
                           w.filt + 2, mode, eps, seed, 0, w.bcount, (const int64_t*)w.boff, Lo.z, Lo.x, cptr<CT>(Lo.c), ss);
        HIPCHK(hipGetLastError());
        cur ^= 1;
        return read_small(0);
    }

    // k-th smallest |c| (0-based) of the live set by an 8-pass radix select
    // over the IEEE bit patterns (non-negative doubles order as integers).
    template <typename CT>
    double win_select_abs(int64_t n, int cur, int64_t k) {
        WinBuf& w = win;
        const DevSet& Li = ib[0].L[cur];
        uint64_t prefix = 0, pmask = 0;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 2048));
        unsigned long long h[kRadix];
        for (int shift = 56; shift >= 0; shift -= 8) {
            HIPCHK(hipMemsetAsync(w.qhist, 0, kRadix * sizeof(unsigned long long), stream));
            hipLaunchKernelGGL(kq_hist<CT>, dim3(grid), dim3(kBlock), 0, stream, cptr<CT>(Li.c), n, shift, prefix, pmask, w.qhist);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(h, w.qhist, sizeof(h), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            int d = 0;
            for (; d < kRadix; ++d) {
                if ((int64_t)h[d] > k) break;
                k -= (int64_t)h[d];
            }
            if (d == kRadix) throw PPError(PP_EHIP, "radix select lost its rank");
            prefix |= (uint64_t)d << shift;
            pmask |= 0xffull << shift;
        }
        double v;
        std::memcpy(&v, &prefix, sizeof(v));
        return v;
    }

    // Radix select of the k-th largest (key, n - 1 - i) pair (0-based from the
    // top): eight byte passes over the keys, then over the index among the
    // terms whose key equals the selected one.
    std::pair<uint64_t, int64_t> win_select_sample(const uint64_t* key, int64_t n, int64_t k) {
        WinBuf& w = win;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 2048));
        unsigned long long h[kRadix];
        uint64_t val[2] = {0, 0};
        for (int stage = 0; stage < 2; ++stage) {
            uint64_t prefix = 0, pmask = 0;
            for (int shift = 56; shift >= 0; shift -= 8) {
                HIPCHK(hipMemsetAsync(w.qhist, 0, kRadix * sizeof(unsigned long long), stream));
                hipLaunchKernelGGL(kq_hist_sel, dim3(grid), dim3(kBlock), 0, stream, key, n, stage, val[0], shift,
                                   prefix, pmask, w.qhist);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(h, w.qhist, sizeof(h), hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
                int d = kRadix - 1;
                for (; d >= 0; --d) {  // from the largest byte value down
                    if ((int64_t)h[d] > k) break;
                    k -= (int64_t)h[d];
                }
                if (d < 0) throw PPError(PP_EHIP, "radix select lost its rank");
                prefix |= (uint64_t)d << shift;
                pmask |= 0xffull << shift;
            }
            val[stage] = prefix;
        }
        return {val[0], (int64_t)val[1]};
    }

    // StochasticSamplingTruncation (spv_evolve.jl:69-99): keep n_keep terms
    // drawn by the exponential-key method, rescaled to the original Σ|c|^2.
    template <int Wt, typename CT>
    int64_t win_sample(int64_t n, int& cur, int64_t n_keep, uint64_t seed) {
        if (n <= n_keep) return n;
        const double norm_sq = norm_sum_of<CT>(ib[0].L[cur].c, n, 0, 2.0);
        uint64_t* key = dalloc<uint64_t>((size_t)n);
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 2048));
        hipLaunchKernelGGL(kw_sample_keys<CT>, dim3(grid), dim3(kBlock), 0, stream, cptr<CT>(ib[0].L[cur].c), n, seed, key);
        HIPCHK(hipGetLastError());
        const auto sel = win_select_sample(key, n, n_keep - 1);
        SampleSel ss{key, sel.first, sel.second, n};
        n = win_compact<Wt, CT>(n, cur, 2, 0.0, 0, ss);
        HIPCHK(hipStreamSynchronize(stream));
        dfree(key);
        const double kept = norm_sum_of<CT>(ib[0].L[cur].c, n, 0, 2.0);
        if (kept > 0.0) {
            // c *= scale, componentwise for ComplexF64 (Complex * Real)
            hipLaunchKernelGGL(kw_scale, dim3(grid), dim3(kBlock), 0, stream, ib[0].L[cur].c, n * cw,
                               std::sqrt(norm_sq / kept));
            HIPCHK(hipGetLastError());
        }
        return n;
    }

    void win_filter(int slot, const pp_filter* f) {
        FilterDev d = make_filter_dev(f);
        HIPCHK(hipMemcpyAsync(win.filt + slot, &d, sizeof(d), hipMemcpyHostToDevice, stream));
    }

    // real(⟨ψ|v|ψ⟩) over live + pending appends: what EnergyCorrection measures
    // (truncation.jl:288)
    template <int Wt, typename CT>
    double win_expect(int64_t n, int cur, int64_t an, const uint64_t* ket) {
        WinBuf& w = win;
        Key<Wt> kk;
        for (int q = 0; q < Wt; ++q) {
            kk.z[q] = ket[q];
            kk.x[q] = 0;
        }
        const DevSet& Li = ib[0].L[cur];
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n + an, kBlock), 1024));
        hipLaunchKernelGGL((kw_expect<Wt, CT>), dim3(grid), dim3(kBlock), 0, stream, Li.z, Li.x, cptr<CT>(Li.c), n, w.A.z, w.A.x,
                           cptr<CT>(w.A.c), an, kk, w.partial, 0);
        hipLaunchKernelGGL(kw_expect_final, dim3(1), dim3(kBlock), 0, stream, w.partial, (int64_t)grid, w.out);
        HIPCHK(hipGetLastError());
        double v = 0.0;
        HIPCHK(hipMemcpyAsync(&v, w.out, sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return v;
    }

    // variance(O, ψ) of the merged live set (statistics.jl:6-22): the terms
    // copied with x first into the append buffer (free at a merged state),
    // radix-sorted by (x, z), then summed per x group.
    template <int Wt, typename CT>
    double win_variance(int64_t n, int cur, const uint64_t* ket) {
        WinBuf& w = win;
        if (n == 0) return 0.0;
        grow_set(w.A, n, 0);
        const DevSet& Li = ib[0].L[cur];
        const unsigned g1 = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 4096));
        hipLaunchKernelGGL((kv_swap<Wt, CT>), dim3(g1), dim3(kBlock), 0, stream, Li.z, Li.x, cptr<CT>(Li.c), n, w.A.z, w.A.x,
                           cptr<CT>(w.A.c));
        const DevSet* S = win_sort<Wt, CT>(n);
        Key<Wt> kk;
        for (int q = 0; q < Wt; ++q) {
            kk.z[q] = ket[q];
            kk.x[q] = 0;
        }
        const int64_t ntiles = cdiv(n, kVarTile);
        double2* tile_v = w.vs.get<double2>(0, (size_t)ntiles);
        int32_t* tile_h = w.vs.get<int32_t>(1, (size_t)ntiles);
        double2* carry = w.vs.get<double2>(2, (size_t)ntiles);
        const unsigned grid = (unsigned)std::min<int64_t>(ntiles, 1024);
        hipLaunchKernelGGL((kv_tiles<Wt, false, CT>), dim3(grid), dim3(kBlock), 0, stream, S->z, S->x, cptr<CT>(S->c), n, kk,
                           tile_v, tile_h, carry, w.partial, w.out);
        hipLaunchKernelGGL(kv_carry, dim3(1), dim3(kBlock), 0, stream, tile_v, tile_h, ntiles, carry, w.out);
        hipLaunchKernelGGL((kv_tiles<Wt, true, CT>), dim3(grid), dim3(kBlock), 0, stream, S->z, S->x, cptr<CT>(S->c), n, kk,
                           tile_v, tile_h, carry, w.partial, w.out);
        hipLaunchKernelGGL(kw_expect_final, dim3(1), dim3(kBlock), 0, stream, w.partial, (int64_t)grid, w.out);
        HIPCHK(hipGetLastError());
        // real(e2 - e1 * e1) with e1 = σ[ψ] = (v[1], v[2]) (statistics.jl:19-21)
        double v[3] = {0.0, 0.0, 0.0};
        HIPCHK(hipMemcpyAsync(v, w.out, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return v[0] - (v[1] * v[1] - v[2] * v[2]);
    }

    // _apply! of a non-compilable strategy: its operations in order.
    template <int Wt, typename CT>
    int64_t win_apply_ops(int64_t n, int& cur, const pp_strategy_op* ops, int nops) {
        for (int o = 0; o < nops; ++o) {
            const pp_strategy_op& op = ops[o];
            if (op.kind == PP_OP_FILTER) {
                win_filter(2, &op.filter);
                n = win_compact<Wt, CT>(n, cur, 0, 0.0, 0);
            } else if (op.kind == PP_OP_ADAPTIVE) {
                // spv_evolve.jl:103-114: threshold = the (n - max_terms)-th
                // smallest |c| (1-based), then coeff_clip!
                pp_filter f = {INT64_MAX, INT64_MAX, INT64_MAX, op.value, 0.0, -1.0, 0.0, -1.0};
                if (n > op.max_terms) f.thresh = win_select_abs<CT>(n, cur, n - op.max_terms - 1);
                win_filter(2, &f);
                n = win_compact<Wt, CT>(n, cur, 0, 0.0, 0);
            } else if (op.kind == PP_OP_STOCHASTIC_COEFF) {
                const uint64_t seed = op.seed ^ (0x9e3779b97f4a7c15ull * (++win.rng_calls));
                n = win_compact<Wt, CT>(n, cur, 1, op.value, seed);
            } else if (op.kind == PP_OP_STOCHASTIC_SAMPLING) {
                const uint64_t seed = op.seed ^ (0x9e3779b97f4a7c15ull * (++win.rng_calls));
                n = win_sample<Wt, CT>(n, cur, op.max_terms, seed);
            } else {
                throw PPError(PP_EARG, "unknown strategy operation kind " + std::to_string(op.kind));
            }
        }
        return n;
    }

    struct WinArgs {
        const pp_strategy_op* ops = nullptr;  // non-compilable boundary strategy (nops > 0)
        int nops = 0;
        const uint64_t* ket = nullptr;        // EnergyCorrection ket (null: none)
        double energy = 0.0;
        bool var = false;                     // EnergyVarianceCorrection (with ket)
        double variance = 0.0;
    };

    template <int Wt, typename CT>
    void evolve_window(const std::vector<GenDev>& gens, int64_t window, const pp_filter* strict, const pp_filter* local,
                       WinArgs& a, pp_counters* cnt) {
        win_init();
        sync_state();
        SegState& st = h_state[0];
        int cur = st.cur;
        int64_t n = st.n[cur];
        int64_t an = 0;
        int64_t app_cap = st.app_cap_v, live_cap = st.live_cap_v;
        const int64_t L = (int64_t)gens.size();
        WinBuf& w = win;
        if (w.cap_gens < L) {
            dfree(w.gens);
            w.cap_gens = L;
            w.gens = dalloc<GenDev>(L);
        }
        HIPCHK(hipMemcpyAsync(w.gens, gens.data(), (size_t)L * sizeof(GenDev), hipMemcpyHostToDevice, stream));
        // EnergyVarianceCorrection merges with NOFILTER and applies the strict
        // filter as a compaction afterwards (spv_evolve.jl:264-270)
        win_filter(0, a.nops > 0 || a.var ? nullptr : strict);
        win_filter(1, local);
        pp_counters c{};
        double t_merge = 0.0;
        auto boundary = [&](bool early) {
            const auto t0 = std::chrono::steady_clock::now();
            if (a.var) {
                const int64_t m = an;
                DevSet* S = win_sort<Wt, CT>(m);
                const int64_t out = win_merge<Wt, CT>(n, cur, S, m, 0);
                if (out > live_cap) live_cap = std::max(2 * live_cap, n + m);
                c.merge_in += n + m;
                c.merge_out += out;
                cur ^= 1;
                n = out;
                an = 0;
                const double e0 = win_expect<Wt, CT>(n, cur, 0, a.ket), v0 = win_variance<Wt, CT>(n, cur, a.ket);
                if (a.nops > 0) {
                    n = win_apply_ops<Wt, CT>(n, cur, a.ops, a.nops);
                } else if (strict) {
                    win_filter(2, strict);
                    n = win_compact<Wt, CT>(n, cur, 0, 0.0, 0);
                }
                a.energy += win_expect<Wt, CT>(n, cur, 0, a.ket) - e0;
                a.variance += win_variance<Wt, CT>(n, cur, a.ket) - v0;
                if (early) c.early_merges += 1;
                t_merge += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                return;
            }
            const double before = a.ket ? win_expect<Wt, CT>(n, cur, an, a.ket) : 0.0;
            const int64_t m = an;
            DevSet* S = win_sort<Wt, CT>(m);
            const int64_t out = win_merge<Wt, CT>(n, cur, S, m, 0);
            // _try_merge! overflow -> _grow_live!(v, n + m) (spv_kernels.jl:377-386)
            if (out > live_cap) live_cap = std::max(2 * live_cap, n + m);
            c.merge_in += n + m;
            c.merge_out += out;
            cur ^= 1;
            n = out;
            an = 0;
            if (a.nops > 0) n = win_apply_ops<Wt, CT>(n, cur, a.ops, a.nops);
            if (a.ket) a.energy += win_expect<Wt, CT>(n, cur, 0, a.ket) - before;
            if (early) c.early_merges += 1;
            t_merge += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        int since = 0;
        const auto tstart = std::chrono::steady_clock::now();
        for (int64_t i = 0; i < L; ++i) {
            if (2 * an + n > app_cap) {  // spv_evolve.jl:356-362
                boundary(true);
                since = 0;
                if (n > app_cap) app_cap = std::max(2 * app_cap, n);
            }
            c.terms_swept += n + an;
            c.rotations += 1;
            const int64_t created = win_rotate<Wt, CT>(w.gens + i, n, an, cur);
            c.terms_created += created;
            an += created;
            since += 1;
            if ((i + 1) % window == 0 || i == L - 1) {
                boundary(false);
                since = 0;
            }
        }
        (void)since;
        const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
        c.t_merge = t_merge;
        c.t_rotate = total - t_merge;
        st.cur = cur;
        st.n[cur] = n;
        st.n[cur ^ 1] = 0;
        st.flip = 0;
        st.app_cap_v = app_cap;
        st.live_cap_v = live_cap;
        st.rotations += c.rotations;
        st.swept += c.terms_swept;
        st.created += c.terms_created;
        st.merge_in += c.merge_in;
        st.merge_out += c.merge_out;
        st.early += c.early_merges;
        n_known[0] = n;
        put_state();
        upload_desc();
        HIPCHK(hipStreamSynchronize(stream));
        if (cnt) *cnt = c;
    }

    // truncate!(v, strategy, corr) for any strategy (truncation.jl:330-337).
    template <int Wt, typename CT>
    void truncate_ops(const pp_strategy_op* ops, int nops, const uint64_t* ket, double* energy, double* variance) {
        win_init();
        sync_state();
        SegState& st = h_state[0];
        int cur = st.cur;
        int64_t n = st.n[cur];
        const double before = ket ? win_expect<Wt, CT>(n, cur, 0, ket) : 0.0;
        const double vbefore = ket && variance ? win_variance<Wt, CT>(n, cur, ket) : 0.0;
        n = win_apply_ops<Wt, CT>(n, cur, ops, nops);
        if (ket) *energy += win_expect<Wt, CT>(n, cur, 0, ket) - before;
        if (ket && variance) *variance += win_variance<Wt, CT>(n, cur, ket) - vbefore;
        st.cur = cur;
        st.n[cur] = n;
        st.n[cur ^ 1] = 0;
        st.flip = 0;
        n_known[0] = n;
        put_state();
        upload_desc();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // ----------------------------------------------------------------------
    // Lookups and reductions on vector 0 (spv_ops.jl, type_SparsePauliVector.jl)

    const DevSet& live0(int64_t& n) {
        sync_state();
        const SegState& st = h_state[0];
        n = st.n[st.cur];
        return ib[0].L[st.cur];
    }

    template <int Wt, typename CT>
    void lookup(const uint64_t* z, const uint64_t* x, int64_t count, double* out, int32_t* found) {
        if (count <= 0) return;
        int64_t n = 0;
        const DevSet& L = live0(n);
        uint64_t* qz = dalloc<uint64_t>((size_t)count * Wt);
        uint64_t* qx = dalloc<uint64_t>((size_t)count * Wt);
        double* dout = dalloc<double>((size_t)count * cw);
        int32_t* dfound = dalloc<int32_t>((size_t)count);
        HIPCHK(hipMemcpyAsync(qz, z, (size_t)count * Wt * 8, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(qx, x, (size_t)count * Wt * 8, hipMemcpyHostToDevice, stream));
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(count, kBlock), 4096));
        hipLaunchKernelGGL((ko_lookup<Wt, CT>), dim3(grid), dim3(kBlock), 0, stream, L.z, L.x, cptr<CT>(L.c), n, qz, qx,
                           count, cptr<CT>(dout), dfound, (int64_t*)nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, (size_t)count * cb(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(found, dfound, (size_t)count * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        dfree(qz); dfree(qx); dfree(dout); dfree(dfound);
    }

    // lookup with the staged queries and results in `base` (the scratch array
    // of pp_lookup_noalloc; count <= kLookupNoAllocMax): no device allocation.
    template <int Wt, typename CT>
    void lookup_noalloc(const uint64_t* z, const uint64_t* x, int64_t count, double* out, int32_t* found,
                        unsigned char* base) {
        if (count <= 0) return;
        int64_t n = 0;
        const DevSet& L = live0(n);
        uint64_t* qz = reinterpret_cast<uint64_t*>(base);
        uint64_t* qx = qz + kLookupNoAllocMax * 2;
        double* dout = reinterpret_cast<double*>(qx + kLookupNoAllocMax * 2);
        int32_t* dfound = reinterpret_cast<int32_t*>(dout + kLookupNoAllocMax * 2);
        HIPCHK(hipMemcpyAsync(qz, z, (size_t)count * Wt * 8, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(qx, x, (size_t)count * Wt * 8, hipMemcpyHostToDevice, stream));
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(count, kBlock), 4096));
        hipLaunchKernelGGL((ko_lookup<Wt, CT>), dim3(grid), dim3(kBlock), 0, stream, L.z, L.x, cptr<CT>(L.c), n, qz, qx,
                           count, cptr<CT>(dout), dfound, (int64_t*)nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, (size_t)count * cb(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(found, dfound, (size_t)count * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // Device indices of `count` keys in the live set of vector 0 (-1: absent).
    template <int Wt, typename CT>
    std::vector<int64_t> find_indices(const uint64_t* z, const uint64_t* x, int64_t count, int64_t* d_idx_out) {
        int64_t n = 0;
        const DevSet& L = live0(n);
        uint64_t* qz = dalloc<uint64_t>((size_t)count * Wt);
        uint64_t* qx = dalloc<uint64_t>((size_t)count * Wt);
        double* dout = dalloc<double>((size_t)count * cw);
        int32_t* dfound = dalloc<int32_t>((size_t)count);
        HIPCHK(hipMemcpyAsync(qz, z, (size_t)count * Wt * 8, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(qx, x, (size_t)count * Wt * 8, hipMemcpyHostToDevice, stream));
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(count, kBlock), 4096));
        hipLaunchKernelGGL((ko_lookup<Wt, CT>), dim3(grid), dim3(kBlock), 0, stream, L.z, L.x, cptr<CT>(L.c), n, qz, qx,
                           count, cptr<CT>(dout), dfound, d_idx_out);
        HIPCHK(hipGetLastError());
        std::vector<int64_t> idx((size_t)count);
        HIPCHK(hipMemcpyAsync(idx.data(), d_idx_out, (size_t)count * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        dfree(qz); dfree(qx); dfree(dout); dfree(dfound);
        return idx;
    }

    void commit_live0(int cur, int64_t n, int64_t live_cap) {
        SegState& st = h_state[0];
        st.cur = cur;
        st.n[cur] = n;
        st.n[cur ^ 1] = 0;
        st.flip = 0;
        st.live_cap_v = live_cap;
        n_known[0] = n;
        put_state();
        upload_desc();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // setindex! for many keys (type_SparsePauliVector.jl:315-335); the last
    // of repeated keys wins, as with sequential calls.
    template <int Wt, typename CT>
    void setindex(const uint64_t* z, const uint64_t* x, const double* c, int64_t count) {
        if (count <= 0) return;
        // last occurrence of each key
        std::vector<int64_t> order((size_t)count);
        std::iota(order.begin(), order.end(), 0);
        auto kless = [&](int64_t a, int64_t b) {
            for (int w = Wt - 1; w >= 0; --w)
                if (z[a * Wt + w] != z[b * Wt + w]) return z[a * Wt + w] < z[b * Wt + w];
            for (int w = Wt - 1; w >= 0; --w)
                if (x[a * Wt + w] != x[b * Wt + w]) return x[a * Wt + w] < x[b * Wt + w];
            return false;
        };
        std::stable_sort(order.begin(), order.end(), kless);
        std::vector<int64_t> uniq;
        for (size_t i = 0; i < order.size(); ++i)
            if (i + 1 == order.size() || kless(order[i], order[i + 1])) uniq.push_back(order[i]);
        const int64_t u = (int64_t)uniq.size();
        std::vector<uint64_t> uz((size_t)u * Wt), ux((size_t)u * Wt);
        std::vector<double> uc((size_t)u * cw);
        for (int64_t i = 0; i < u; ++i) {
            for (int w = 0; w < Wt; ++w) {
                uz[i * Wt + w] = z[uniq[i] * Wt + w];
                ux[i * Wt + w] = x[uniq[i] * Wt + w];
            }
            for (int q = 0; q < cw; ++q) uc[i * cw + q] = c[uniq[i] * cw + q];
        }
        int64_t* didx = dalloc<int64_t>((size_t)u);
        const std::vector<int64_t> idx = find_indices<Wt, CT>(uz.data(), ux.data(), u, didx);
        dfree(didx);
        std::vector<int64_t> fi;
        std::vector<double> fv;
        std::vector<uint64_t> nz, nx;
        std::vector<double> nc;
        for (int64_t i = 0; i < u; ++i) {
            if (idx[i] >= 0) {
                fi.push_back(idx[i]);
                for (int q = 0; q < cw; ++q) fv.push_back(uc[i * cw + q]);
            } else {
                for (int w = 0; w < Wt; ++w) {
                    nz.push_back(uz[i * Wt + w]);
                    nx.push_back(ux[i * Wt + w]);
                }
                for (int q = 0; q < cw; ++q) nc.push_back(uc[i * cw + q]);
            }
        }
        SegState& st = h_state[0];
        int cur = st.cur;
        int64_t n = st.n[cur];
        int64_t live_cap = st.live_cap_v;
        if (!fi.empty()) {
            int64_t* di = dalloc<int64_t>(fi.size());
            double* dv = dalloc<double>(fv.size());
            HIPCHK(hipMemcpyAsync(di, fi.data(), fi.size() * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(dv, fv.data(), fv.size() * 8, hipMemcpyHostToDevice, stream));
            const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv((int64_t)fi.size(), kBlock), 4096));
            hipLaunchKernelGGL(ko_scatter<CT>, dim3(grid), dim3(kBlock), 0, stream, cptr<CT>(ib[0].L[cur].c),
                               (const int64_t*)di, cptr<CT>((const double*)dv), (int64_t)fi.size());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
            dfree(di);
            dfree(dv);
        }
        const int64_t m = (int64_t)nc.size() / cw;
        if (m > 0) {
            // the reference inserts one key at a time, growing when full
            for (int64_t i = 0; i < m; ++i) {
                if (n + i == live_cap) live_cap = std::max(2 * live_cap, n + i + 1);
            }
            win_init();
            WinBuf& w = win;
            grow_set(w.A, m, 0);
            HIPCHK(hipMemcpyAsync(w.A.z, nz.data(), (size_t)m * Wt * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(w.A.x, nx.data(), (size_t)m * Wt * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(w.A.c, nc.data(), (size_t)m * cb(), hipMemcpyHostToDevice, stream));
            DevSet* S = win_sort<Wt, CT>(m);
            win_filter(0, nullptr);
            n = win_merge<Wt, CT>(n, cur, S, m, 0);
            cur ^= 1;
        }
        commit_live0(cur, n, live_cap);
    }

    // sum!(v1, v2) (spv_ops.jl:182-193): v2's sorted live terms merged into
    // vector 0 with NOFILTER; _try_merge! growth when the union overflows.
    template <int Wt, typename CT>
    void sum_from(Batch& other) {
        int64_t n2 = 0;
        const DevSet& L2 = other.live0(n2);
        sync_state();
        SegState& st = h_state[0];
        int cur = st.cur;
        const int64_t n = st.n[cur];
        int64_t live_cap = st.live_cap_v;
        if (n2 == 0) return;
        win_init();
        WinBuf& w = win;
        grow_set(w.A, n2, 0);
        HIPCHK(hipMemcpyAsync(w.A.z, L2.z, (size_t)n2 * kb(), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(w.A.x, L2.x, (size_t)n2 * kb(), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(w.A.c, L2.c, (size_t)n2 * cb(), hipMemcpyDeviceToDevice, stream));
        win_filter(0, nullptr);
        const int64_t out = win_merge<Wt, CT>(n, cur, &w.A, n2, 0);
        if (out > live_cap) live_cap = std::max(2 * live_cap, n + n2);
        commit_live0(cur ^ 1, out, live_cap);
    }

    // mul!(v, a) (spv_ops.jl:12-17): a real a scales every component; a
    // complex a (ComplexF64 coefficients only) multiplies as Complex * Complex.
    void scale(double ar, double ai, bool complex_scalar) {
        int64_t n = 0;
        const DevSet& L = live0(n);
        if (n == 0) return;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 2048));
        if (complex_scalar)
            hipLaunchKernelGGL(kw_scale_c, dim3(grid), dim3(kBlock), 0, stream, cptr<pp_cplx>(L.c), n, ar, ai);
        else
            hipLaunchKernelGGL(kw_scale, dim3(grid), dim3(kBlock), 0, stream, L.c, n * cw, ar);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }

    // delete! for many keys (type_SparsePauliVector.jl:345-358); absent keys
    // are ignored.
    template <int Wt, typename CT>
    void delete_keys(const uint64_t* z, const uint64_t* x, int64_t count) {
        if (count <= 0) return;
        int64_t* didx = dalloc<int64_t>((size_t)count);
        const std::vector<int64_t> idx = find_indices<Wt, CT>(z, x, count, didx);
        std::vector<int64_t> hit;
        for (int64_t v : idx)
            if (v >= 0) hit.push_back(v);
        dfree(didx);
        SegState& st = h_state[0];
        int cur = st.cur;
        int64_t n = st.n[cur];
        if (!hit.empty()) {
            uint8_t* mark = dalloc<uint8_t>((size_t)n);
            int64_t* di = dalloc<int64_t>(hit.size());
            HIPCHK(hipMemsetAsync(mark, 0, (size_t)n, stream));
            HIPCHK(hipMemcpyAsync(di, hit.data(), hit.size() * 8, hipMemcpyHostToDevice, stream));
            const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv((int64_t)hit.size(), kBlock), 4096));
            hipLaunchKernelGGL(ko_mark, dim3(grid), dim3(kBlock), 0, stream, mark, (const int64_t*)di, (int64_t)hit.size());
            HIPCHK(hipGetLastError());
            win_init();
            SampleSel ss{};
            ss.mark = mark;
            n = win_compact<Wt, CT>(n, cur, 3, 0.0, 0, ss);
            HIPCHK(hipStreamSynchronize(stream));
            dfree(mark);
            dfree(di);
        }
        commit_live0(cur, n, st.live_cap_v);
    }

    double reduce_partials(double2* part, unsigned grid, bool is_max) {
        double* dout = dalloc<double>(1);
        if (is_max) hipLaunchKernelGGL(ko_max_final, dim3(1), dim3(kBlock), 0, stream, part, (int64_t)grid, dout);
        else hipLaunchKernelGGL(kw_expect_final, dim3(1), dim3(kBlock), 0, stream, part, (int64_t)grid, dout);
        HIPCHK(hipGetLastError());
        double v = 0.0;
        HIPCHK(hipMemcpyAsync(&v, dout, sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        dfree(dout);
        return v;
    }

    // kind 0: Σ abs2(c), 1: Σ |c|, 2: max |c|, 3: Σ |c|^p
    template <typename CT>
    double norm_sum(int kind, double p) {
        int64_t n = 0;
        const DevSet& L = live0(n);
        return norm_sum_of<CT>(L.c, n, kind, p);
    }

    template <typename CT>
    double norm_sum_of(const double* c, int64_t n, int kind, double p) {
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 1024));
        double2* part = dalloc<double2>(grid);
        hipLaunchKernelGGL(ko_norm<CT>, dim3(grid), dim3(kBlock), 0, stream, cptr<CT>(c), n, kind, p, part);
        HIPCHK(hipGetLastError());
        const double v = reduce_partials(part, grid, kind == 2);
        dfree(part);
        return v;
    }

    // Σ over this vector's terms of ko_join's mode term against `other`.
    template <int Wt, typename CT>
    double join(Batch& other, int mode, int comp = 0) {
        int64_t na = 0, nb = 0;
        const DevSet& A = live0(na);
        const DevSet& Bs = other.live0(nb);
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(na, kBlock), 1024));
        double2* part = dalloc<double2>(grid);
        hipLaunchKernelGGL((ko_join<Wt, CT>), dim3(grid), dim3(kBlock), 0, stream, A.z, A.x, cptr<CT>(A.c), na, Bs.z, Bs.x, cptr<CT>(Bs.c), nb,
                           mode, comp, part);
        HIPCHK(hipGetLastError());
        const double v = reduce_partials(part, grid, false);
        dfree(part);
        return v;
    }

    // copy(v) (type_SparsePauliVector.jl:249-254): vector 0 takes src's live
    // set and the reference's virtual capacities, device to device.
    void copy_from(Batch& src) {
        int64_t n = 0;
        const DevSet& S = src.live0(n);
        const int64_t lcap = src.h_state[0].live_cap_v, acap = src.h_state[0].app_cap_v;
        sync_state();
        ensure_caps(0, std::max<int64_t>(2 * n + 1024, 4096), std::max<int64_t>(n + 1024, ib[0].C.cap), ib[0].maxlev,
                    0, 0);
        put_live0(S.z, S.x, S.c, n, lcap, acap);
    }

    // The live set of vector 0 := n device terms (sorted, unique), with the
    // given virtual capacities; buffers already hold n terms.
    void put_live0(const uint64_t* z, const uint64_t* x, const double* c, int64_t n, int64_t live_cap,
                   int64_t app_cap) {
        DevSet& d = ib[0].L[0];
        if (n > 0) {
            HIPCHK(hipMemcpyAsync(d.z, z, (size_t)n * kb(), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(d.x, x, (size_t)n * kb(), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(d.c, c, (size_t)n * cb(), hipMemcpyDeviceToDevice, stream));
        }
        SegState& st = h_state[0];
        st.n[0] = n;
        st.n[1] = 0;
        st.cur = 0;
        st.flip = 0;
        st.m = 0;
        st.stalled = 0;
        st.compact = 0;
        st.live_cap_v = live_cap;
        st.app_cap_v = app_cap;
        n_known[0] = n;
        put_state();
        upload_desc();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // otimes(v1, v2) (spv_ops.jl:342-358) into vector 0 (N = N1 + N2): every
    // pair's key and product, radix-sorted (the keys are distinct, so the
    // reference's sort + _triples_to_live! is the sorted copy); the capacities
    // of the reference's result, live max(16, n) and append 16.
    template <int Wt, typename CT>
    void otimes_from(Batch& a, int Na, Batch& b) {
        int64_t na = 0, nb = 0;
        const DevSet& A = a.live0(na);
        const DevSet& Bv = b.live0(nb);
        sync_state();
        const int64_t m = na * nb;
        if (m > 0) {
            win_init();
            grow_set(win.A, m, 0);
            const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(m, kBlock), 8192));
            hipLaunchKernelGGL(ko_otimes<CT>, dim3(grid), dim3(kBlock), 0, stream, A.z, A.x, cptr<CT>(A.c), na, a.W,
                               Bv.z, Bv.x, cptr<CT>(Bv.c), nb, b.W, Na, win.A.z, win.A.x, cptr<CT>(win.A.c), W);
            HIPCHK(hipGetLastError());
        }
        const DevSet* S = m > 0 ? win_sort<Wt, CT>(m) : nullptr;
        ensure_caps(0, std::max<int64_t>(2 * m + 1024, 4096), std::max<int64_t>(m + 1024, ib[0].C.cap), ib[0].maxlev,
                    0, 0);
        put_live0(S ? S->z : nullptr, S ? S->x : nullptr, S ? S->c : nullptr, m, std::max<int64_t>(16, m), 16);
    }

    // largest / largest_diag (spv_ops.jl:520-553): index of the first term with
    // the largest |c| (among x == 0 terms for diag_only); -1 if none.
    template <int Wt, typename CT>
    int64_t largest(int diag_only) {
        int64_t n = 0;
        const DevSet& L = live0(n);
        if (n == 0) return -1;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 1024));
        double2* part = dalloc<double2>(grid);
        int64_t* dout = dalloc<int64_t>(1);
        hipLaunchKernelGGL((ko_argmax<Wt, CT>), dim3(grid), dim3(kBlock), 0, stream, L.x, cptr<CT>(L.c), n, diag_only,
                           part);
        hipLaunchKernelGGL(ko_argmax_final, dim3(1), dim3(kBlock), 0, stream, part, (int64_t)grid, dout);
        HIPCHK(hipGetLastError());
        int64_t idx = -1;
        HIPCHK(hipMemcpyAsync(&idx, dout, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        dfree(part);
        dfree(dout);
        return idx;
    }

    // Term i of the live set of vector 0 to the host.
    void fetch_term(int64_t i, uint64_t* z, uint64_t* x, double* c) {
        int64_t n = 0;
        const DevSet& L = live0(n);
        if (i < 0 || i >= n) throw PPError(PP_EARG, "term index out of range");
        HIPCHK(hipMemcpyAsync(z, L.z + i * W, kb(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(x, L.x + i * W, kb(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(c, L.c + i * cw, cb(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // diag(v) / offdiag(v) in place (spv_ops.jl:555-587): the terms with
    // x == 0 (or x != 0), order kept.
    template <int Wt, typename CT>
    void select_diag(int offdiag) {
        win_init();
        sync_state();
        SegState& st = h_state[0];
        int cur = st.cur;
        int64_t n = st.n[cur];
        n = win_compact<Wt, CT>(n, cur, offdiag ? 5 : 4, 0.0, 0);
        commit_live0(cur, n, st.live_cap_v);
    }

    // Matrix(v) (spv_ops.jl:109-115) for N <= 13: 2^N x 2^N entries,
    // column-major, cw doubles each, into host memory.
    template <int Wt, typename CT>
    void matrix(int N, double* out) {
        int64_t n = 0;
        const DevSet& L = live0(n);
        const int64_t D = (int64_t)1 << N;
        const size_t bytes = (size_t)(D * D) * cb();
        win_init();
        WinBuf& w = win;
        double* dout = dalloc<double>((size_t)(D * D) * cw);
        const DevSet* S = nullptr;
        if (n > 0) {
            grow_set(w.A, n, 0);
            const unsigned g1 = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kBlock), 4096));
            hipLaunchKernelGGL((kv_swap<Wt, CT>), dim3(g1), dim3(kBlock), 0, stream, L.z, L.x, cptr<CT>(L.c), n,
                               w.A.z, w.A.x, cptr<CT>(w.A.c));
            S = win_sort<Wt, CT>(n);
        }
        HIPCHK(hipMemsetAsync(w.err, 0, sizeof(int32_t), stream));
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(D * D, kBlock), 16384));
        hipLaunchKernelGGL(ko_matrix<CT>, dim3(grid), dim3(kBlock), 0, stream, S ? S->z : w.A.z, S ? S->x : w.A.x,
                           cptr<CT>(S ? S->c : w.A.c), n, Wt, N, dout, w.err);
        HIPCHK(hipGetLastError());
        int32_t e = 0;
        HIPCHK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(&e, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        dfree(dout);
        HIPCHK(hipMemsetAsync(w.err, 0, sizeof(int32_t), stream));
        if (e)
            throw PPError(PP_EARG, "InexactError: Matrix(v) of a Float64 vector with an imaginary entry "
                                   "(a term with an odd number of Y factors); use ComplexF64 coefficients");
    }

    void upload(int b, const uint64_t* z, const uint64_t* x, const double* c, int64_t n, bool sorted) {
        const int Wl = W;
        std::vector<uint64_t> hz(z, z + n * Wl), hx(x, x + n * Wl);
        const int C = cw;
        std::vector<double> hc(c, c + n * C);
        auto less = [&](int64_t a, int64_t q) {
            for (int w = Wl - 1; w >= 0; --w)
                if (hz[a * Wl + w] != hz[q * Wl + w]) return hz[a * Wl + w] < hz[q * Wl + w];
            for (int w = Wl - 1; w >= 0; --w)
                if (hx[a * Wl + w] != hx[q * Wl + w]) return hx[a * Wl + w] < hx[q * Wl + w];
            return false;
        };
        if (!sorted) {
            std::vector<int64_t> perm((size_t)n);
            std::iota(perm.begin(), perm.end(), 0);
            std::sort(perm.begin(), perm.end(), less);
            std::vector<uint64_t> sz((size_t)n * Wl), sx((size_t)n * Wl);
            std::vector<double> sc((size_t)n * C);
            for (int64_t i = 0; i < n; ++i) {
                for (int w = 0; w < Wl; ++w) {
                    sz[i * Wl + w] = hz[perm[i] * Wl + w];
                    sx[i * Wl + w] = hx[perm[i] * Wl + w];
                }
                for (int q = 0; q < C; ++q) sc[i * C + q] = hc[perm[i] * C + q];
            }
            hz.swap(sz);
            hx.swap(sx);
            hc.swap(sc);
        }
        for (int64_t i = 1; i < n; ++i)
            if (!less(i - 1, i))
                throw PPError(PP_EARG, sorted ? "keys not strictly increasing" : "duplicate keys in upload");
        sync_state();
        const int64_t cap = std::max<int64_t>(2 * n + 1024, 4096);
        ensure_caps(b, cap, std::max<int64_t>(n + 1024, ib[b].C.cap), ib[b].maxlev, 0, 0);
        DevSet& s = ib[b].L[0];
        if (n > 0) {
            HIPCHK(hipMemcpyAsync(s.z, hz.data(), (size_t)n * Wl * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(s.x, hx.data(), (size_t)n * Wl * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(s.c, hc.data(), (size_t)n * cb(), hipMemcpyHostToDevice, stream));
        }
        SegState& st = h_state[b];
        st.n[0] = n;
        st.n[1] = 0;
        st.cur = 0;
        st.flip = 0;
        st.m = 0;
        st.stalled = 0;
        st.compact = 0;
        if (n > st.live_cap_v) st.live_cap_v = n;
        n_known[b] = n;
        put_state();
        upload_desc();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // Pauli-sum products (spv_ops.jl:255-484) on the windowed path's scratch
    // (the append buffer is empty between calls): pairs in generation order,
    // stable radix sort by key, run sums.  Host arrays in and out; returns
    // the output count (terms written only when it fits `cap`).
    template <int Wt, typename CT>
    int64_t algebra(int op, int cplx, const pp_terms& a, const pp_terms& b, uint64_t* oz, uint64_t* ox, double* ore,
                    double* oim, int64_t cap) {
        win_init();
        WinBuf& w = win;
        const int64_t na = a.n, nb = b.n;
        if (na == 0 || nb == 0) return 0;
        if (na > (int64_t)1 << 31 || nb > (int64_t)1 << 31 || na * nb > ((int64_t)1 << 40))
            throw PPError(PP_EARG, "too many pairs");
        DevArena& tmp = w.alg;
        auto upload = [&](const pp_terms& t, AlgOperand& d, int slot) {
            uint64_t* z = tmp.get<uint64_t>(slot, (size_t)t.n * Wt);
            uint64_t* x = tmp.get<uint64_t>(slot + 1, (size_t)t.n * Wt);
            double* re = tmp.get<double>(slot + 2, (size_t)t.n);
            double* im = cplx ? tmp.get<double>(slot + 3, (size_t)t.n) : nullptr;
            HIPCHK(hipMemcpyAsync(z, t.z, (size_t)t.n * Wt * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(x, t.x, (size_t)t.n * Wt * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(re, t.re, (size_t)t.n * 8, hipMemcpyHostToDevice, stream));
            if (cplx) HIPCHK(hipMemcpyAsync(im, t.im, (size_t)t.n * 8, hipMemcpyHostToDevice, stream));
            d = AlgOperand{z, x, re, im, t.n, t.conj};
        };
        AlgOperand A, Bo;
        upload(a, A, 0);
        upload(b, Bo, 4);
        if (op != kAlgMul) A.conj = Bo.conj = 0;
        HIPCHK(hipMemsetAsync(w.err, 0, sizeof(int32_t), stream));
        // pairs: count, scan, write
        const int64_t total = na * nb;
        const int64_t nblk = cdiv(total, kRB);
        win_blocks(nblk);
        hipLaunchKernelGGL(ka_pairs<Wt>, dim3((unsigned)nblk), dim3(kBlock), 0, stream, A, Bo, op, cplx, 1, w.bcount,
                           (const int64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr, (double*)nullptr,
                           (double*)nullptr, (double*)nullptr, w.err);
        scan_i64(w.bcount, nblk, w.boff, w.small);
        const int64_t m = read_small(0);
        if (m == 0) return 0;
        grow_set(w.A, m, 0);
        double* pre = tmp.get<double>(8, (size_t)m);
        double* pim = cplx ? tmp.get<double>(9, (size_t)m) : nullptr;
        hipLaunchKernelGGL(ka_pairs<Wt>, dim3((unsigned)nblk), dim3(kBlock), 0, stream, A, Bo, op, cplx, 0, w.bcount,
                           (const int64_t*)w.boff, w.A.z, w.A.x, w.A.c, pre, pim, w.err);
        HIPCHK(hipGetLastError());
        int32_t e = 0;
        HIPCHK(hipMemcpyAsync(&e, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (e & 16)
            throw PPError(PP_EARG, "InexactError: a product coefficient is not real (real coefficient type)");
        // the pair records carry one double each, whatever the handle's type
        const DevSet* S = win_sort<Wt, double>(m);
        // runs: count, scan, write
        const int64_t mblk = cdiv(m, kRB);
        win_blocks(mblk);
        hipLaunchKernelGGL(ka_runs<Wt>, dim3((unsigned)mblk), dim3(kBlock), 0, stream, S->z, S->x, S->c, m,
                           (const double*)pre, (const double*)pim, cplx, 1, w.bcount, (const int64_t*)nullptr,
                           (uint64_t*)nullptr, (uint64_t*)nullptr, (double*)nullptr, (double*)nullptr);
        scan_i64(w.bcount, mblk, w.boff, w.small);
        const int64_t nout = read_small(0);
        if (nout > cap) return nout;
        uint64_t* dz = tmp.get<uint64_t>(10, (size_t)nout * Wt);
        uint64_t* dx = tmp.get<uint64_t>(11, (size_t)nout * Wt);
        double* dre = tmp.get<double>(12, (size_t)nout);
        double* dim = cplx ? tmp.get<double>(13, (size_t)nout) : nullptr;
        hipLaunchKernelGGL(ka_runs<Wt>, dim3((unsigned)mblk), dim3(kBlock), 0, stream, S->z, S->x, S->c, m,
                           (const double*)pre, (const double*)pim, cplx, 0, w.bcount, (const int64_t*)w.boff, dz, dx,
                           dre, dim);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(oz, dz, (size_t)nout * Wt * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(ox, dx, (size_t)nout * Wt * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(ore, dre, (size_t)nout * 8, hipMemcpyDeviceToHost, stream));
        if (cplx) HIPCHK(hipMemcpyAsync(oim, dim, (size_t)nout * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return nout;
    }

    int64_t download(int b, uint64_t* z, uint64_t* x, double* c, int64_t cap) {
        sync_state();
        const SegState& st = h_state[b];
        const int64_t n = st.n[st.cur];
        if (cap < n) throw PPError(PP_EARG, "output capacity too small");
        const DevSet& s = ib[b].L[st.cur];
        if (n > 0) {
            HIPCHK(hipMemcpyAsync(z, s.z, (size_t)n * W * 8, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipMemcpyAsync(x, s.x, (size_t)n * W * 8, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipMemcpyAsync(c, s.c, (size_t)n * cb(), hipMemcpyDeviceToHost, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        return n;
    }
};

// ---------------------------------------------------------------------------
// Rotation tables

// g must arrive zeroed (levels past nlev stay zero).
static void make_gen(GenDev& g, int W, const uint64_t* gz, const uint64_t* gx, double cosv, double sinv) {
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
        for (uint64_t m = gx[w]; m; m &= m - 1) g.lev[nlev++] = (uint8_t)(64 * w + __builtin_ctzll(m));
    for (int w = 0; w < W; ++w)  // z bits above
        for (uint64_t m = gz[w]; m; m &= m - 1) g.lev[nlev++] = (uint8_t)(64 * W + 64 * w + __builtin_ctzll(m));
    g.nlev = nlev;
}

// Steps for B vectors: generator i of vector b is gz[(i * B + b) * W]; active
// may be null (all active).  The first active rotation of each vector skips
// the per-step precheck (the sequence-start precheck covers it).
// st: nsteps * B entries that are zero or hold an earlier call's steps.
static void make_steps(const StepSpan& st, int W, int B, int64_t nsteps, const uint64_t* gz, const uint64_t* gx,
                       const double* cosv, const double* sinv, const uint8_t* active, std::vector<int32_t>& en) {
    en.assign(B, 0);
    std::vector<char> seen(B, 0);
    for (int64_t s = 0; s < nsteps; ++s)
        for (int b = 0; b < B; ++b) {
            const size_t k = (size_t)(s * B + b);
            StepDev& d = st[k];
            const int old = std::min<int>(std::max<int>(d.g.nlev, 0), kMaxLevels);
            std::memset(&d.g, 0, offsetof(GenDev, lev));
            std::memset(d.g.lev, 0, (size_t)old);  // levels past the new nlev stay zero
            make_gen(d.g, W, gz + k * W, gx + k * W, cosv[k], sinv[k]);
            d.active = active ? (active[k] ? 1 : 0) : 1;
            d.precheck = 0;
            if (d.active) {
                d.precheck = seen[b] ? 1 : 0;
                seen[b] = 1;
                en[b] = 1;
            }
        }
}

static void add_counters(pp_counters* c, const SegState& a, const SegState& b) {
    c->rotations += b.rotations - a.rotations;
    c->terms_swept += b.swept - a.swept;
    c->terms_created += b.created - a.created;
    c->merge_in += b.merge_in - a.merge_in;
    c->merge_out += b.merge_out - a.merge_out;
    c->early_merges += b.early - a.early;
}

struct pp_handle {
    Batch bt;
    int N = 0;
};

struct pp_batch {
    Batch bt;
};

#define API_BEGIN try {
#define API_END                             \
    }                                       \
    catch (const PPError& e) {              \
        return fail(e.code, e.msg);         \
    }                                       \
    catch (const std::exception& e) {       \
        return fail(PP_EHIP, e.what());     \
    }                                       \
    g_last_error.clear();                   \
    return PP_OK;

template <typename T>
struct TypeTag {
    using type = T;
};

// f(words per key, coefficient type) for a batch's layout.
template <typename F>
static void dispatch_wc(const Batch& bt, F&& f) {
    using W1 = std::integral_constant<int, 1>;
    using W2 = std::integral_constant<int, 2>;
    if (bt.cw == 1) {
        if (bt.W == 1) f(W1(), TypeTag<double>());
        else f(W2(), TypeTag<double>());
    } else {
        if (bt.W == 1) f(W1(), TypeTag<pp_cplx>());
        else f(W2(), TypeTag<pp_cplx>());
    }
}

static void batch_evolve(Batch& bt, int64_t nsteps, const uint64_t* gz, const uint64_t* gx, const double* cosv,
                         const double* sinv, const uint8_t* active, const pp_filter* strict, const pp_filter* local,
                         pp_counters* counters) {
    std::vector<int32_t> en;
    bt.sync_state();
    const StepSpan st = bt.host_steps((size_t)(nsteps * bt.B));
    make_steps(st, bt.W, bt.B, nsteps, gz, gx, cosv, sinv, active, en);
    std::vector<SegState> before(bt.h_state, bt.h_state + bt.B);
    bt.set_filters(strict, local);
    dispatch_wc(bt, [&](auto wc, auto ct) { bt.run_sequence<decltype(wc)::value, typename decltype(ct)::type>(st, nsteps, en); });
    if (counters)
        for (int b = 0; b < bt.B; ++b) {
            std::memset(counters + b, 0, sizeof(pp_counters));
            add_counters(counters + b, before[b], bt.h_state[b]);
        }
}

extern "C" {

const char* pp_last_error(void) { return g_last_error.c_str(); }

const char* pp_version(void) { return "pauliprop 0.2 gfx950"; }

double pp_julia_powi(double x, int64_t n) { return julia_powi(x, n); }

// --- single vector ------------------------------------------------------------

int pp_create(int N, int64_t live_cap, int64_t append_cap, pp_handle** out) {
    return pp_create_typed(N, live_cap, append_cap, 0, out);
}

int pp_create_typed(int N, int64_t live_cap, int64_t append_cap, int complex_coeffs, pp_handle** out) {
    if (!out) return fail(PP_EARG, "out is NULL");
    *out = nullptr;
    if (N < 1 || N > 128) return fail(PP_EARG, "N must be in 1:128");
    if (live_cap < 1 || append_cap < 1) return fail(PP_EARG, "capacities must be >= 1");
    if (complex_coeffs != 0 && complex_coeffs != 1) return fail(PP_EARG, "complex_coeffs must be 0 or 1");
    pp_handle* h = new pp_handle();
    try {
        h->N = N;
        h->bt.init(N <= 64 ? 1 : 2, {N}, {live_cap}, {append_cap}, complex_coeffs ? 2 : 1);
    } catch (const PPError& e) {
        delete h;
        return fail(e.code, e.msg);
    }
    const int rc = scratch_setup(h->bt.dev);
    if (rc != PP_OK) {
        delete h;
        return rc;
    }
    *out = h;
    g_last_error.clear();
    return PP_OK;
}

int pp_coeff_type(pp_handle* h, int* complex_coeffs) {
    if (!h || !complex_coeffs) return fail(PP_EARG, "null argument");
    *complex_coeffs = h->bt.cw == 2 ? 1 : 0;
    g_last_error.clear();
    return PP_OK;
}

int pp_destroy(pp_handle* h) {
    if (!h) return PP_OK;
    (void)hipStreamSynchronize(h->bt.stream);
    delete h;
    return PP_OK;
}

int pp_upload(pp_handle* h, const uint64_t* z, const uint64_t* x, const double* c, int64_t n, int sorted) {
    if (!h) return fail(PP_EARG, "null handle");
    if (n < 0 || (n > 0 && (!z || !x || !c))) return fail(PP_EARG, "bad arrays");
    API_BEGIN
    h->bt.upload(0, z, x, c, n, sorted != 0);
    API_END
}

int pp_algebra(pp_handle* h, int op, int complex_coeffs, const pp_terms* a, const pp_terms* b, uint64_t* oz,
               uint64_t* ox, double* ore, double* oim, int64_t cap, int64_t* nout) {
    if (!h || !a || !b || !nout) return fail(PP_EARG, "null argument");
    if (op < PP_ALG_MUL || op > PP_ALG_ANTICOMMUTATOR) return fail(PP_EARG, "unknown operation");
    if (op == PP_ALG_COMMUTATOR && !complex_coeffs)
        return fail(PP_EARG, "commutator coefficients are imaginary; use a complex coefficient type (T=ComplexF64)");
    for (const pp_terms* t : {a, b}) {
        if (t->n < 0 || (t->n > 0 && (!t->z || !t->x || !t->re || (complex_coeffs && !t->im))))
            return fail(PP_EARG, "bad operand arrays");
    }
    if (cap < 0 || (cap > 0 && (!oz || !ox || !ore || (complex_coeffs && !oim))))
        return fail(PP_EARG, "bad output arrays");
    API_BEGIN
    int64_t n = 0;
    dispatch_wc(h->bt, [&](auto wc, auto ct) {
        n = h->bt.algebra<decltype(wc)::value, typename decltype(ct)::type>(op, complex_coeffs ? 1 : 0, *a, *b, oz, ox, ore, oim, cap);
    });
    *nout = n;
    if (n > cap) return fail(PP_EOVERFLOW, "output capacity too small (" + std::to_string(n) + " terms)");
    API_END
}

int pp_download(pp_handle* h, uint64_t* z, uint64_t* x, double* c, int64_t cap, int64_t* n) {
    if (!h || !n) return fail(PP_EARG, "null argument");
    API_BEGIN
    *n = h->bt.download(0, z, x, c, cap);
    API_END
}

int pp_size(pp_handle* h, int64_t* n) {
    if (!h || !n) return fail(PP_EARG, "null argument");
    API_BEGIN
    h->bt.sync_state();
    *n = h->bt.n_known[0];
    API_END
}

int pp_capacities(pp_handle* h, int64_t* live_cap, int64_t* append_cap) {
    if (!h || !live_cap || !append_cap) return fail(PP_EARG, "null argument");
    API_BEGIN
    h->bt.sync_state();
    *live_cap = h->bt.h_state[0].live_cap_v;
    *append_cap = h->bt.h_state[0].app_cap_v;
    API_END
}

int pp_evolve(pp_handle* h, const uint64_t* gz, const uint64_t* gx, const double* cosv, const double* sinv,
              int64_t L, int64_t window, const pp_filter* strict, const pp_filter* local, pp_counters* counters) {
    if (!h) return fail(PP_EARG, "null handle");
    if (window < 1) return fail(PP_EARG, "window must be >= 1");
    if (L < 0) return fail(PP_EARG, "L must be >= 0");
    if (L > 0 && (!gz || !gx || !cosv || !sinv)) return fail(PP_EARG, "null generator arrays");
    if (counters) std::memset(counters, 0, sizeof(*counters));
    if (L == 0) return PP_OK;
    if (window != 1) return pp_evolve_ex(h, gz, gx, cosv, sinv, L, window, strict, nullptr, 0, local, nullptr,
                                         nullptr, nullptr, counters);
    API_BEGIN
    Batch& bt = h->bt;
    bt.timing = counters != nullptr;
    bt.t_merge = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (counters) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, bt.stream));
    }
    batch_evolve(bt, L, gz, gx, cosv, sinv, nullptr, strict, local, counters);
    if (counters) {
        HIPCHK(hipEventRecord(e1, bt.stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        counters->t_merge = bt.t_merge;
        counters->t_rotate = ms * 1e-3 - bt.t_merge;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    bt.timing = false;
    API_END
}

static void check_ops(const pp_strategy_op* ops, int nops) {
    if (nops < 0 || (nops > 0 && !ops)) throw PPError(PP_EARG, "bad strategy operations");
    for (int o = 0; o < nops; ++o) {
        const pp_strategy_op& op = ops[o];
        if (op.kind == PP_OP_ADAPTIVE && op.max_terms < 0) throw PPError(PP_EARG, "max_terms must be >= 0");
        if (op.kind == PP_OP_STOCHASTIC_COEFF && !(op.value > 0.0)) throw PPError(PP_EARG, "epsilon must be > 0");
        if (op.kind == PP_OP_STOCHASTIC_SAMPLING && op.max_terms < 0) throw PPError(PP_EARG, "n_keep must be >= 0");
        if (op.kind < PP_OP_FILTER || op.kind > PP_OP_STOCHASTIC_SAMPLING)
            throw PPError(PP_EARG, "unknown strategy operation kind " + std::to_string(op.kind));
    }
}

int pp_evolve_ex(pp_handle* h, const uint64_t* gz, const uint64_t* gx, const double* cosv, const double* sinv,
                 int64_t L, int64_t window, const pp_filter* strict, const pp_strategy_op* ops, int nops,
                 const pp_filter* local, const uint64_t* energy_ket, double* energy_acc, double* variance_acc,
                 pp_counters* counters) {
    if (!h) return fail(PP_EARG, "null handle");
    if (window < 1) return fail(PP_EARG, "window must be >= 1");
    if (L < 0) return fail(PP_EARG, "L must be >= 0");
    if (L > 0 && (!gz || !gx || !cosv || !sinv)) return fail(PP_EARG, "null generator arrays");
    if (energy_ket && !energy_acc) return fail(PP_EARG, "energy_acc is NULL");
    if (variance_acc && !energy_ket) return fail(PP_EARG, "the variance correction needs energy_ket");
    if (counters) std::memset(counters, 0, sizeof(*counters));
    if (L == 0) return PP_OK;
    API_BEGIN
    check_ops(ops, nops);
    Batch& bt = h->bt;
    std::vector<GenDev> gens((size_t)L);  // value-initialised: zero
    for (int64_t i = 0; i < L; ++i) make_gen(gens[(size_t)i], bt.W, gz + i * bt.W, gx + i * bt.W, cosv[i], sinv[i]);
    Batch::WinArgs a;
    a.ops = ops;
    a.nops = nops;
    a.ket = energy_ket;
    a.energy = energy_ket ? *energy_acc : 0.0;
    a.var = variance_acc != nullptr;
    a.variance = variance_acc ? *variance_acc : 0.0;
    dispatch_wc(bt, [&](auto wc, auto ct) { bt.evolve_window<decltype(wc)::value, typename decltype(ct)::type>(gens, window, strict, local, a, counters); });
    if (energy_ket) *energy_acc = a.energy;
    if (variance_acc) *variance_acc = a.variance;
    API_END
}

int pp_truncate(pp_handle* h, const pp_strategy_op* ops, int nops, const uint64_t* energy_ket, double* energy_acc,
                double* variance_acc) {
    if (!h) return fail(PP_EARG, "null handle");
    if (energy_ket && !energy_acc) return fail(PP_EARG, "energy_acc is NULL");
    if (variance_acc && !energy_ket) return fail(PP_EARG, "the variance correction needs energy_ket");
    API_BEGIN
    check_ops(ops, nops);
    double e = energy_ket ? *energy_acc : 0.0;
    double v = variance_acc ? *variance_acc : 0.0;
    dispatch_wc(h->bt, [&](auto wc, auto ct) {
        h->bt.truncate_ops<decltype(wc)::value, typename decltype(ct)::type>(ops, nops, energy_ket, &e, variance_acc ? &v : nullptr);
    });
    if (energy_ket) *energy_acc = e;
    if (variance_acc) *variance_acc = v;
    API_END
}

int pp_variance_ket(pp_handle* h, const uint64_t* ket, double* out) {
    if (!h || !ket || !out) return fail(PP_EARG, "null argument");
    API_BEGIN
    Batch& bt = h->bt;
    bt.win_init();
    bt.sync_state();
    const SegState& st = bt.h_state[0];
    const int64_t n = st.n[st.cur];
    const int cur = st.cur;
    dispatch_wc(bt, [&](auto wc, auto ct) { *out = bt.win_variance<decltype(wc)::value, typename decltype(ct)::type>(n, cur, ket); });
    API_END
}

int pp_lookup(pp_handle* h, const uint64_t* z, const uint64_t* x, int64_t count, double* c, int32_t* found) {
    if (!h || count < 0 || (count > 0 && (!z || !x || !c || !found))) return fail(PP_EARG, "bad arguments");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.lookup<decltype(wc)::value, typename decltype(ct)::type>(z, x, count, c, found); });
    API_END
}

int pp_lookup_noalloc(pp_handle* h, const uint64_t* z, const uint64_t* x, int64_t count, double* c,
                      int32_t* found) {
    if (!h || count < 0 || (count > 0 && (!z || !x || !c || !found))) return fail(PP_EARG, "bad arguments");
    if (count > kLookupNoAllocMax)
        return fail(PP_EARG, "pp_lookup_noalloc: count exceeds the scratch capacity of " +
                                 std::to_string(kLookupNoAllocMax) + " queries per call");
    const int dev = h->bt.dev;
    unsigned char* base = dev >= 0 && dev < kScratchDevices ? static_cast<unsigned char*>(g_scratch_ptr[dev]) : nullptr;
    if (!base) return fail(PP_EHIP, "pp_lookup_noalloc: no scratch address for this device");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.lookup_noalloc<decltype(wc)::value, typename decltype(ct)::type>(z, x, count, c, found, base); });
    API_END
}

int pp_setindex(pp_handle* h, const uint64_t* z, const uint64_t* x, const double* c, int64_t count) {
    if (!h || count < 0 || (count > 0 && (!z || !x || !c))) return fail(PP_EARG, "bad arguments");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.setindex<decltype(wc)::value, typename decltype(ct)::type>(z, x, c, count); });
    API_END
}

int pp_delete(pp_handle* h, const uint64_t* z, const uint64_t* x, int64_t count) {
    if (!h || count < 0 || (count > 0 && (!z || !x))) return fail(PP_EARG, "bad arguments");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.delete_keys<decltype(wc)::value, typename decltype(ct)::type>(z, x, count); });
    API_END
}

int pp_sum(pp_handle* a, pp_handle* b) {
    if (!a || !b) return fail(PP_EARG, "null argument");
    if (a == b) return fail(PP_EARG, "sum!(v, v) needs two distinct vectors");
    if (a->N != b->N) return fail(PP_EARG, "qubit counts differ");
    if (a->bt.cw != b->bt.cw) return fail(PP_EARG, "coefficient types differ");
    API_BEGIN
    dispatch_wc(a->bt, [&](auto wc, auto ct) { a->bt.sum_from<decltype(wc)::value, typename decltype(ct)::type>(b->bt); });
    API_END
}

int pp_scale(pp_handle* h, double a) {
    if (!h) return fail(PP_EARG, "null handle");
    API_BEGIN
    h->bt.scale(a, 0.0, false);
    API_END
}

int pp_scale_complex(pp_handle* h, double re, double im) {
    if (!h) return fail(PP_EARG, "null handle");
    if (h->bt.cw != 2)
        return fail(PP_EARG, "InexactError: a complex scalar needs ComplexF64 coefficients");
    API_BEGIN
    h->bt.scale(re, im, true);
    API_END
}

int pp_norm(pp_handle* h, double p, double* out) {
    if (!h || !out) return fail(PP_EARG, "null argument");
    if (!(p > 0)) return fail(PP_EARG, "p must be > 0");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto, auto ct) {
        using CT = typename decltype(ct)::type;
        Batch& bt = h->bt;
        if (p == 2.0) *out = std::sqrt(bt.norm_sum<CT>(0, 2.0));
        else if (p == 1.0) *out = bt.norm_sum<CT>(1, 1.0);
        else if (std::isinf(p)) *out = bt.norm_sum<CT>(2, p);
        else *out = std::pow(bt.norm_sum<CT>(3, p), 1.0 / p);
    });
    API_END
}

int pp_inner_product(pp_handle* a, pp_handle* b, double* out) {
    if (!a || !b || !out) return fail(PP_EARG, "null argument");
    if (a->N != b->N) return fail(PP_EARG, "qubit counts differ");
    if (a->bt.cw != b->bt.cw) return fail(PP_EARG, "coefficient types differ");
    API_BEGIN
    dispatch_wc(a->bt, [&](auto wc, auto ct) {
        for (int comp = 0; comp < a->bt.cw; ++comp)
            out[comp] = a->bt.join<decltype(wc)::value, typename decltype(ct)::type>(b->bt, 0, comp);
    });
    API_END
}

int pp_diff_norm(pp_handle* a, pp_handle* b, double* out) {
    if (!a || !b || !out) return fail(PP_EARG, "null argument");
    if (a->N != b->N) return fail(PP_EARG, "qubit counts differ");
    if (a->bt.cw != b->bt.cw) return fail(PP_EARG, "coefficient types differ");
    API_BEGIN
    dispatch_wc(a->bt, [&](auto wc, auto ct) {
        constexpr int Wt = decltype(wc)::value;
        using CT = typename decltype(ct)::type;
        const double s1 = a->bt.join<Wt, CT>(b->bt, 1);
        const double s2 = b->bt.join<Wt, CT>(a->bt, 2);
        *out = std::sqrt(s1 + s2);
    });
    API_END
}

int pp_evolve_single(pp_handle* h, const uint64_t* gz, const uint64_t* gx, double cosv, double sinv) {
    if (!h || !gz || !gx) return fail(PP_EARG, "null argument");
    API_BEGIN
    // spv_evolve.jl:228-236: grow the append length if short, rotate, merge
    // with NOFILTER; no early merge.
    Batch& bt = h->bt;
    bt.sync_state();
    SegState& st = bt.h_state[0];
    const int64_t n = st.n[st.cur];
    if (n > st.app_cap_v) st.app_cap_v = std::max(2 * st.app_cap_v, n);
    bt.put_state();
    std::vector<int32_t> en;
    std::vector<StepDev> one(1);
    const StepSpan steps{one.data(), 1};
    make_steps(steps, bt.W, 1, 1, gz, gx, &cosv, &sinv, nullptr, en);
    bt.set_filters(nullptr, nullptr);
    // app_cap_v >= n now, so the sequence-start precheck cannot fire
    dispatch_wc(bt, [&](auto wc, auto ct) { bt.run_sequence<decltype(wc)::value, typename decltype(ct)::type>(steps, 1, en); });
    API_END
}

int pp_compact(pp_handle* h, const pp_filter* f) {
    if (!h || !f) return fail(PP_EARG, "null argument");
    API_BEGIN
    h->bt.set_filters(f, nullptr);
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.compact<decltype(wc)::value, typename decltype(ct)::type>({1}); });
    API_END
}

int pp_scale_by_popcount(pp_handle* h, int kind, const uint64_t* mask, const double* t0, const double* t1,
                         const double* t2) {
    if (!h || !mask || !t0) return fail(PP_EARG, "null argument");
    if (kind < 0 || kind > 4) return fail(PP_EARG, "kind must be in 0:4");
    if (kind == 4 && (!t1 || !t2)) return fail(PP_EARG, "general channel needs three tables");
    API_BEGIN
    Batch& bt = h->bt;
    std::vector<double> tabs(3 * kExpTab, 0.0);
    for (int k = 0; k <= h->N; ++k) {
        tabs[k] = t0[k];
        if (kind == 4) {
            tabs[kExpTab + k] = t1[k];
            tabs[2 * kExpTab + k] = t2[k];
        }
    }
    ChanDev ch{};
    ch.kind = kind;
    for (int w = 0; w < bt.W; ++w) ch.mask[w] = mask[w];
    ch.t0 = bt.d_tabs;
    ch.t1 = bt.d_tabs + kExpTab;
    ch.t2 = bt.d_tabs + 2 * kExpTab;
    dispatch_wc(bt, [&](auto wc, auto ct) { bt.channel<decltype(wc)::value, typename decltype(ct)::type>({ch}, tabs, {1}); });
    HIPCHK(hipStreamSynchronize(bt.stream));
    API_END
}

int pp_expectation_ket(pp_handle* h, const uint64_t* ket, double* out) {
    if (!h || !ket || !out) return fail(PP_EARG, "null argument");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.expectation<decltype(wc)::value, typename decltype(ct)::type>(ket, out); });
    API_END
}

int pp_copy(pp_handle* dst, pp_handle* src) {
    if (!dst || !src) return fail(PP_EARG, "null argument");
    if (dst == src) return PP_OK;
    if (dst->N != src->N) return fail(PP_EARG, "qubit counts differ");
    if (dst->bt.cw != src->bt.cw) return fail(PP_EARG, "coefficient types differ");
    API_BEGIN
    dst->bt.copy_from(src->bt);
    API_END
}

int pp_otimes(pp_handle* out, pp_handle* a, pp_handle* b) {
    if (!out || !a || !b) return fail(PP_EARG, "null argument");
    if (out == a || out == b) return fail(PP_EARG, "the result must be a third vector");
    if (out->N != a->N + b->N) return fail(PP_EARG, "the result needs N1 + N2 qubits");
    if (a->bt.cw != b->bt.cw || out->bt.cw != a->bt.cw) return fail(PP_EARG, "coefficient types differ");
    API_BEGIN
    dispatch_wc(out->bt, [&](auto wc, auto ct) {
        out->bt.otimes_from<decltype(wc)::value, typename decltype(ct)::type>(a->bt, a->N, b->bt);
    });
    API_END
}

int pp_largest(pp_handle* h, int diag_only, uint64_t* z, uint64_t* x, double* c, int64_t* index) {
    if (!h || !z || !x || !c || !index) return fail(PP_EARG, "null argument");
    API_BEGIN
    int64_t idx = -1;
    dispatch_wc(h->bt, [&](auto wc, auto ct) {
        idx = h->bt.largest<decltype(wc)::value, typename decltype(ct)::type>(diag_only ? 1 : 0);
    });
    if (idx < 0) {
        // largest: a vector of only NaN coefficients gives its first term (the
        // reference's bi = 1); an empty vector or no diagonal term is an error
        int64_t n = 0;
        h->bt.sync_state();
        n = h->bt.n_known[0];
        if (diag_only || n == 0) throw PPError(PP_EARG, "collection must be non-empty");
        idx = 0;
    }
    h->bt.fetch_term(idx, z, x, c);
    *index = idx;
    API_END
}

int pp_select_diag(pp_handle* h, int offdiag) {
    if (!h) return fail(PP_EARG, "null handle");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) {
        h->bt.select_diag<decltype(wc)::value, typename decltype(ct)::type>(offdiag ? 1 : 0);
    });
    API_END
}

int pp_matrix(pp_handle* h, double* out) {
    if (!h || !out) return fail(PP_EARG, "null argument");
    if (h->N > 13) return fail(PP_EARG, "Matrix(v) is limited to N <= 13 (2^26 entries)");
    API_BEGIN
    dispatch_wc(h->bt, [&](auto wc, auto ct) {
        h->bt.matrix<decltype(wc)::value, typename decltype(ct)::type>(h->N, out);
    });
    API_END
}

int pp_check(pp_handle* h, int* ok) {
    if (!h || !ok) return fail(PP_EARG, "null argument");
    API_BEGIN
    int32_t o = 0;
    dispatch_wc(h->bt, [&](auto wc, auto ct) { h->bt.check<decltype(wc)::value, typename decltype(ct)::type>(&o); });
    *ok = o;
    API_END
}

// --- batches -----------------------------------------------------------------

int pp_batch_create(int B, const int* N, const int64_t* live_cap, const int64_t* append_cap, pp_batch** out) {
    if (!out || !N || !live_cap || !append_cap) return fail(PP_EARG, "null argument");
    *out = nullptr;
    if (B < 1 || B >= (1 << pp::kBlkVecBits)) return fail(PP_EARG, "B must be in [1, 2^20)");
    int maxN = 0;
    for (int b = 0; b < B; ++b) {
        if (N[b] < 1 || N[b] > 128) return fail(PP_EARG, "N must be in 1:128");
        if (live_cap[b] < 1 || append_cap[b] < 1) return fail(PP_EARG, "capacities must be >= 1");
        maxN = std::max(maxN, N[b]);
    }
    pp_batch* p = new pp_batch();
    try {
        p->bt.init(maxN <= 64 ? 1 : 2, std::vector<int>(N, N + B), std::vector<int64_t>(live_cap, live_cap + B),
                   std::vector<int64_t>(append_cap, append_cap + B));
    } catch (const PPError& e) {
        delete p;
        return fail(e.code, e.msg);
    }
    *out = p;
    g_last_error.clear();
    return PP_OK;
}

int pp_batch_destroy(pp_batch* p) {
    if (!p) return PP_OK;
    (void)hipStreamSynchronize(p->bt.stream);
    delete p;
    return PP_OK;
}

int pp_batch_words(pp_batch* p, int* W) {
    if (!p || !W) return fail(PP_EARG, "null argument");
    *W = p->bt.W;
    return PP_OK;
}

int pp_batch_upload(pp_batch* p, int b, const uint64_t* z, const uint64_t* x, const double* c, int64_t n, int sorted) {
    if (!p) return fail(PP_EARG, "null handle");
    if (b < 0 || b >= p->bt.B) return fail(PP_EARG, "vector index out of range");
    if (n < 0 || (n > 0 && (!z || !x || !c))) return fail(PP_EARG, "bad arrays");
    API_BEGIN
    p->bt.upload(b, z, x, c, n, sorted != 0);
    API_END
}

int pp_batch_download(pp_batch* p, int b, uint64_t* z, uint64_t* x, double* c, int64_t cap, int64_t* n) {
    if (!p || !n) return fail(PP_EARG, "null argument");
    if (b < 0 || b >= p->bt.B) return fail(PP_EARG, "vector index out of range");
    API_BEGIN
    *n = p->bt.download(b, z, x, c, cap);
    API_END
}

int pp_batch_sizes(pp_batch* p, int64_t* n) {
    if (!p || !n) return fail(PP_EARG, "null argument");
    API_BEGIN
    p->bt.sync_state();
    for (int b = 0; b < p->bt.B; ++b) n[b] = p->bt.n_known[b];
    API_END
}

int pp_batch_evolve(pp_batch* p, int64_t nsteps, const uint64_t* gz, const uint64_t* gx, const double* cosv,
                    const double* sinv, const uint8_t* active, const pp_filter* strict, const pp_filter* local,
                    pp_counters* counters) {
    if (!p) return fail(PP_EARG, "null handle");
    if (nsteps < 0) return fail(PP_EARG, "nsteps must be >= 0");
    if (nsteps > 0 && (!gz || !gx || !cosv || !sinv)) return fail(PP_EARG, "null generator arrays");
    if (counters) std::memset(counters, 0, sizeof(pp_counters) * (size_t)p->bt.B);
    if (nsteps == 0) return PP_OK;
    API_BEGIN
    batch_evolve(p->bt, nsteps, gz, gx, cosv, sinv, active, strict, local, counters);
    API_END
}

int pp_batch_scale_by_popcount(pp_batch* p, const int* kind, const uint64_t* mask, const double* tables) {
    if (!p || !kind || !mask || !tables) return fail(PP_EARG, "null argument");
    API_BEGIN
    Batch& bt = p->bt;
    std::vector<double> tabs((size_t)bt.B * 3 * kExpTab, 0.0);
    std::vector<ChanDev> ch((size_t)bt.B);
    std::vector<int32_t> en((size_t)bt.B, 0);
    for (int b = 0; b < bt.B; ++b) {
        if (kind[b] < -1 || kind[b] > 4) throw PPError(PP_EARG, "kind must be in -1:4");
        ch[b].kind = kind[b];
        en[b] = kind[b] >= 0;
        for (int w = 0; w < bt.W; ++w) ch[b].mask[w] = mask[(size_t)b * bt.W + w];
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k <= bt.N[b]; ++k)
                tabs[((size_t)b * 3 + t) * kExpTab + k] = tables[((size_t)b * 3 + t) * 129 + k];
        ch[b].t0 = bt.d_tabs + ((size_t)b * 3 + 0) * kExpTab;
        ch[b].t1 = bt.d_tabs + ((size_t)b * 3 + 1) * kExpTab;
        ch[b].t2 = bt.d_tabs + ((size_t)b * 3 + 2) * kExpTab;
    }
    dispatch_wc(bt, [&](auto wc, auto ct) { bt.channel<decltype(wc)::value, typename decltype(ct)::type>(ch, tabs, en); });
    HIPCHK(hipStreamSynchronize(bt.stream));
    API_END
}

int pp_batch_reserve(pp_batch* p, const int64_t* cap) {
    if (!p || !cap) return fail(PP_EARG, "null argument");
    API_BEGIN
    Batch& bt = p->bt;
    bt.sync_state();
    for (int b = 0; b < bt.B; ++b) {
        if (cap[b] < 0) throw PPError(PP_EARG, "capacity must be >= 0");
        const SegState& st = bt.h_state[b];
        bt.ensure_caps(b, std::max<int64_t>(cap[b], 1), std::max<int64_t>(cap[b] / 2, 1), bt.ib[b].maxlev,
                       st.n[st.cur], st.cur);
    }
    bt.upload_desc();
    HIPCHK(hipStreamSynchronize(bt.stream));
    API_END
}

int pp_batch_set_tiles(pp_batch* p, int tile, int merge_tile) {
    if (!p) return fail(PP_EARG, "null handle");
    if (tile != 0 && tile != 512 && tile != 2048) return fail(PP_EARG, "tile must be 0, 512 or 2048");
    if (merge_tile != 0 && merge_tile != 512 && merge_tile != 1024 && merge_tile != 2048)
        return fail(PP_EARG, "merge_tile must be 0, 512, 1024 or 2048");
    if (merge_tile == 2048 && p->bt.W == 2) return fail(PP_EARG, "merge_tile 2048 needs N <= 64");
    p->bt.tile_req = tile;
    p->bt.mtile_req = merge_tile;
    return PP_OK;
}

int pp_batch_set_graphs(pp_batch* p, int on) {
    if (!p) return fail(PP_EARG, "null handle");
    p->bt.graphs = on != 0;
    return PP_OK;
}

int pp_batch_graph_stats(pp_batch* p, int64_t* captures, int64_t* launches) {
    if (!p || !captures || !launches) return fail(PP_EARG, "null argument");
    *captures = p->bt.graph_captures;
    *launches = p->bt.graph_launches;
    return PP_OK;
}

int pp_batch_set_timing(pp_batch* p, int on) {
    if (!p) return fail(PP_EARG, "null handle");
    p->bt.timing = on != 0;
    return PP_OK;
}

int pp_batch_merge_time(pp_batch* p, double* seconds, int64_t* launches, int reset) {
    if (!p || !seconds || !launches) return fail(PP_EARG, "null argument");
    *seconds = p->bt.t_merge;
    *launches = p->bt.merge_launches;
    if (reset) {
        p->bt.t_merge = 0.0;
        p->bt.merge_launches = 0;
    }
    return PP_OK;
}

static int set_resident(Batch& bt, int mode, int cap_limit) {
    if (mode < 0 || mode > 2) return fail(PP_EARG, "resident mode must be 0, 1 or 2");
    if (cap_limit < 0) return fail(PP_EARG, "cap_limit must be >= 0");
    bt.res_mode = mode;
    bt.res_cap_limit = cap_limit;
    g_last_error.clear();
    return PP_OK;
}

int pp_batch_set_resident(pp_batch* p, int mode, int cap_limit) {
    if (!p) return fail(PP_EARG, "null handle");
    return set_resident(p->bt, mode, cap_limit);
}

int pp_set_resident(pp_handle* h, int mode, int cap_limit) {
    if (!h) return fail(PP_EARG, "null handle");
    return set_resident(h->bt, mode, cap_limit);
}

// Look-back tags: move the base forward (tests of the wrap near 2^22).
static int set_epoch(Batch& bt, uint32_t e) {
    if (e == 0 || (uint64_t)e >= kEpochMask) return fail(PP_EARG, "epoch must be in 1 .. 2^22 - 2");
    API_BEGIN
    bt.sync_state();
    if ((uint64_t)e < bt.epoch_used) throw PPError(PP_EARG, "the look-back tag base only moves forward");
    bt.epoch_used = e;
    HIPCHK(hipMemcpyAsync(bt.d_epoch, &e, sizeof(uint32_t), hipMemcpyHostToDevice, bt.stream));
    if (e > bt.win.epoch) bt.win.epoch = e;
    HIPCHK(hipStreamSynchronize(bt.stream));
    API_END
}

int pp_batch_set_epoch(pp_batch* p, uint32_t epoch) {
    if (!p) return fail(PP_EARG, "null handle");
    return set_epoch(p->bt, epoch);
}

int pp_set_epoch(pp_handle* h, uint32_t epoch) {
    if (!h) return fail(PP_EARG, "null handle");
    return set_epoch(h->bt, epoch);
}

int pp_batch_resident_stats(pp_batch* p, int64_t* launches, int64_t* fallbacks, int64_t* aborts) {
    if (!p || !launches || !fallbacks || !aborts) return fail(PP_EARG, "null argument");
    *launches = p->bt.res_launches;
    *fallbacks = p->bt.res_fallbacks;
    *aborts = p->bt.res_aborts;
    return PP_OK;
}

int pp_batch_resident_time(pp_batch* p, double* seconds, int reset) {
    if (!p || !seconds) return fail(PP_EARG, "null argument");
    API_BEGIN
    p->bt.sync_state();
    *seconds = p->bt.t_res;
    if (reset) p->bt.t_res = 0.0;
    API_END
}

int pp_batch_resident_work(pp_batch* p, int64_t* rotations, int64_t* swept, int64_t* merge_out, int64_t* rebalances,
                           int reset) {
    if (!p || !rotations || !swept || !merge_out || !rebalances) return fail(PP_EARG, "null argument");
    API_BEGIN
    Batch& bt = p->bt;
    unsigned long long w[4] = {0, 0, 0, 0};
    if (bt.d_rwork) {
        HIPCHK(hipMemcpyAsync(w, bt.d_rwork, sizeof(w), hipMemcpyDeviceToHost, bt.stream));
        if (reset) HIPCHK(hipMemsetAsync(bt.d_rwork, 0, sizeof(w), bt.stream));
        HIPCHK(hipStreamSynchronize(bt.stream));
    }
    *rotations = (int64_t)w[0];
    *swept = (int64_t)w[1];
    *merge_out = (int64_t)w[2];
    *rebalances = (int64_t)w[3];
    API_END
}

int pp_batch_expectation_ket(pp_batch* p, const uint64_t* kets, double* out) {
    if (!p || !kets || !out) return fail(PP_EARG, "null argument");
    API_BEGIN
    dispatch_wc(p->bt, [&](auto wc, auto ct) { p->bt.expectation<decltype(wc)::value, typename decltype(ct)::type>(kets, out); });
    API_END
}

}  // extern "C"
