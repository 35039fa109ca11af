// Shared host/device definitions for the SparsePauliVector engine (gfx950).
#pragma once
#include <stdint.h>

namespace pp {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // 4 waves per workgroup
constexpr int kTileMin = 512;      // smallest sweep tile (scratch is sized for it)
constexpr int kMaxLevels = 256;    // set bits of a combined (z, x) mask, N <= 128
constexpr int kExpTab = 257;       // exp(-alpha * w) for w = 0..256
constexpr int16_t kHdbNone = 0x3fff;  // "no neighbour": above every bit position

// Tile shapes are template parameters of the kernels: the sweep tile T
// (terms per workgroup in kb_flags/kb_emit: 512 or 2048) and the merge tile
// TM (outputs per workgroup in kb_merge: 512, 1024, 2048 for W = 1; 512, 1024
// for W = 2, whose LDS window holds TM x 40 B).  The host picks them per call
// from the autotuned table.

// One rotation, prepared on the host (spv_evolve.jl:333-343 computes the same
// packed generator, n_g = popcount(gz & gx), cos θ and sin θ once per generator).
struct GenDev {
    uint64_t gz[2];
    uint64_t gx[2];
    double cosv;
    double sinv;
    int32_t ng;
    int32_t nlev;                // set bits of the combined key mask (z above x)
    uint8_t lev[kMaxLevels];     // their combined-key bit positions
};

// MergeFilter (spv_kernels.jl:33-42) plus host-computed exp tables for the
// damped criteria: exptab_w[w] = exp((-alpha_w) * w).
struct FilterDev {
    int64_t wmax, xwmax, mwmax;
    double thresh, alpha_w, dthresh_w, alpha_xw, dthresh_xw;
    double exptab_w[kExpTab];
    double exptab_xw[kExpTab];
};

// Per-instance device buffers.  A batch holds B independent SparsePauliVectors
// of the same word count; the single-vector API is a batch of one.
struct SegDesc {
    uint64_t* Lz[2];      // live terms, ping-pong (W words per term)
    uint64_t* Lx[2];
    double* Lc[2];
    uint64_t* Cz;         // children of the current rotation, in key order
    uint64_t* Cx;
    double* Cc;
    uint64_t* flags;      // one bit per live term
    uint16_t* wpre;       // in-tile exclusive prefix per flag word
    int32_t* tilecnt;     // flagged terms per sweep tile
    int64_t* tileoff;     // exclusive scan of tilecnt (ntiles + 1 entries)
    int64_t* bnd;         // boundary segment counts, 4 per (tile, level)
    uint64_t* agg;        // segment summaries, 4 levels per sweep tile
    uint64_t* status;     // merge look-back words
    const struct FilterDev* fs;  // strict filter
    const struct FilterDev* fl;  // local filter
    int64_t cap;          // capacity of each live buffer (terms)
    int64_t capC;         // capacity of the child buffer (terms)
    int32_t maxlev;       // levels per tile held by bnd
    int32_t local_active; // local filter can drop children
};

// Per-instance device state, advanced by the kernels without host round trips.
struct SegState {
    int64_t n[2];         // live counts of Lz[0] / Lz[1]
    int64_t m;            // children of the current rotation
    int64_t app_cap_v;    // the reference's append-buffer length (decides early merges)
    int64_t live_cap_v;   // the reference's live-buffer length
    int64_t rotations, swept, created, merge_in, merge_out, early;  // WindowCounters sums
    int64_t stall_step;   // step at which the instance stalled
    int64_t need;         // size that did not fit
    int64_t nstep;        // live count entering the current rotation or compaction
    uint32_t ticket;      // merge tile ticket
    int32_t cur;          // live buffer parity
    int32_t flip;         // live parity = cur ^ flip (a merge toggles flip)
    int32_t stalled;      // 1: child buffer too small, 2: live buffer too small
    int32_t compact;      // early-merge compaction requested at this layer start
    int32_t err;          // internal consistency bits
};

// One rotation of one instance at one step.
struct StepDev {
    GenDev g;
    int32_t active;       // 0: the instance skips this step
    int32_t precheck;     // run the reference's capacity precheck before it
};

}  // namespace pp
