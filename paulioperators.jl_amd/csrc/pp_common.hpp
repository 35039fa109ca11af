// Shared host/device definitions for the SparsePauliVector engine (gfx950).
#pragma once
#include <stdint.h>

namespace pp {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // 4 waves per workgroup
constexpr int kTile = 2048;        // sweep tile: 32 flag words, 8 elements per lane
constexpr int kTileWords = kTile / 64;
constexpr int kMaxLevels = 256;    // set bits of a combined (z, x) mask, N <= 128
constexpr int kExpTab = 257;       // exp(-alpha * w) for w = 0..256
constexpr int16_t kHdbNone = 0x3fff;  // "no neighbour": above every bit position

// Merge tile (outputs per workgroup) per word count: the LDS tile holds the L
// and C slices of one merge-path window (keys + coefficients).
template <int W> struct MergeTile;
template <> struct MergeTile<1> { static constexpr int kSize = 2048; };
template <> struct MergeTile<2> { static constexpr int kSize = 1024; };

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

}  // namespace pp
