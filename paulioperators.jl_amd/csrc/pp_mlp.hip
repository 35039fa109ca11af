// libpauliprop: learned surrogate of <O>(theta) — an ensemble of E MLPs
// din -> H -> ... -> H -> 1 (GELU), trained with Adam on labels produced by
// the propagation engine.  BF16 operands on the matrix cores (gfx950
// v_mfma_f32_32x32x16_bf16), FP32 accumulation, FP32 master weights and
// optimizer state.  This part has no counterpart in the reference.
//
// Layouts (per member e, row-major):
//   input       X[b][KIN]          bf16, KIN = din rounded up to 64 (zero pad)
//   hidden l    W_l[H][K_l] bf16 copy of the FP32 master, b_l[H] fp32
//   output      w_o[H], b_o        fp32
//   activations Z_l[b][H] fp32 (pre-activation), A_l[b][H] bf16 = gelu(Z_l)
// Every GEMM is C[m][n] = sum_k A(m,k) B(n,k) with element strides, so the
// forward (X W^T), the input gradient (dZ W) and the weight gradient
// (dZ^T X) use one kernel; operands are staged through LDS in 64 x 32 tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <type_traits>

#include "../../include/pauliprop.h"

namespace pp {
void set_last_error(const std::string& msg);  // pp_last_error() of the library
}

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 64, kBN = 64, kBK = 32, kLDK = kBK + 8;
constexpr int kKin = 64;

struct MlpError {
    int code;
    std::string msg;
};

#define MCHK(expr)                                                                                   \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess)                                                                        \
            throw MlpError{_e == hipErrorOutOfMemory ? PP_ENOMEM : PP_EHIP,                          \
                           std::string(#expr) + ": " + hipGetErrorString(_e)};                       \
    } while (0)

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU through erf by Abramowitz & Stegun, Handbook of Mathematical
// Functions, 7.1.26 (|error| <= 1.5e-7 on the real line): for u >= 0,
// erf(u) = 1 - P(t) exp(-u^2), P(t) = a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5,
// t = 1 / (1 + p u); erf is odd.  With u = |x| / sqrt 2 and E = exp(-x^2 / 2):
//   gelu(x) = 0.5 x (1 + erf(x / sqrt 2)) = max(x, 0) - 0.5 |x| P E
//   gelu'(x) = 0.5 (1 + erf(x / sqrt 2)) + x E / sqrt(2 pi)
//            = [x >= 0] - sign(x) 0.5 P E + x E / sqrt(2 pi)
// (no cancellation for negative x).  The error is far below the bf16
// rounding (2^-9 relative) that every activation and gradient goes through.
// half_p(x) = 0.5 P(t) at u = |x| / sqrt 2 (coefficients pre-halved).
__device__ __forceinline__ float half_p(float x) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f * 0.70710678118654752f, fabsf(x), 1.0f));
    float q = fmaf(0.5f * 1.061405429f, t, 0.5f * -1.453152027f);
    q = fmaf(q, t, 0.5f * 1.421413741f);
    q = fmaf(q, t, 0.5f * -0.284496736f);
    q = fmaf(q, t, 0.5f * 0.254829592f);
    return q * t;
}
__device__ __forceinline__ float gauss(float x) { return __expf(-0.5f * x * x); }  // exp(-x^2 / 2)
__device__ __forceinline__ float gelu(float x) { return fmaf(-fabsf(x) * half_p(x), gauss(x), fmaxf(x, 0.0f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float hp = half_p(x);
    return fmaf(fmaf(x, 0.39894228040143268f, copysignf(hp, -x)), gauss(x), x >= 0.0f ? 1.0f : 0.0f);
}

struct GemmArgs {
    const uint16_t* A;
    int64_t sam, sak, sae;
    const uint16_t* B;
    int64_t sbn, sbk, sbe;
    int M, N, K;
    // EPI 0: C = acc (weight gradient).  EPI 1 (training forward) and 5 (its
    // last hidden layer): C = acc + bias (pre-activation), H = gelu of it.
    // EPI 2: H = acc * gelu'(Zp) (back-propagated dZ).  EPI 3 (inference):
    // H only.  EPI 4 (inference, last hidden layer): neither.
    float* C;
    int64_t ldc, sce;
    uint16_t* H;
    int64_t ldh, she;
    const float* bias;
    int64_t sbias;
    const float* Zp;
    int64_t ldz, sze;
    float* db;  // EPI 2: db[col] += the column's sum of the bf16 outputs (bias gradient)
    int64_t sdb;
    // EPI 4 / 5: the output layer fused in, yp[e][slot][row] = sum over the
    // slot's columns of bf16(gelu(z)) wo[col], one slot per wave's column
    // range
    float* yp;
    int64_t syp;
    const float* wo;
    int64_t swo;
    // split K (EPI 0): blockIdx.z = part * ... + e over kparts parts of K / kparts
    // each; part p's product goes to C + p scp
    int kparts;
    int64_t scp;
};

// Sums N per-lane values over the 32 lanes of each half wave (N = 16 or 32)
// by halving the held values at each exchange with lane ^ M (M = 16, 8, ..,
// 1): afterwards v[0] of lane l is the total of value index (l mod 32) >>
// (5 - log2 N), complete in the lanes whose low 5 - log2 N bits are zero.
template <int N, int M>
__device__ __forceinline__ void half_wave_transpose_sum(float* v, int lane) {
    if constexpr (M >= 1) {
        if constexpr (N > 1) {
            const bool hi = (lane & M) != 0;
#pragma unroll
            for (int k = 0; k < N / 2; ++k) {
                const float keep = hi ? v[k + N / 2] : v[k];
                const float send = hi ? v[k] : v[k + N / 2];
                v[k] = keep + __shfl_xor(send, M, 64);
            }
            half_wave_transpose_sum<N / 2, M / 2>(v, lane);
        } else {
            v[0] += __shfl_xor(v[0], M, 64);
            half_wave_transpose_sum<1, M / 2>(v, lane);
        }
    }
}

// Stage a 64 x 32 tile (rows r0.., k0..) of a bf16 operand into LDS as
// dst[row][k].  Either the k stride or the row stride is 1 (checked by the
// host); 8 contiguous elements per thread, 16-byte loads.
__device__ __forceinline__ void stage_tile(const uint16_t* __restrict__ src, int64_t srow, int64_t sk, int r0, int k0,
                                           uint16_t* __restrict__ dst) {
    const int t = threadIdx.x;
    if (sk == 1) {
        const int row = t >> 2, kc = (t & 3) * 8;
        const uint4 v = *reinterpret_cast<const uint4*>(src + (int64_t)(r0 + row) * srow + k0 + kc);
        *reinterpret_cast<uint4*>(dst + row * kLDK + kc) = v;
    } else {
        const int k = t >> 3, rc = (t & 7) * 8;
        const uint4 v = *reinterpret_cast<const uint4*>(src + (int64_t)(k0 + k) * sk + r0 + rc);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[(rc + j) * kLDK + k] = h[j];
    }
}

constexpr bool epi_act(int epi) { return epi == 1 || epi >= 3; }  // bias + GELU
constexpr bool epi_keep_z(int epi) { return epi == 1 || epi == 5; }
constexpr bool epi_keep_h(int epi) { return epi == 1 || epi == 3 || epi == 5; }
constexpr bool epi_out(int epi) { return epi >= 4; }

template <int EPI>
__global__ void __launch_bounds__(256) kg_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[kBM * kLDK];
    __shared__ __attribute__((aligned(16))) uint16_t sB[kBN * kLDK];
    const int e = blockIdx.z % (gridDim.z / g.kparts), part = blockIdx.z / (gridDim.z / g.kparts);
    const int kl = g.K / g.kparts, kb = part * kl;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const uint16_t* A = g.A + e * g.sae;
    const uint16_t* B = g.B + e * g.sbe;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = kb; k0 < kb + kl; k0 += kBK) {
        stage_tile(A, g.sam, g.sak, m0, k0, sA);
        stage_tile(B, g.sbn, g.sbk, n0, k0, sB);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 16) {
            // lane l supplies row (l mod 32), k = kk + 8 (l div 32) .. +7
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(sA + (wm * 32 + (lane & 31)) * kLDK + kk + 8 * (lane >> 5));
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(sB + (wn * 32 + (lane & 31)) * kLDK + kk + 8 * (lane >> 5));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator r of lane l: row 8 (r div 4) + 4 (l div 32) + (r mod 4), column l mod 32
    const int col = n0 + wn * 32 + (lane & 31);
    float dbs = 0.0f;
    float ys[16];
    const float wcol = epi_out(EPI) ? g.wo[e * g.swo + col] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const float v = acc[r];
        ys[r] = 0.0f;
        if (EPI == 0) {
            g.C[part * g.scp + e * g.sce + (int64_t)row * g.ldc + col] = v;
        } else if (epi_act(EPI)) {
            const float z = v + g.bias[e * g.sbias + col];
            if (epi_keep_z(EPI)) g.C[e * g.sce + (int64_t)row * g.ldc + col] = z;
            const uint16_t h = f2bf(gelu(z));
            if (epi_keep_h(EPI)) g.H[e * g.she + (int64_t)row * g.ldh + col] = h;
            ys[r] = bf2f(h) * wcol;
        } else {
            const float zp = g.Zp[e * g.sze + (int64_t)row * g.ldz + col];
            const uint16_t h = f2bf(v * gelu_grad(zp));
            g.H[e * g.she + (int64_t)row * g.ldh + col] = h;
            dbs += bf2f(h);
        }
    }
    if (EPI == 2 && g.db) {
        dbs += __shfl_xor(dbs, 32, 64);  // lanes l and l + 32 hold the same column
        if (lane < 32) atomicAdd(g.db + e * g.sdb + col, dbs);
    }
    if (epi_out(EPI)) {
        // slot = the wave's 32 columns; value r of lane l is row 8 (r div 4) + 4 (l div 32) + (r mod 4)
        half_wave_transpose_sum<16, 16>(ys, lane);
        const float t = ys[0];
        if ((lane & 1) == 0) {
            const int r = (lane & 31) >> 1;
            const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            g.yp[e * g.syp + (int64_t)((n0 + wn * 32) / 32) * g.M + row] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// kg_gemm2: the same product and epilogues for shapes that are multiples of
// 128 (M), 128 (N) and 64 (K).  A workgroup computes a 128 x 128 tile with
// four waves of 64 x 64 (2 x 2 blocks of v_mfma_f32_32x32x16_bf16, each
// fragment used twice; a 128 x 256 tile holds one workgroup per CU and
// measured 20% slower on 65536-row inference); K advances 64 per stage
// through one LDS stage of 36 KiB (three workgroups per CU; with two stages
// only two fit, measured 16% slower): the next stage's global loads are
// issued into registers before the current stage's MFMAs and written after
// a barrier that ends the stage's reads.  AK / BK: the
// operand's k stride is 1 (rows of 8 contiguous k per 16-byte load);
// otherwise its row stride is 1 and each thread loads a block of 8 rows x 4 k
// (four 16-byte loads), transposes it in registers and writes 8-byte rows to
// LDS, the 16 lanes of a group covering one row's 64 k (no bank conflict).
constexpr int kTM = 128, kTK = 64, kLDT = kTK + 8;  // LDS row: 72 bf16 (144 B)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// One operand tile of ROWS x 64 in flight: ROWS / 32 (k-contiguous) or
// ROWS / 128 (row-contiguous, four loads each) 16-byte loads per thread.
template <int ROWS>
struct Stage {
    u32x4 v[ROWS / 32];
};

template <bool KC, int ROWS>
__device__ __forceinline__ void stage_load(const uint16_t* __restrict__ src, int64_t srow, int64_t sk, int r0, int k0,
                                           Stage<ROWS>& st) {
    const int t = threadIdx.x;
    if constexpr (KC) {
#pragma unroll
        for (int i = 0; i < ROWS / 32; ++i) {
            const int c = t + 256 * i;  // ROWS x 8 chunks of 8 k
            const int row = c >> 3, kc = (c & 7) * 8;
            st.v[i] = *reinterpret_cast<const u32x4*>(src + (int64_t)(r0 + row) * srow + k0 + kc);
        }
    } else {
#pragma unroll
        for (int h = 0; h < ROWS / 128; ++h) {
            const int rb = ((t >> 4) + 16 * h) * 8, kb = (t & 15) * 4;  // row blocks of 8, k quads
#pragma unroll
            for (int i = 0; i < 4; ++i)
                st.v[4 * h + i] = *reinterpret_cast<const u32x4*>(src + (int64_t)(k0 + kb + i) * sk + r0 + rb);
        }
    }
}

template <bool KC, int ROWS>
__device__ __forceinline__ void stage_store(const Stage<ROWS>& st, uint16_t* __restrict__ dst) {
    const int t = threadIdx.x;
    if constexpr (KC) {
#pragma unroll
        for (int i = 0; i < ROWS / 32; ++i) {
            const int c = t + 256 * i;
            const int row = c >> 3, kc = (c & 7) * 8;
            *reinterpret_cast<u32x4*>(dst + row * kLDT + kc) = st.v[i];
        }
    } else {
#pragma unroll
        for (int h = 0; h < ROWS / 128; ++h) {
            const int rb = ((t >> 4) + 16 * h) * 8, kb = (t & 15) * 4;
            // st.v[4h + i] holds rows rb..rb+7 at k = kb + i; row r of the
            // output is (k0, k1) (k2, k3) as two 32-bit words
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int q = r >> 1, sh = (r & 1) * 16;
                const uint32_t a0 = (st.v[4 * h][q] >> sh) & 0xffffu, a1 = (st.v[4 * h + 1][q] >> sh) & 0xffffu;
                const uint32_t a2 = (st.v[4 * h + 2][q] >> sh) & 0xffffu, a3 = (st.v[4 * h + 3][q] >> sh) & 0xffffu;
                u32x2 o;
                o.x = a0 | (a1 << 16);
                o.y = a2 | (a3 << 16);
                *reinterpret_cast<u32x2*>(dst + (rb + r) * kLDT + kb) = o;
            }
        }
    }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

template <int EPI, bool AK, bool BK>
__global__ void __launch_bounds__(256, 3) kg_gemm2(GemmArgs g) {
    constexpr int TN = 128;
    // one LDS stage (A then B, 36 KiB): three workgroups per CU
    __shared__ __attribute__((aligned(16))) uint16_t smem[(kTM + TN) * kLDT];
    uint16_t* const sA = smem;
    uint16_t* const sB = smem + kTM * kLDT;
    // split K only for the weight gradient (EPI 0)
    const int kparts = EPI == 0 ? g.kparts : 1;
    const int e = EPI == 0 ? blockIdx.z % (gridDim.z / kparts) : blockIdx.z;
    const int part = EPI == 0 ? blockIdx.z / (gridDim.z / kparts) : 0;
    const int kb = part * (g.K / kparts);
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * TN;
    const uint16_t* A = g.A + e * g.sae;
    const uint16_t* B = g.B + e * g.sbe;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;  // the wave's 64 x 64 quadrant
    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int nk = g.K / kparts / kTK;
    Stage<kTM> pa;
    Stage<TN> pb;
    stage_load<AK, kTM>(A, g.sam, g.sak, m0, kb, pa);
    stage_load<BK, TN>(B, g.sbn, g.sbk, n0, kb, pb);
    stage_store<AK, kTM>(pa, sA);
    stage_store<BK, TN>(pb, sB);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) {
            stage_load<AK, kTM>(A, g.sam, g.sak, m0, kb + (kt + 1) * kTK, pa);
            stage_load<BK, TN>(B, g.sbn, g.sbk, n0, kb + (kt + 1) * kTK, pb);
        }
        const uint16_t* a_s = sA + (wm * 64 + (lane & 31)) * kLDT + 8 * (lane >> 5);
        const uint16_t* b_s = sB + (wn * 64 + (lane & 31)) * kLDT + 8 * (lane >> 5);
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            // lane l supplies row (l mod 32), k = kk + 8 (l div 32) .. +7
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(a_s + i * 32 * kLDT + kk);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(b_s + j * 32 * kLDT + kk);
            // B rows first: the product block is C transposed, so a lane's
            // accumulators run along a row of C (four adjacent columns each)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) {
            __syncthreads();  // every wave has read the stage
            stage_store<AK, kTM>(pa, sA);
            stage_store<BK, TN>(pb, sB);
        }
        __syncthreads();
    }
    // accumulator 4 c + q of lane l in block (i, j): row 64 wm + 32 i + (l mod
    // 32), column 64 wn + 32 j + 8 c + 4 (l div 32) + q of the tile.  bf16
    // outputs go through LDS (the stage's buffer, free after the last
    // barrier) so that the global stores are whole 256-byte rows.
    constexpr bool kOut = epi_out(EPI);
    constexpr bool kH = epi_keep_h(EPI) || EPI == 2;
    constexpr int kLDH = 136;  // bf16 per LDS row of the output tile (272 B)
    static_assert(kTM * kLDH <= (kTM + TN) * kLDT, "output tile exceeds the LDS stage");
    uint16_t* sH = smem;
    float ys[2] = {0.0f, 0.0f};  // fused output layer: this lane's row sums
    float cs[32];                // EPI 2: column sums over this lane's rows
#pragma unroll
    for (int q = 0; q < 32; ++q) cs[q] = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int row = m0 + wm * 64 + i * 32 + (lane & 31);
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int col = n0 + wn * 64 + j * 32 + 8 * c + 4 * (lane >> 5);
                floatx4 v;
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] = acc[i][j][4 * c + q];
                if constexpr (EPI == 0) {
                    *reinterpret_cast<floatx4*>(g.C + part * g.scp + e * g.sce + (int64_t)row * g.ldc + col) = v;
                } else if constexpr (epi_act(EPI)) {
                    const floatx4 z = v + *reinterpret_cast<const floatx4*>(g.bias + e * g.sbias + col);
                    if constexpr (epi_keep_z(EPI))
                        *reinterpret_cast<floatx4*>(g.C + e * g.sce + (int64_t)row * g.ldc + col) = z;
                    floatx4 a;
#pragma unroll
                    for (int q = 0; q < 4; ++q) a[q] = gelu(z[q]);
                    const bf16x4 h = __builtin_convertvector(a, bf16x4);
                    if constexpr (epi_keep_h(EPI))
                        *reinterpret_cast<bf16x4*>(sH + (row - m0) * kLDH + col - n0) = h;
                    if constexpr (kOut) {
                        const floatx4 w = *reinterpret_cast<const floatx4*>(g.wo + e * g.swo + col);
#pragma unroll
                        for (int q = 0; q < 4; ++q) ys[i] += (float)h[q] * w[q];
                    }
                } else {
                    const floatx4 zp = *reinterpret_cast<const floatx4*>(g.Zp + e * g.sze + (int64_t)row * g.ldz + col);
                    floatx4 d;
#pragma unroll
                    for (int q = 0; q < 4; ++q) d[q] = v[q] * gelu_grad(zp[q]);
                    const bf16x4 h = __builtin_convertvector(d, bf16x4);
                    *reinterpret_cast<bf16x4*>(sH + (row - m0) * kLDH + col - n0) = h;
#pragma unroll
                    for (int q = 0; q < 4; ++q) cs[16 * j + 4 * c + q] += (float)h[q];
                }
            }
    }
    if constexpr (kH) {
        __syncthreads();
        const int t = threadIdx.x;
#pragma unroll
        for (int pass = 0; pass < 8; ++pass) {
            const int row = pass * 16 + (t >> 4), cc = (t & 15) * 8;
            *reinterpret_cast<u32x4*>(g.H + e * g.she + (int64_t)(m0 + row) * g.ldh + n0 + cc) =
                *reinterpret_cast<const u32x4*>(sH + row * kLDH + cc);
        }
    }
    if constexpr (kOut) {
        // lanes l and l + 32 hold the two halves of the wave's 64 columns;
        // slot = the wave's column range
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float t = ys[i] + __shfl_xor(ys[i], 32, 64);
            if (lane < 32)
                g.yp[e * g.syp + (int64_t)((n0 + wn * 64) / 64) * g.M + m0 + wm * 64 + i * 32 + lane] = t;
        }
    }
    if constexpr (EPI == 2) {
        if (g.db) {
            // value 16 j + 4 c + q: column 64 wn + 32 j + 8 c + 4 (l div 32) + q;
            // summed over the 32 rows of each half wave
            half_wave_transpose_sum<32, 16>(cs, lane);
            const int v = lane & 31;
            const int col = n0 + wn * 64 + 32 * (v >> 4) + 8 * ((v >> 2) & 3) + 4 * (lane >> 5) + (v & 3);
            atomicAdd(g.db + e * g.sdb + col, cs[0]);
        }
    }
}

// Minibatch gather: Xb[e][i][:] = X[perm[e][off + i]][:], Yb[e][i] likewise.
__global__ void kg_gather(const uint16_t* __restrict__ X, const float* __restrict__ Y, const int32_t* __restrict__ perm,
                          int64_t n, int64_t off, int Bsz, uint16_t* __restrict__ Xb, float* __restrict__ Yb) {
    const int e = blockIdx.y;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int c = threadIdx.x & 63;
    if (i >= Bsz) return;
    const int64_t src = perm[(int64_t)e * n + off + i];
    Xb[((int64_t)e * Bsz + i) * kKin + c] = X[src * kKin + c];
    if (c == 0) Yb[(int64_t)e * Bsz + i] = Y[src];
}

// Split-K sum: dst[e][i] = sum over parts (in order) of ws[part][e][i], i < cnt
// (a multiple of 4; dst rows contiguous).
__global__ void kg_splitk_sum(const float* __restrict__ ws, int parts, int E, int64_t cnt, float* __restrict__ dst,
                              int64_t sdst) {
    const int e = blockIdx.y;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= cnt) return;
    floatx4 acc = *reinterpret_cast<const floatx4*>(ws + e * cnt + i);
    for (int q = 1; q < parts; ++q) acc += *reinterpret_cast<const floatx4*>(ws + ((int64_t)q * E + e) * cnt + i);
    *reinterpret_cast<floatx4*>(dst + e * sdst + i) = acc;
}

// Zero the gradient slots the backward pass accumulates into with atomics
// (every hidden bias, the output weights and bias) and the minibatch loss;
// the weight gradients are stored whole by their GEMMs.
struct AccSlots {
    int n;
    int64_t off[17], len[17];
};
__global__ void kg_zero_acc(float* __restrict__ g, int64_t P, AccSlots sl, float* __restrict__ loss) {
    const int e = blockIdx.y, r = blockIdx.x;
    if (r < sl.n)
        for (int64_t i = threadIdx.x; i < sl.len[r]; i += blockDim.x) g[e * P + sl.off[r] + i] = 0.0f;
    if (r == 0 && threadIdx.x == 0) loss[e] = 0.0f;
}

// The output layer's value from the last hidden GEMM's partial sums:
// y = sum over slots (in slot order) of yp[e][slot][row], + bo[e].
__device__ __forceinline__ float out_value(const float* __restrict__ yp, int nslot, int64_t syp, int64_t M,
                                           const float* __restrict__ bo, int64_t sbo, int e, int64_t row) {
    float s = 0.0f;
    for (int q = 0; q < nslot; ++q) s += yp[e * syp + q * M + row];
    return s + bo[e * sbo];
}

// dy = 2 (y - t) / B (mean squared error per member), loss += (y - t)^2 / B,
// the output bias gradient db_o += sum_b dy and the epoch's loss sum.
__global__ void kg_loss(const float* __restrict__ yp, int nslot, int64_t syp, const float* __restrict__ bo,
                        int64_t sbo, const float* __restrict__ t, int Bsz, float* __restrict__ dy,
                        float* __restrict__ loss, float* __restrict__ dbo, int64_t sdbo, double* __restrict__ epoch) {
    const int e = blockIdx.y;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f, d = 0.0f;
    if (b < Bsz) {
        const float r = out_value(yp, nslot, syp, Bsz, bo, sbo, e, b) - t[(int64_t)e * Bsz + b];
        d = 2.0f * r / (float)Bsz;
        dy[(int64_t)e * Bsz + b] = d;
        l = r * r / (float)Bsz;
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        l += __shfl_xor(l, s, 64);
        d += __shfl_xor(d, s, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(loss + e, l);
        atomicAdd(dbo + e * sdbo, d);
        atomicAdd(epoch + e, (double)l);
    }
}

// The output layer's backward in one pass over [B][H] per member: dZ_L =
// bf16(dy w_o gelu'(Z_L)), dw_o[o] += sum_b dy[b] A_L[b][o] and the last
// hidden bias gradient db[o] += sum_b dZ_L[b][o] (of the stored bf16 values).
__global__ void kg_out_bwd(const float* __restrict__ dy, const float* __restrict__ w, int64_t swe,
                           const float* __restrict__ Z, const uint16_t* __restrict__ A, int H, int Bsz, int rows_per,
                           uint16_t* __restrict__ dZ, float* __restrict__ gwo, float* __restrict__ gb, int64_t sg) {
    const int e = blockIdx.z;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= H) return;
    const int b0 = blockIdx.y * rows_per;
    const int b1 = min(Bsz, b0 + rows_per);
    const float wo = w[e * swe + o];
    float sw = 0.0f, sb = 0.0f;
    for (int b = b0; b < b1; ++b) {
        const float d = dy[(int64_t)e * Bsz + b];
        const int64_t q = ((int64_t)e * Bsz + b) * H + o;
        sw += bf2f(A[q]) * d;
        const uint16_t h = f2bf(d * wo * gelu_grad(Z[q]));
        dZ[q] = h;
        sb += bf2f(h);
    }
    atomicAdd(gwo + e * sg + o, sw);
    atomicAdd(gb + e * sg + o, sb);
}

// Adam and the bf16 copies of the hidden weight matrices in one pass: the
// master layout of W_l[H][K_l] per member is the bf16 copy's layout.
struct CastMap {
    int L;
    int64_t off[16], size[16];
    uint16_t* dst[16];
};
__global__ void kg_adam_cast(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                             float* __restrict__ v, int64_t n, int64_t P, float lr, float b1, float b2, float eps,
                             float c1, float c2, CastMap cm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    const int64_t e = i / P, off = i - e * P;
    for (int l = 0; l < cm.L; ++l)
        if (off >= cm.off[l] && off < cm.off[l] + cm.size[l]) cm.dst[l][e * cm.size[l] + (off - cm.off[l])] = f2bf(pi);
}

// bf16 copy of a weight block [rows][cols] (row stride ld in the master).
__global__ void kg_cast(const float* __restrict__ src, int64_t ld_src, int64_t se_src, uint16_t* __restrict__ dst,
                        int64_t ld_dst, int64_t se_dst, int rows, int cols) {
    const int e = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    dst[e * se_dst + (int64_t)r * ld_dst + c] = f2bf(src[e * se_src + (int64_t)r * ld_src + c]);
}

// fp32 [n][din] -> bf16 [n][kKin] with zero padding
__global__ void kg_pack_input(const float* __restrict__ x, int64_t n, int din, uint16_t* __restrict__ X) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kKin) return;
    const int64_t r = i / kKin;
    const int c = (int)(i % kKin);
    X[i] = c < din ? f2bf(x[r * din + c]) : (uint16_t)0;
}

// Each member's output y[e][i] (from the fused partial sums), then mean
// and (population) standard deviation over the E members of each row.
__global__ void kg_ensemble(const float* __restrict__ yp, int nslot, int64_t syp, const float* __restrict__ bo,
                            int64_t sbo, int E, int64_t n, int64_t stride, float* __restrict__ y,
                            double* __restrict__ mean, double* __restrict__ sd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int e = 0; e < E; ++e) {
        const float v = out_value(yp, nslot, syp, n, bo, sbo, e, i);
        y[e * stride + i] = v;
        s += (double)v;
    }
    const double mu = s / E;
    double q = 0.0;
    for (int e = 0; e < E; ++e) {
        const double d = (double)y[e * stride + i] - mu;
        q += d * d;
    }
    mean[i] = mu;
    sd[i] = sqrt(q / E);
}

template <typename T>
T* mal(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    MCHK(hipMalloc(&p, count * sizeof(T)));
    return static_cast<T*>(p);
}

int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

// ---------------------------------------------------------------------------
struct pp_mlp {
    int E = 0, din = 0, H = 0, L = 0;  // L hidden layers
    hipStream_t stream = nullptr;
    // FP32 master layout per member (padded): W_1[H][kKin], b_1[H],
    // then (W_l[H][H], b_l[H]) for l = 2..L, then w_o[H], b_o[1]
    std::vector<int64_t> offW, offb;  // per hidden layer
    int64_t offwo = 0, offbo = 0, P = 0;  // P: floats per member
    float* p = nullptr;
    float* g = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    int64_t step = 0;
    std::vector<uint16_t*> Wb;  // bf16 copies per hidden layer: [E][H][K_l]
    // data
    uint16_t* X = nullptr;
    float* Y = nullptr;
    int64_t n = 0;
    int32_t* perm = nullptr;
    // minibatch workspace (capacity cap_b rows)
    int cap_b = 0;
    uint16_t* Xb = nullptr;
    float* Yb = nullptr;
    std::vector<float*> Z;       // [E][B][H] per layer
    std::vector<uint16_t*> A;    // [E][B][H] per layer
    uint16_t* dZa = nullptr;     // [E][B][H] ping
    uint16_t* dZb = nullptr;     // [E][B][H] pong
    float* y = nullptr;  // [E][H / 32][cap_b] output-layer partial sums
    float* dy = nullptr;
    float* loss = nullptr;
    double* loss_acc = nullptr;  // [E] per-epoch sum of the minibatch losses
    // inference workspace
    int64_t cap_q = 0;
    uint16_t* Xq = nullptr;
    uint16_t* Qa = nullptr;
    uint16_t* Qb = nullptr;
    float* yq = nullptr;
    float* ypq = nullptr;  // [E][H / 32][cap_q] output-layer partial sums
    float* ws = nullptr;   // split-K partial products of the weight gradients
    size_t ws_cap = 0;
    int ncu = 256;         // compute units of the device
    float* xq_f = nullptr;
    double* mq = nullptr;
    double* sq = nullptr;

    int K(int l) const { return l == 0 ? kKin : H; }

    ~pp_mlp() {
        auto F = [](void* q) {
            if (q) (void)hipFree(q);
        };
        F(p); F(g); F(m); F(v);
        for (auto q : Wb) F(q);
        F(X); F(Y); F(perm); F(Xb); F(Yb);
        for (auto q : Z) F(q);
        for (auto q : A) F(q);
        F(dZa); F(dZb); F(y); F(dy); F(loss); F(loss_acc);
        F(Xq); F(Qa); F(Qb); F(yq); F(ypq); F(xq_f); F(mq); F(sq); F(ws);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void init(int E_, int din_, int H_, int L_) {
        E = E_; din = din_; H = H_; L = L_;
        MCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        {
            int dev = 0;
            MCHK(hipGetDevice(&dev));
            hipDeviceProp_t prop;
            MCHK(hipGetDeviceProperties(&prop, dev));
            if (prop.multiProcessorCount > 0) ncu = prop.multiProcessorCount;
        }
        int64_t off = 0;
        for (int l = 0; l < L; ++l) {
            offW.push_back(off);
            off += (int64_t)H * K(l);
            offb.push_back(off);
            off += H;
        }
        offwo = off;
        off += H;
        offbo = off;
        off += 1;
        P = (off + 63) / 64 * 64;
        p = mal<float>((size_t)E * P);
        g = mal<float>((size_t)E * P);
        m = mal<float>((size_t)E * P);
        v = mal<float>((size_t)E * P);
        MCHK(hipMemsetAsync(p, 0, sizeof(float) * E * P, stream));
        MCHK(hipMemsetAsync(g, 0, sizeof(float) * E * P, stream));  // padding stays zero
        MCHK(hipMemsetAsync(m, 0, sizeof(float) * E * P, stream));
        MCHK(hipMemsetAsync(v, 0, sizeof(float) * E * P, stream));
        for (int l = 0; l < L; ++l) Wb.push_back(mal<uint16_t>((size_t)E * H * K(l)));
        loss = mal<float>(E);
        loss_acc = mal<double>(E);
        MCHK(hipStreamSynchronize(stream));
    }

    void ensure_batch(int Bsz) {
        if (cap_b >= Bsz) return;
        auto F = [](void* q) {
            if (q) (void)hipFree(q);
        };
        F(Xb); F(Yb); F(dZa); F(dZb); F(y); F(dy);
        for (auto q : Z) F(q);
        for (auto q : A) F(q);
        Z.clear();
        A.clear();
        cap_b = Bsz;
        const size_t act = (size_t)E * Bsz * H;
        Xb = mal<uint16_t>((size_t)E * Bsz * kKin);
        Yb = mal<float>((size_t)E * Bsz);
        for (int l = 0; l < L; ++l) {
            Z.push_back(mal<float>(act));
            A.push_back(mal<uint16_t>(act));
        }
        dZa = mal<uint16_t>(act);
        dZb = mal<uint16_t>(act);
        y = mal<float>((size_t)E * (H / 32) * Bsz);  // output-layer partial sums [E][slot][Bsz]
        dy = mal<float>((size_t)E * Bsz);
    }

    void cast_weights() {
        for (int l = 0; l < L; ++l) {
            const int64_t cnt = (int64_t)H * K(l);
            hipLaunchKernelGGL(kg_cast, dim3((unsigned)cdiv64(cnt, 256), E), dim3(256), 0, stream, p + offW[l],
                               (int64_t)K(l), P, Wb[l], (int64_t)K(l), (int64_t)H * K(l), H, K(l));
        }
        MCHK(hipGetLastError());
    }

    // kg_gemm2 takes the shape (else kg_gemm)
    static bool wide_tiles(int M, int N, int Kd) { return M % kTM == 0 && N % 128 == 0 && Kd % kTK == 0; }
    // output-layer partial sums per row of the last hidden GEMM (M x H):
    // one per wave's column range, 64 columns in kg_gemm2, 32 in kg_gemm
    int out_slots(int M) const { return wide_tiles(M, H, K(L - 1)) ? H / 64 : H / 32; }

    template <int EPI>
    void gemm(const uint16_t* Aop, int64_t sam, int64_t sak, int64_t sae, const uint16_t* Bop, int64_t sbn, int64_t sbk,
              int64_t sbe, int M, int N, int Kd, float* C, int64_t ldc, int64_t sce, uint16_t* Hout, int64_t ldh,
              int64_t she, const float* bias, int64_t sbias, const float* Zp, int64_t ldz, int64_t sze,
              float* db = nullptr, int64_t sdb = 0, float* yp = nullptr, int64_t syp = 0, const float* wo = nullptr,
              int64_t swo = 0) {
        if (M % kBM || N % kBN || Kd % kBK) throw MlpError{PP_EARG, "GEMM shape not a multiple of the tile"};
        if ((sak != 1 && sam != 1) || (sbk != 1 && sbn != 1)) throw MlpError{PP_EARG, "GEMM operand not contiguous"};
        GemmArgs a{Aop, sam, sak, sae, Bop, sbn, sbk, sbe, M, N, Kd, C, ldc, sce, Hout, ldh, she, bias, sbias, Zp, ldz, sze,
                   db, sdb, yp, syp, wo, swo, 1, 0};
        const bool wide = wide_tiles(M, N, Kd);
        // a weight gradient with fewer workgroups than CUs (K = the batch)
        // splits K into parts written to a workspace and summed in order
        int parts = 1;
        if (EPI == 0 && ldc == N) {
            const int64_t wgs = wide ? (int64_t)(N / 128) * (M / kTM) * E : (int64_t)(N / kBN) * (M / kBM) * E;
            const int kt = wide ? kTK : kBK;
            while (wgs * parts < ncu && Kd % (2 * parts * kt) == 0 && Kd / (2 * parts) >= 2 * kt) parts *= 2;
        }
        if (parts > 1) {
            const size_t need = (size_t)parts * E * M * N;
            if (ws_cap < need) {
                if (ws) (void)hipFree(ws);
                ws = nullptr;
                ws_cap = 0;
                ws = mal<float>(need);
                ws_cap = need;
            }
            a.C = ws;
            a.sce = (int64_t)M * N;
            a.scp = (int64_t)E * M * N;
            a.kparts = parts;
        }
        if (wide) {
            const dim3 grid(N / 128, M / kTM, E * parts);
            if (sak == 1 && sbk == 1)
                hipLaunchKernelGGL((kg_gemm2<EPI, true, true>), grid, dim3(256), 0, stream, a);
            else if (sak == 1)
                hipLaunchKernelGGL((kg_gemm2<EPI, true, false>), grid, dim3(256), 0, stream, a);
            else if (sbk == 1)
                hipLaunchKernelGGL((kg_gemm2<EPI, false, true>), grid, dim3(256), 0, stream, a);
            else
                hipLaunchKernelGGL((kg_gemm2<EPI, false, false>), grid, dim3(256), 0, stream, a);
        } else {
            hipLaunchKernelGGL(kg_gemm<EPI>, dim3(N / kBN, M / kBM, E * parts), dim3(256), 0, stream, a);
        }
        if (parts > 1) {
            const int64_t cnt = (int64_t)M * N;
            hipLaunchKernelGGL(kg_splitk_sum, dim3((unsigned)cdiv64(cnt / 4, 256), E), dim3(256), 0, stream, ws, parts, E,
                               cnt, C, sce);
        }
        MCHK(hipGetLastError());
    }

    // Forward over a minibatch already gathered in Xb (train: keep Z).
    void forward_train(int Bsz) {
        const uint16_t* in = Xb;
        int64_t ld_in = kKin;
        for (int l = 0; l < L; ++l) {
            auto run = [&](auto epi) {
                gemm<decltype(epi)::value>(in, ld_in, 1, (int64_t)Bsz * ld_in, Wb[l], K(l), 1, (int64_t)H * K(l), Bsz,
                                           H, K(l), Z[l], H, (int64_t)Bsz * H, A[l], H, (int64_t)Bsz * H, p + offb[l],
                                           P, nullptr, 0, 0, nullptr, 0, y, (int64_t)out_slots(Bsz) * Bsz, p + offwo,
                                           P);
            };
            // the last hidden layer's epilogue also forms the output layer's partial sums
            if (l == L - 1) run(std::integral_constant<int, 5>());
            else run(std::integral_constant<int, 1>());
            in = A[l];
            ld_in = H;
        }
    }

    // Loss and gradients into g (its atomically accumulated slots zeroed
    // first, the weight gradients stored whole), for the minibatch in Xb / Yb.
    void backward(int Bsz) {
        {
            AccSlots sl{};
            sl.n = L + 1;
            for (int l = 0; l < L; ++l) {
                sl.off[l] = offb[l];
                sl.len[l] = H;
            }
            sl.off[L] = offwo;  // w_o and b_o are adjacent
            sl.len[L] = offbo + 1 - offwo;
            hipLaunchKernelGGL(kg_zero_acc, dim3((unsigned)sl.n, E), dim3(256), 0, stream, g, P, sl, loss);
        }
        // dy, loss, db_o and the epoch loss; then the output layer's dZ with
        // dw_o and the last hidden layer's bias gradient
        const int ns = out_slots(Bsz);
        hipLaunchKernelGGL(kg_loss, dim3((unsigned)cdiv64(Bsz, 256), E), dim3(256), 0, stream, y, ns,
                           (int64_t)ns * Bsz, p + offbo, P, Yb, Bsz, dy, loss, g + offbo, P, loss_acc);
        const int rows = 16;
        const int64_t act = (int64_t)Bsz * H;
        hipLaunchKernelGGL(kg_out_bwd, dim3((unsigned)cdiv64(H, 256), (unsigned)cdiv64(Bsz, rows), E), dim3(256), 0,
                           stream, dy, p + offwo, P, Z[L - 1], A[L - 1], H, Bsz, rows, dZa, g + offwo,
                           g + offb[L - 1], P);
        MCHK(hipGetLastError());
        uint16_t* dz = dZa;
        uint16_t* dzn = dZb;
        for (int l = L - 1; l >= 0; --l) {
            const uint16_t* in = l == 0 ? Xb : A[l - 1];
            const int64_t ld_in = l == 0 ? kKin : H;
            // dW_l[o][i] = sum_b dZ[b][o] in[b][i]: M = H (o), N = K_l (i), K = B
            gemm<0>(dz, 1, H, act, in, 1, ld_in, (int64_t)Bsz * ld_in, H, K(l), Bsz, g + offW[l], K(l), P, nullptr, 0,
                    0, nullptr, 0, nullptr, 0, 0);
            if (l > 0) {
                // dZ_{l-1}[b][i] = (sum_o dZ[b][o] W_l[o][i]) gelu'(Z_{l-1}[b][i]): M = B, N = H (i), K = H (o);
                // its epilogue adds db_{l-1} = sum_b dZ_{l-1}[b][:]
                gemm<2>(dz, H, 1, act, Wb[l], 1, K(l), (int64_t)H * K(l), Bsz, H, H, nullptr, 0, 0, dzn, H, act,
                        nullptr, 0, Z[l - 1], H, act, g + offb[l - 1], P);
                std::swap(dz, dzn);
            }
        }
    }

    void adam(float lr) {
        step += 1;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.0f - (float)std::pow((double)b1, (double)step);
        const float c2 = 1.0f - (float)std::pow((double)b2, (double)step);
        const int64_t cnt = (int64_t)E * P;
        CastMap cm{};
        cm.L = L;
        for (int l = 0; l < L; ++l) {
            cm.off[l] = offW[l];
            cm.size[l] = (int64_t)H * K(l);
            cm.dst[l] = Wb[l];
        }
        hipLaunchKernelGGL(kg_adam_cast, dim3((unsigned)cdiv64(cnt, 256)), dim3(256), 0, stream, p, g, m, v, cnt, P, lr,
                           b1, b2, eps, c1, c2, cm);
        MCHK(hipGetLastError());
    }

    void ensure_query(int64_t cq) {
        if (cap_q >= cq) return;
        auto F = [](void* q) {
            if (q) (void)hipFree(q);
        };
        F(Xq); F(Qa); F(Qb); F(yq); F(ypq); F(xq_f); F(mq); F(sq);
        cap_q = cq;
        Xq = mal<uint16_t>((size_t)cq * kKin);
        Qa = mal<uint16_t>((size_t)E * cq * H);
        Qb = mal<uint16_t>((size_t)E * cq * H);
        yq = mal<float>((size_t)E * cq);
        ypq = mal<float>((size_t)E * (H / 32) * cq);
        xq_f = mal<float>((size_t)cq * din);
        mq = mal<double>((size_t)cq);
        sq = mal<double>((size_t)cq);
    }

    // Forward of `cnt` rows (multiple of 64) of Xq; the output layer's
    // partial sums land in ypq[E][slot][cnt].
    void forward_query(int64_t cnt) {
        const uint16_t* in = Xq;
        int64_t ld_in = kKin, se_in = 0;  // the input is shared by every member
        uint16_t* outs[2] = {Qa, Qb};
        for (int l = 0; l < L; ++l) {
            uint16_t* o = outs[l & 1];
            auto run = [&](auto epi) {
                gemm<decltype(epi)::value>(in, ld_in, 1, se_in, Wb[l], K(l), 1, (int64_t)H * K(l), (int)cnt, H, K(l),
                                           nullptr, 0, 0, o, H, cnt * H, p + offb[l], P, nullptr, 0, 0, nullptr, 0,
                                           ypq, (int64_t)out_slots((int)cnt) * cnt, p + offwo, P);
            };
            // the last hidden layer stores no activations: its epilogue forms
            // the output layer's partial sums in ypq
            if (l == L - 1) run(std::integral_constant<int, 4>());
            else run(std::integral_constant<int, 3>());
            in = o;
            ld_in = H;
            se_in = cnt * H;
        }
    }
};

namespace {

int mfail(int code, const std::string& msg) {
    pp::set_last_error(msg);
    return code;
}

}  // namespace

#define MLP_BEGIN try {
#define MLP_END                                 \
    }                                           \
    catch (const MlpError& e) {                 \
        return mfail(e.code, e.msg);            \
    }                                           \
    catch (const std::exception& e) {           \
        return mfail(PP_EHIP, e.what());        \
    }                                           \
    pp::set_last_error("");                     \
    return PP_OK;

extern "C" {

int pp_mlp_create(int E, int din, int hidden, int layers, pp_mlp** out) {
    if (!out) return mfail(PP_EARG, "out is NULL");
    *out = nullptr;
    if (E < 1 || E > 64) return mfail(PP_EARG, "E must be in 1:64");
    if (din < 1 || din > kKin) return mfail(PP_EARG, "din must be in 1:64");
    if (hidden < 64 || hidden % 64) return mfail(PP_EARG, "hidden must be a positive multiple of 64");
    if (layers < 1 || layers > 16) return mfail(PP_EARG, "layers must be in 1:16");
    pp_mlp* h = new pp_mlp();
    try {
        h->init(E, din, hidden, layers);
    } catch (const MlpError& e) {
        delete h;
        return mfail(e.code, e.msg);
    }
    *out = h;
    return PP_OK;
}

int pp_mlp_destroy(pp_mlp* h) {
    if (!h) return PP_OK;
    (void)hipStreamSynchronize(h->stream);
    delete h;
    return PP_OK;
}

int pp_mlp_param_count(pp_mlp* h, int64_t* count) {
    if (!h || !count) return mfail(PP_EARG, "null argument");
    // host layout per member: W_1[H][din], b_1, (W_l[H][H], b_l)..., w_o[H], b_o
    int64_t c = (int64_t)h->H * h->din + h->H;
    for (int l = 1; l < h->L; ++l) c += (int64_t)h->H * h->H + h->H;
    *count = c + h->H + 1;
    return PP_OK;
}

static void host_layout(pp_mlp* h, const float* src, std::vector<float>& dst) {
    // host layout -> padded device layout of one member
    dst.assign((size_t)h->P, 0.0f);
    int64_t q = 0;
    for (int l = 0; l < h->L; ++l) {
        const int kin = l == 0 ? h->din : h->H;
        for (int r = 0; r < h->H; ++r)
            for (int c = 0; c < kin; ++c) dst[h->offW[l] + (int64_t)r * h->K(l) + c] = src[q++];
        for (int r = 0; r < h->H; ++r) dst[h->offb[l] + r] = src[q++];
    }
    for (int r = 0; r < h->H; ++r) dst[h->offwo + r] = src[q++];
    dst[h->offbo] = src[q++];
}

static void device_layout(pp_mlp* h, const float* src, float* dst) {
    int64_t q = 0;
    for (int l = 0; l < h->L; ++l) {
        const int kin = l == 0 ? h->din : h->H;
        for (int r = 0; r < h->H; ++r)
            for (int c = 0; c < kin; ++c) dst[q++] = src[h->offW[l] + (int64_t)r * h->K(l) + c];
        for (int r = 0; r < h->H; ++r) dst[q++] = src[h->offb[l] + r];
    }
    for (int r = 0; r < h->H; ++r) dst[q++] = src[h->offwo + r];
    dst[q++] = src[h->offbo];
}

int pp_mlp_set_params(pp_mlp* h, const float* params) {
    if (!h || !params) return mfail(PP_EARG, "null argument");
    MLP_BEGIN
    int64_t cnt = 0;
    pp_mlp_param_count(h, &cnt);
    std::vector<float> all((size_t)h->E * h->P), one;
    for (int e = 0; e < h->E; ++e) {
        host_layout(h, params + (size_t)e * cnt, one);
        std::copy(one.begin(), one.end(), all.begin() + (size_t)e * h->P);
    }
    MCHK(hipMemcpyAsync(h->p, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    MCHK(hipMemsetAsync(h->m, 0, sizeof(float) * h->E * h->P, h->stream));
    MCHK(hipMemsetAsync(h->v, 0, sizeof(float) * h->E * h->P, h->stream));
    h->step = 0;
    h->cast_weights();
    MCHK(hipStreamSynchronize(h->stream));
    MLP_END
}

int pp_mlp_get_params(pp_mlp* h, float* params) {
    if (!h || !params) return mfail(PP_EARG, "null argument");
    MLP_BEGIN
    int64_t cnt = 0;
    pp_mlp_param_count(h, &cnt);
    std::vector<float> all((size_t)h->E * h->P);
    MCHK(hipMemcpyAsync(all.data(), h->p, all.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    for (int e = 0; e < h->E; ++e) device_layout(h, all.data() + (size_t)e * h->P, params + (size_t)e * cnt);
    MLP_END
}

int pp_mlp_load_data(pp_mlp* h, const float* x, const float* y, int64_t n) {
    if (!h || !x || !y || n < 1) return mfail(PP_EARG, "bad arguments");
    MLP_BEGIN
    if (h->X) (void)hipFree(h->X);
    if (h->Y) (void)hipFree(h->Y);
    if (h->perm) (void)hipFree(h->perm);
    h->X = mal<uint16_t>((size_t)n * kKin);
    h->Y = mal<float>((size_t)n);
    h->perm = mal<int32_t>((size_t)h->E * n);
    h->n = n;
    float* xf = mal<float>((size_t)n * h->din);
    MCHK(hipMemcpyAsync(xf, x, sizeof(float) * n * h->din, hipMemcpyHostToDevice, h->stream));
    MCHK(hipMemcpyAsync(h->Y, y, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(kg_pack_input, dim3((unsigned)cdiv64(n * kKin, 256)), dim3(256), 0, h->stream, xf, n, h->din,
                       h->X);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(h->stream));
    (void)hipFree(xf);
    MLP_END
}

// One epoch: floor(n / batch) Adam steps; member e visits rows in the order
// perm[e*n .. e*n + n).  loss[e] = mean minibatch loss over the epoch.
int pp_mlp_train_epoch(pp_mlp* h, const int32_t* perm, int batch, float lr, float* loss_out) {
    if (!h || !perm || !loss_out) return mfail(PP_EARG, "null argument");
    if (batch < 64 || batch % 64) return mfail(PP_EARG, "batch must be a positive multiple of 64");
    if (!h->X) return mfail(PP_EARG, "no training data loaded");
    if (batch > h->n) return mfail(PP_EARG, "batch larger than the data set");
    MLP_BEGIN
    for (int64_t i = 0; i < (int64_t)h->E * h->n; ++i)
        if (perm[i] < 0 || perm[i] >= h->n) throw MlpError{PP_EARG, "perm entry out of range"};
    MCHK(hipMemcpyAsync(h->perm, perm, sizeof(int32_t) * h->E * h->n, hipMemcpyHostToDevice, h->stream));
    h->ensure_batch(batch);
    const int64_t steps = h->n / batch;
    std::vector<double> acc((size_t)h->E, 0.0);
    // the steps are enqueued back to back; the losses are summed on the device
    // and read once at the end of the epoch
    MCHK(hipMemsetAsync(h->loss_acc, 0, sizeof(double) * h->E, h->stream));
    for (int64_t s = 0; s < steps; ++s) {
        hipLaunchKernelGGL(kg_gather, dim3((unsigned)cdiv64(batch, 4), h->E), dim3(256), 0, h->stream, h->X, h->Y,
                           h->perm, h->n, s * batch, batch, h->Xb, h->Yb);
        h->forward_train(batch);
        h->backward(batch);
        h->adam(lr);
    }
    MCHK(hipMemcpyAsync(acc.data(), h->loss_acc, sizeof(double) * h->E, hipMemcpyDeviceToHost, h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    for (int e = 0; e < h->E; ++e) loss_out[e] = (float)(acc[e] / (double)std::max<int64_t>(steps, 1));
    MLP_END
}

// Gradients of the mean squared error of every member on one batch (x[b][din],
// t[b], shared by all members), in the host parameter layout: grads[e][count].
int pp_mlp_gradients(pp_mlp* h, const float* x, const float* t, int batch, float* grads, float* loss_out) {
    if (!h || !x || !t || !grads) return mfail(PP_EARG, "null argument");
    if (batch < 64 || batch % 64) return mfail(PP_EARG, "batch must be a positive multiple of 64");
    MLP_BEGIN
    h->ensure_batch(batch);
    h->ensure_query(batch);
    MCHK(hipMemcpyAsync(h->xq_f, x, sizeof(float) * batch * h->din, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(kg_pack_input, dim3((unsigned)cdiv64((int64_t)batch * kKin, 256)), dim3(256), 0, h->stream,
                       h->xq_f, (int64_t)batch, h->din, h->Xq);
    for (int e = 0; e < h->E; ++e) {
        MCHK(hipMemcpyAsync(h->Xb + (size_t)e * batch * kKin, h->Xq, sizeof(uint16_t) * batch * kKin,
                            hipMemcpyDeviceToDevice, h->stream));
        MCHK(hipMemcpyAsync(h->Yb + (size_t)e * batch, t, sizeof(float) * batch, hipMemcpyHostToDevice, h->stream));
    }
    h->forward_train(batch);
    h->backward(batch);
    std::vector<float> all((size_t)h->E * h->P);
    MCHK(hipMemcpyAsync(all.data(), h->g, all.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    std::vector<float> l((size_t)h->E);
    MCHK(hipMemcpyAsync(l.data(), h->loss, sizeof(float) * h->E, hipMemcpyDeviceToHost, h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    int64_t cnt = 0;
    pp_mlp_param_count(h, &cnt);
    for (int e = 0; e < h->E; ++e) device_layout(h, all.data() + (size_t)e * h->P, grads + (size_t)e * cnt);
    if (loss_out)
        for (int e = 0; e < h->E; ++e) loss_out[e] = l[e];
    MLP_END
}

// Ensemble prediction for n query rows x[n][din]: per-member outputs
// (members[e*n + i], may be NULL), mean and population std over members.
// seconds (may be NULL): device time of the forward passes (HIP events).
int pp_mlp_predict(pp_mlp* h, const float* x, int64_t n, float* members, double* mean, double* sd, double* seconds) {
    if (!h || !x || !mean || !sd || n < 0) return mfail(PP_EARG, "bad arguments");
    if (n == 0) return PP_OK;
    MLP_BEGIN
    const int64_t chunk = std::min<int64_t>(65536, (n + 63) / 64 * 64);
    h->ensure_query(chunk);
    hipEvent_t e0, e1;
    MCHK(hipEventCreate(&e0));
    MCHK(hipEventCreate(&e1));
    double t = 0.0;
    std::vector<float> yh;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int64_t cnt = std::min(chunk, n - r0);
        const int64_t cpad = (cnt + 63) / 64 * 64;
        MCHK(hipMemsetAsync(h->xq_f, 0, sizeof(float) * cpad * h->din, h->stream));
        MCHK(hipMemcpyAsync(h->xq_f, x + r0 * h->din, sizeof(float) * cnt * h->din, hipMemcpyHostToDevice, h->stream));
        MCHK(hipEventRecord(e0, h->stream));
        hipLaunchKernelGGL(kg_pack_input, dim3((unsigned)cdiv64(cpad * kKin, 256)), dim3(256), 0, h->stream, h->xq_f,
                           cpad, h->din, h->Xq);
        h->forward_query(cpad);
        const int ns = h->out_slots((int)cpad);
        hipLaunchKernelGGL(kg_ensemble, dim3((unsigned)cdiv64(cpad, 256)), dim3(256), 0, h->stream, h->ypq, ns,
                           (int64_t)ns * cpad, h->p + h->offbo, h->P, h->E, cpad, cpad, h->yq, h->mq, h->sq);
        MCHK(hipGetLastError());
        MCHK(hipEventRecord(e1, h->stream));
        MCHK(hipMemcpyAsync(mean + r0, h->mq, sizeof(double) * cnt, hipMemcpyDeviceToHost, h->stream));
        MCHK(hipMemcpyAsync(sd + r0, h->sq, sizeof(double) * cnt, hipMemcpyDeviceToHost, h->stream));
        if (members) {
            yh.resize((size_t)h->E * cpad);
            MCHK(hipMemcpyAsync(yh.data(), h->yq, sizeof(float) * h->E * cpad, hipMemcpyDeviceToHost, h->stream));
        }
        MCHK(hipStreamSynchronize(h->stream));
        float ms = 0.f;
        MCHK(hipEventElapsedTime(&ms, e0, e1));
        t += ms * 1e-3;
        if (members)
            for (int e = 0; e < h->E; ++e)
                std::memcpy(members + (size_t)e * n + r0, yh.data() + (size_t)e * cpad, sizeof(float) * cnt);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = t;
    MLP_END
}

}  // extern "C"
