/*
 * lzr_net.hip -- the learned LM initializer on gfx950: a small CNN whose forward and
 * backward passes run on the BF16 matrix cores (MFMA, FP32 accumulation), trained with
 * Adam on FP32 master weights, and the C ABI of include/localizer_hip.h (lzr_net_*).
 *
 * Network (11 x 11 patch, half-width 5):
 *   x  = bf16((patch - median) / scale)                       121
 *   a1 = bf16(relu(conv3x3(x; W1) + b1))        9 x 9 x 32    (valid padding)
 *   a2 = bf16(relu(conv3x3(a1; W2) + b2))       7 x 7 x 32
 *   h  = bf16(relu(W_fc1 flatten(a2) + b_fc1))  64            flatten index (y*7+x)*32+c
 *   t  = W_fc2 h + b_fc2                        5             FP32
 *
 * MFMA tiling.  v_mfma_f32_16x16x32_bf16: lane l holds A[m = l%16][k-slot (l/16, i)],
 * B[k-slot (l/16, i)][n = l%16] (i < 8) and C[m = 4(l/16) + i][n = l%16] (i < 4).  Any
 * bijection of k-slots to K indices works as long as A and B use the same one.  The
 * forward pass computes every layer transposed (features x patches): weights are the A
 * operand, activations the B operand.  One K-chunk of 32 is one 3x3 tap over the 32
 * input channels (conv2) or one 7x7 position of the flattened map (fc1).  conv1 (K = 9
 * taps) uses v_mfma_f32_16x16x16bf16_1k with the taps in k-slots 0..8.
 *
 * Forward kernel: persistent, one 8-wave block per CU, 16 patches per group (the MFMA
 * N dimension).  conv1 writes a1 to LDS (8-byte stores of 4 channels); conv2 at
 * position q reads 8 channels of one patch per tap (16-byte LDS reads), and the 8
 * outputs a lane holds are its fc1 B operand for K-chunk q, so the 7x7x32 map never
 * leaves registers (inference) and fc1 accumulates while conv2 runs.  Weights packed
 * in MFMA operand order (conv2: 72 VGPRs per lane) are loaded once per wave; fc1's are
 * prefetched one position ahead.
 *
 * Backward (training): activations are stored transposed ([feature][patch], patch
 * fastest) so that every weight-gradient GEMM reads contiguous 16-byte K-slices
 * (K = patches); gradients are accumulated in FP32 with atomics.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/localizer_hip.h"
#include "lzr_order.h"
#include "lzr_netinit.h"

#define NET_CHECK(call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) throw std::string(#call ": ") + hipGetErrorString(e_);        \
    } while (0)

namespace lzr_net {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

constexpr int kH = LZR_NET_HALF_WIDTH; /* 5 */
constexpr int kN = 2 * kH + 1;         /* 11 */
constexpr int kPix = kN * kN;          /* 121 */
constexpr int kXRow = 128;             /* bf16 per patch row of the input buffer */
constexpr int kC = 32;                 /* channels */
constexpr int kS1 = 9, kP1 = 81;       /* conv1 output 9 x 9 */
constexpr int kS2 = 7, kP2 = 49;       /* conv2 output 7 x 7 */
constexpr int kF = kP2 * kC;           /* 1568 */
constexpr int kHid = 64;
constexpr int kOut = LZR_NET_OUTPUTS;  /* 5 */
constexpr int kGroup = 16;             /* patches per forward group (one MFMA M-tile) */
constexpr int kWaves = 8; /* 2 per SIMD */
constexpr int kXsStride = kXRow + 8; /* LDS row of a patch: 272 B, 16 patches hit 16 banks */
constexpr int kThreads = 64 * kWaves;
constexpr int kA1Stride = kP1 * kC + 8; /* bf16 per patch in LDS (padded) */
constexpr int kHsStride = kHid + 8;

constexpr int OFF_W1 = LZR_NET_OFF_CONV1_W, OFF_B1 = LZR_NET_OFF_CONV1_B;
constexpr int OFF_W2 = LZR_NET_OFF_CONV2_W, OFF_B2 = LZR_NET_OFF_CONV2_B;
constexpr int OFF_WF1 = LZR_NET_OFF_FC1_W, OFF_BF1 = LZR_NET_OFF_FC1_B;
constexpr int OFF_WF2 = LZR_NET_OFF_FC2_W, OFF_BF2 = LZR_NET_OFF_FC2_B;
constexpr int NPARAMS = LZR_NET_NPARAMS;
static_assert(OFF_B1 == 32 * 9 && OFF_W2 == OFF_B1 + 32 && OFF_B2 == OFF_W2 + 32 * 32 * 9, "layout");
static_assert(OFF_WF1 == OFF_B2 + 32 && OFF_BF1 == OFF_WF1 + 64 * kF && OFF_WF2 == OFF_BF1 + 64, "layout");
static_assert(OFF_BF2 == OFF_WF2 + 5 * 64 && NPARAMS == OFF_BF2 + 5, "layout");

/* Packed BF16 operand buffers (element counts). */
constexpr int PK_W1 = 2 * 64 * 4;         /* [tile 2][lane][4]             conv1^T A */
constexpr int PK_W2 = 9 * 2 * 64 * 8;     /* [tap][tile 2][lane][8]        conv2^T A */
constexpr int PK_W2T = 9 * 2 * 64 * 8;    /* [tap][ntile 2][lane][8]       conv2^T B */
constexpr int PK_WF1 = 49 * 4 * 64 * 8;   /* [q][mtile 4][lane][8]         fc1^T A   */
constexpr int PK_WF1T = 98 * 2 * 64 * 8;  /* [ntile 98][kchunk 2][lane][8] fc1^T B   */
constexpr int PK_WF2 = 2 * 64 * 8;        /* [kchunk 2][lane][8]           fc2^T A   */
constexpr int PK_TOTAL = PK_W1 + PK_W2 + PK_W2T + PK_WF1 + PK_WF1T + PK_WF2;

/* ---- BF16 helpers ---------------------------------------------------------------- */

/* FP32 -> BF16, round to nearest even (NaN kept quiet). */
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
/* bf16 strictly positive (the ReLU mask of a stored activation) */
__device__ inline bool bf_pos(uint16_t b) { return b != 0 && !(b & 0x8000u); }

__device__ inline f32x4_t mfma32(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}
__device__ inline f32x4_t mfma16(s16x4_t a, s16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ inline u32x4_t ld16(const uint16_t *p) { return *reinterpret_cast<const u32x4_t *>(p); }
__device__ inline void st8(uint16_t *p, uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
    u32x2_t v;
    v[0] = (uint32_t)a | ((uint32_t)b << 16);
    v[1] = (uint32_t)c | ((uint32_t)d << 16);
    *reinterpret_cast<u32x2_t *>(p) = v;
}
__device__ inline void ld8(const uint16_t *p, uint16_t (&o)[4]) {
    const u32x2_t v = *reinterpret_cast<const u32x2_t *>(p);
    o[0] = (uint16_t)(v[0] & 0xffffu);
    o[1] = (uint16_t)(v[0] >> 16);
    o[2] = (uint16_t)(v[1] & 0xffffu);
    o[3] = (uint16_t)(v[1] >> 16);
}
/* orders this wave's LDS accesses (LDS instructions of one wave complete in order) */
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Per-patch statistics written by the prep kernel. */
struct PatchStat {
    float median, scale;
    int xoff, yoff;
    int valid;
};

struct Packed {
    const uint16_t *w1, *w2, *w2t, *wf1, *wf1t, *wf2;
    const float *wf2f; /* bf16-rounded fc2 weights as FP32 [5][64] */
    const float *params;
};

/* Activation stores of the training forward (patch-fastest, leading dimension ldb). */
struct TrainStore {
    uint16_t *xT;  /* [128][ldb]      */
    uint16_t *a1T; /* [81*32][ldb]    */
    uint16_t *a2T; /* [1568][ldb]     */
    uint16_t *h;   /* [ldb][64]       */
    int ldb;
};

/* ---- prep: patch gather, median / scale, normalized BF16 input ------------------- */

/* Two candidates per wave, one per half-wave of 32 lanes, four pixels per lane (index
 * l + 32 q, q < 4, for l + 32 q < 121).  The median (sorted key 60) is selected bit by bit
 * from the top with wave ballots, both halves at once: the largest t < 2^16 with
 * #{keys < t} <= 60 is the key of rank 60, exactly -- about a dozen uniform steps of four
 * compares and four ballots instead of the 28 exchange stages of a 128-key sorting network. */
__device__ inline uint32_t half_max16(uint32_t v) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}

/* Rows 8 rb .. 8 rb + 7 (two per wave of the block). */
__device__ __forceinline__ void prep_rows(const uint16_t *frames, int nx, int ny, int nframes,
                                          const LzrCandidate *cands, int ncands, int nrows, uint16_t *xin,
                                          PatchStat *stats, double k_start, double *start_sse, const int *sel_idx,
                                          int rb) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, hl = lane & 31;
    const int r = (rb * 4 + wave) * 2 + half;
    const bool live = r < nrows;
    if (__ballot(live) == 0) return; /* wave-uniform */
    bool valid = live && r < ncands;
    const int ci = sel_idx && valid ? sel_idx[r] : r; /* candidate of row r */
    int xoff = 0, yoff = 0, f = 0;
    double cx = 0.0, cy = 0.0, ca = 0.0, cb = 0.0;
    if (valid) {
        const LzrCandidate c = cands[ci];
        cx = c.x;
        cy = c.y;
        ca = c.intensity;
        cb = c.background;
        const double fx = c.x - (double)kH, fy = c.y - (double)kH;
        valid = c.frame >= 0 && c.frame < nframes && fx >= 0.0 && fy >= 0.0 && fx < 2147483647.0 &&
                fy < 2147483647.0;
        if (valid) {
            xoff = (int)fx;
            yoff = (int)fy;
            f = c.frame;
            valid = xoff + kN - 1 <= nx - 1 && yoff + kN - 1 <= ny - 1;
        }
    }
    uint32_t u[4];
    bool has[4];
    const uint16_t *src = frames + ((size_t)f * ny + yoff) * nx + xoff;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = hl + 32 * q;
        has[q] = valid && i < kPix;
        u[q] = has[q] ? (uint32_t)src[(size_t)(i / kN) * nx + (i % kN)] : 0u;
    }
    const uint32_t top = half_max16(max(max(u[0], u[1]), max(u[2], u[3])));
    /* rank 60 in each half (uniform loop over the larger top bit of the two halves) */
    const uint32_t top2 = __builtin_amdgcn_readfirstlane(half_max16(top) | (uint32_t)__shfl_xor((int)top, 32, 64));
    const int msb = top2 ? 31 - __builtin_clz(top2) : 0;
    uint32_t lo = 0;
    for (int bit = msb; bit >= 0; --bit) {
        const uint32_t t = lo | (1u << bit);
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const unsigned long long m = __ballot(has[q] && u[q] < t);
            cnt += half ? __popcll(m >> 32) : __popcll(m & 0xffffffffull);
        }
        if (cnt <= kPix / 2) lo = t;
    }
    if (start_sse) { /* kernel argument: uniform */
        /* |f|^2 over the patch at the reference start point (A = intensity, b = background,
         * width w0 = sqrt(0.5 / k_start) at the candidate's position): lanes 0..10 of the half
         * hold the column factors, lanes 11..21 the row factors, every pixel one product */
        const double x0 = cx - (double)xoff, y0 = cy - (double)yoff;
        const double t = hl < kN ? (double)hl - x0 : (double)(hl - kN) - y0;
        const double e = exp(-(t * t) * k_start);
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = hl + 32 * q, row_i = i / kN, col_i = i - row_i * kN;
            const double ex = __shfl(e, half * 32 + col_i, 64);
            const double ey = __shfl(e, half * 32 + kN + row_i, 64); /* row_i <= 11 */
            const double fv = __builtin_fma(ca, ex * ey, cb - (double)u[q]);
            acc = has[q] ? __builtin_fma(fv, fv, acc) : acc;
        }
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (live && hl == 0) start_sse[ci] = valid ? acc : __builtin_nan("");
    }
    uint16_t *row = xin + (size_t)(live ? r : 0) * kXRow;
    if (!live) return;
    if (!valid) {
#pragma unroll
        for (int q = 0; q < 4; ++q) row[hl + 32 * q] = 0;
        if (hl == 0) {
            PatchStat st;
            st.median = 0.f;
            st.scale = 1.f;
            st.xoff = xoff;
            st.yoff = yoff;
            st.valid = 0;
            stats[r] = st;
        }
        return;
    }
    const float m = (float)lo, mx = (float)top; /* keys 60 and 120 */
    float scale = mx - m;
    if (!(scale >= 1.f)) scale = 1.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = hl + 32 * q;
        row[i] = i < kPix ? f2bf(((float)u[q] - m) / scale) : (uint16_t)0;
    }
    if (hl == 0) {
        PatchStat st;
        st.median = m;
        st.scale = scale;
        st.xoff = xoff;
        st.yoff = yoff;
        st.valid = 1;
        stats[r] = st;
    }
}

/* One block per 8 rows; with a device row count (sel_count: rows [0, *sel_count) are the
 * candidates sel_idx names) the launch is sized for the most rows and its blocks stride
 * over the rows the count leaves. */
__global__ void __launch_bounds__(256) prep_kernel(const uint16_t *frames, int nx, int ny, int nframes,
                                                   const LzrCandidate *cands, int ncands, int nrows,
                                                   uint16_t *xin, PatchStat *stats, double k_start,
                                                   double *start_sse, const int *sel_idx,
                                                   const unsigned *sel_count) {
    if (!sel_count) {
        prep_rows(frames, nx, ny, nframes, cands, ncands, nrows, xin, stats, k_start, start_sse, nullptr, blockIdx.x);
        return;
    }
    const int rows = (int)*sel_count;
    for (int rb = blockIdx.x; rb * 8 < rows; rb += gridDim.x)
        prep_rows(frames, nx, ny, nframes, cands, rows, rows, xin, stats, k_start, start_sse, sel_idx, rb);
}

/* ---- fused forward ---------------------------------------------------------------- */

/* Two FP32 -> packed BF16 (round to nearest even; v_cvt_pk_bf16_f32). */
__device__ inline uint32_t cvt_pk(float a, float b) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    const f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

/* ReLU as one integer max on the bit pattern (a float's bits order like the float for
 * non-negative values; negative values, -0 included, have the sign bit set and map to +0). */
__device__ inline float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

/* Every layer is computed transposed, C[feature][patch] = W * X^T: the weights are the
 * A operand (registers / L2) and the activations the B operand (LDS), so a lane's four
 * outputs are four consecutive features of one patch (one 8-byte LDS store), and the
 * eight conv2 outputs a lane holds at position q are directly its fc1 B operand for
 * K-chunk q, with the k-slot (g, i) of that chunk mapped to channel
 * o(g, i) = (i < 4 ? 4g + i : 16 + 4g + i - 4) (the fc1 weights are packed the same way). */
template <bool kTrain>
__global__ void __launch_bounds__(kThreads) forward_kernel(const uint16_t *xin, int n, Packed W, float *out,
                                                           TrainStore T, const unsigned *n_dev) {
    if (n_dev) n = (int)*n_dev; /* rows counted on the device (at most the launch's n) */
    __shared__ __attribute__((aligned(16))) uint16_t xs[kGroup * kXsStride];
    __shared__ __attribute__((aligned(16))) uint16_t a1s[kGroup * kA1Stride];
    __shared__ __attribute__((aligned(16))) uint16_t hs[kGroup * kHsStride];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6); /* uniform: position indices in scalar registers */
    const int lr = lane & 15, lg = lane >> 4;

    /* weights in MFMA A-operand order, once per wave */
    s16x4_t w1r[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const uint16_t *p = W.w1 + (mt * 64 + lane) * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) w1r[mt][i] = (short)p[i];
    }
    u32x4_t w2r[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) w2r[t][mt] = ld16(W.w2 + ((t * 2 + mt) * 64 + lane) * 8);
    u32x4_t wf2r[2];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) wf2r[kc] = ld16(W.wf2 + (kc * 64 + lane) * 8);
    /* biases of this lane's output rows (features mt*16 + 4*lg + i) */
    f32x4_t b1v[2], b2v[2], bf2v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            b1v[mt][i] = W.params[OFF_B1 + mt * 16 + 4 * lg + i];
            b2v[mt][i] = W.params[OFF_B2 + mt * 16 + 4 * lg + i];
        }
        bf2v[i] = 4 * lg + i < kOut ? W.params[OFF_BF2 + 4 * lg + i] : 0.f;
    }
    /* conv1 taps of this lane's k-slots 4*lg + i */
    int toff[4];
    bool kval[4]; /* k-slots beyond the 9 taps read the patch's first pixel and are zeroed */
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * lg + i;
        kval[i] = k < 9;
        toff[i] = k < 9 ? (k / 3) * kN + k % 3 : 0;
    }

    const int ngroups = (n + kGroup - 1) / kGroup;
    /* thread tid < 256 moves 16 bytes of the group's input; the next group's are loaded
     * while conv2 runs */
    const int xr_ = tid >> 4, xc_ = tid & 15;
    u32x4_t xnext = {0u, 0u, 0u, 0u};
    if (tid < kGroup * 16 && (int)blockIdx.x * kGroup + xr_ < n)
        xnext = ld16(xin + (size_t)(blockIdx.x * kGroup + xr_) * kXRow + xc_ * 8);
    for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int roi0 = g * kGroup;
        /* phase 0: the group's normalized patches */
        if (tid < kGroup * 16) {
            const int r = xr_, c = xc_;
            const u32x4_t v = xnext;
            *reinterpret_cast<u32x4_t *>(xs + r * kXsStride + c * 8) = v;
            if (kTrain) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    T.xT[(size_t)(c * 8 + 2 * k) * T.ldb + roi0 + r] = (uint16_t)(v[k] & 0xffffu);
                    T.xT[(size_t)(c * 8 + 2 * k + 1) * T.ldb + roi0 + r] = (uint16_t)(v[k] >> 16);
                }
            }
        }
        __syncthreads();

        /* phase 1: conv1 at position p: a1^T[32 ch][16 patches]; two positions per step so
         * that the LDS reads of the second overlap the first's MFMAs */
        for (int p0 = wave; p0 < kP1; p0 += 2 * kWaves) {
            s16x4_t bx[2];
            int pp[2];
            pp[0] = p0;
            pp[1] = p0 + kWaves < kP1 ? p0 + kWaves : -1; /* uniform across the wave */
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int p = pp[u] < 0 ? p0 : pp[u];
                const int py = p / kS1, px = p - py * kS1;
                const uint16_t *xr = xs + lr * kXsStride + py * kN + px;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const short v = (short)xr[toff[i]]; /* unconditional: no branch per load */
                    bx[u][i] = kval[i] ? v : (short)0;
                }
            }
            /* both positions' products first, so that neither waits on the other's results */
            f32x4_t cc[2][2];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) cc[0][mt] = mfma16(w1r[mt], bx[0], b1v[mt]);
            if (pp[1] >= 0) {
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) cc[1][mt] = mfma16(w1r[mt], bx[1], b1v[mt]);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (pp[u] < 0) continue;
                const int p = pp[u];
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    const f32x4_t c = cc[u][mt];
                    const float v0 = relu(c[0]), v1 = relu(c[1]), v2 = relu(c[2]), v3 = relu(c[3]);
                    u32x2_t pk;
                    pk[0] = cvt_pk(v0, v1);
                    pk[1] = cvt_pk(v2, v3);
                    *reinterpret_cast<u32x2_t *>(a1s + lr * kA1Stride + p * kC + mt * 16 + 4 * lg) = pk;
                    if (kTrain) {
                        const uint16_t o[4] = {(uint16_t)(pk[0] & 0xffffu), (uint16_t)(pk[0] >> 16),
                                               (uint16_t)(pk[1] & 0xffffu), (uint16_t)(pk[1] >> 16)};
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            T.a1T[(size_t)(p * kC + mt * 16 + 4 * lg + i) * T.ldb + roi0 + lr] = o[i];
                    }
                }
            }
        }
        __syncthreads();

        /* phase 2: conv2 at position q (registers) -> fc1 K-chunk q */
        f32x4_t hacc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) hacc[mt] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
        if (tid < kGroup * 16) {
            const int gn = g + (int)gridDim.x;
            xnext = (u32x4_t){0u, 0u, 0u, 0u};
            if (gn < ngroups && gn * kGroup + xr_ < n) xnext = ld16(xin + (size_t)(gn * kGroup + xr_) * kXRow + xc_ * 8);
        }
        u32x4_t wf[4]; /* fc1 weights of this wave's next position, one position ahead */
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) wf[mt] = ld16(W.wf1 + ((wave * 4 + mt) * 64 + lane) * 8);
        for (int q = wave; q < kP2; q += kWaves) {
            const int qy = q / kS2, qx = q - qy * kS2;
            f32x4_t c0 = b2v[0], c1 = b2v[1];
            const uint16_t *ar = a1s + lr * kA1Stride + lg * 8;
            /* the nine B operands read before the products: one LDS wait instead of one per tap */
            u32x4_t b[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) b[t] = ld16(ar + ((qy + t / 3) * kS1 + qx + t % 3) * kC);
            if constexpr (!kTrain) __builtin_amdgcn_sched_barrier(0); /* keep the reads ahead of the products */
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                c0 = mfma32(w2r[t][0], b[t], c0);
                c1 = mfma32(w2r[t][1], b[t], c1);
            }
            u32x4_t bq;
            bq[0] = cvt_pk(relu(c0[0]), relu(c0[1]));
            bq[1] = cvt_pk(relu(c0[2]), relu(c0[3]));
            bq[2] = cvt_pk(relu(c1[0]), relu(c1[1]));
            bq[3] = cvt_pk(relu(c1[2]), relu(c1[3]));
            if (kTrain) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int o = i < 4 ? 4 * lg + i : 16 + 4 * lg + (i - 4);
                    const uint16_t v = (uint16_t)((bq[i >> 1] >> (16 * (i & 1))) & 0xffffu);
                    T.a2T[(size_t)(q * kC + o) * T.ldb + roi0 + lr] = v;
                }
            }
            /* fc1 K-chunk q, then the weights of this wave's next position into the same registers */
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) hacc[mt] = mfma32(wf[mt], bq, hacc[mt]);
            if (q + kWaves < kP2) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) wf[mt] = ld16(W.wf1 + (((q + kWaves) * 4 + mt) * 64 + lane) * 8);
            }
        }
        __syncthreads();
        /* [waves][64 j][kRedStride]: a1 is no longer needed.  A row of 16 patches is padded to 20
         * floats so that the four 16-lane quarters of a wave's stores fall on different banks. */
        constexpr int kRedStride = kGroup + 4;
        static_assert(kWaves * kHid * kRedStride * 4 <= kGroup * kA1Stride * 2, "reduction buffer");
        float *red = reinterpret_cast<float *>(a1s);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[(wave * kHid + mt * 16 + 4 * lg + i) * kRedStride + lr] = hacc[mt][i];
        __syncthreads();

        /* phase 3: h = relu(sum over waves + b) */
        for (int e = tid; e < kGroup * kHid; e += kThreads) {
            const int r = e & (kGroup - 1), j = e >> 4;
            float s = W.params[OFF_BF1 + j];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) s += red[(w * kHid + j) * kRedStride + r];
            const uint16_t hb = f2bf(relu(s));
            hs[r * kHsStride + j] = hb;
            if (kTrain) T.h[(size_t)(roi0 + r) * kHid + j] = hb;
        }
        __syncthreads();

        /* phase 4: fc2, out^T[5][16 patches] */
        if (wave == 0) {
            f32x4_t o = bf2v;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) o = mfma32(wf2r[kc], ld16(hs + lr * kHsStride + kc * 32 + lg * 8), o);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = 4 * lg + i;
                if (t < kOut && roi0 + lr < n) out[(size_t)(roi0 + lr) * kOut + t] = o[i];
            }
        }
        __syncthreads();
    }
}

/* ---- outputs -> start points / targets -------------------------------------------- */

__global__ void decode_kernel(const float *out, const PatchStat *stats, int n, double w, LzrStart *starts,
                              float *raw, const int *sel_idx = nullptr, const unsigned *sel_count = nullptr) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (sel_count) n = (int)*sel_count;
    if (r >= n) return;
    const PatchStat s = stats[r];
    float t[kOut];
    bool ok = s.valid != 0;
#pragma unroll
    for (int k = 0; k < kOut; ++k) {
        t[k] = out[(size_t)r * kOut + k];
        ok = ok && (t[k] == t[k]) && fabsf(t[k]) < 1e6f;
    }
    if (raw) {
#pragma unroll
        for (int k = 0; k < kOut; ++k) raw[(size_t)r * kOut + k] = s.valid ? t[k] : __builtin_nanf("");
    }
    if (!starts) return;
    LzrStart st;
    st.amplitude = 0.0;
    st.width = 0.0;
    st.x = 0.0;
    st.y = 0.0;
    st.background = 0.0;
    if (ok) {
        const double width = w * exp((double)t[1]);
        if (width > 0.0 && width < 1e3) {
            st.amplitude = (double)t[0] * (double)s.scale;
            st.width = width;
            st.x = (double)s.xoff + kH + (double)t[2];
            st.y = (double)s.yoff + kH + (double)t[3];
            st.background = (double)s.median + (double)t[4] * (double)s.scale;
        }
    }
    starts[sel_idx ? sel_idx[r] : r] = st;
}

/* Normalized targets from the LM labels; weight 1 for accepted symmetric fits of a
 * valid patch, else 0.  counter[0] += used rows. */
__global__ void targets_kernel(const LzrFit *labels, const PatchStat *stats, int ncands, int nrows, double w,
                               float *tgt, float *wt, unsigned *counter) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    bool use = false;
    float t[kOut] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (r < ncands) {
        const PatchStat s = stats[r];
        const LzrFit &L = labels[r];
        if (s.valid && L.status == LZR_ACCEPTED && L.model == LZR_MODEL_SYMMETRIC && L.params[1] > 0.0) {
            t[0] = (float)(L.params[0] / s.scale);
            t[1] = (float)log(L.params[1] / w);
            t[2] = (float)(L.params[2] - (s.xoff + kH));
            t[3] = (float)(L.params[3] - (s.yoff + kH));
            t[4] = (float)((L.params[4] - s.median) / s.scale);
            use = true;
#pragma unroll
            for (int k = 0; k < kOut; ++k) use = use && t[k] == t[k] && fabsf(t[k]) < 1e6f;
        }
    }
#pragma unroll
    for (int k = 0; k < kOut; ++k) tgt[(size_t)r * kOut + k] = use ? t[k] : 0.f;
    wt[r] = use ? 1.f : 0.f;
    const unsigned long long b = __ballot(use);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(counter, (unsigned)__popcll(b));
}

/* ---- backward ------------------------------------------------------------------------ */

/* Loss, fc2 backward, dh.  256 patches per block. */
__global__ void __launch_bounds__(256) head_backward_kernel(const float *out, const float *tgt, const float *wt,
                                                            const unsigned *counter, const uint16_t *hg,
                                                            const float *wf2f, int B, float *grad, float *loss,
                                                            uint16_t *dh, uint16_t *dhT) {
    __shared__ float sd[256][kOut];
    __shared__ uint16_t sh[256][kHid];
    __shared__ float sdh[256][kHid + 1];
    __shared__ float sl[256];
    const int tid = threadIdx.x;
    const int r = blockIdx.x * 256 + tid;
    const unsigned nv = *counter;
    const float inv = nv > 0 ? 1.f / (float)nv : 0.f;
    float d[kOut];
    float l = 0.f;
#pragma unroll
    for (int k = 0; k < kOut; ++k) {
        float e = 0.f;
        if (r < B && wt[r] > 0.f) e = out[(size_t)r * kOut + k] - tgt[(size_t)r * kOut + k];
        d[k] = 2.f * e * inv;
        l += e * e * inv;
        sd[tid][k] = d[k];
    }
    sl[tid] = l;
    for (int j = 0; j < kHid; ++j) {
        const uint16_t hb = r < B ? hg[(size_t)r * kHid + j] : (uint16_t)0;
        sh[tid][j] = hb;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < kOut; ++k) s += wf2f[k * kHid + j] * d[k];
        const uint16_t db = f2bf(bf_pos(hb) ? s : 0.f);
        sdh[tid][j] = bf2f(db);
        if (r < B) {
            dh[(size_t)r * kHid + j] = db;
            dhT[(size_t)j * B + r] = db;
        }
    }
    __syncthreads();
    for (int e = tid; e < kOut * kHid; e += 256) {
        const int k = e / kHid, j = e % kHid;
        float s = 0.f;
        for (int q = 0; q < 256; ++q) s += sd[q][k] * bf2f(sh[q][j]);
        atomicAdd(&grad[OFF_WF2 + k * kHid + j], s);
    }
    if (tid < kHid) {
        float s = 0.f;
        for (int q = 0; q < 256; ++q) s += sdh[q][tid];
        atomicAdd(&grad[OFF_BF1 + tid], s);
    }
    if (tid < kOut) {
        float s = 0.f;
        for (int q = 0; q < 256; ++q) s += sd[q][tid];
        atomicAdd(&grad[OFF_BF2 + tid], s);
    }
    if (tid == 0) {
        float s = 0.f;
        for (int q = 0; q < 256; ++q) s += sl[q];
        atomicAdd(loss, s);
    }
}

/* d a2 = (dh W_fc1) * [a2 > 0].  One wave per (16-patch M-tile, 14 N-tiles). */
__global__ void __launch_bounds__(256) fc1_dx_kernel(const uint16_t *dh, const uint16_t *wf1t,
                                                     const uint16_t *a2T, int B, uint16_t *da2, uint16_t *da2T) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
    const int mt = blockIdx.x * 4 + wave;
    if (mt * 16 >= B) return;
    const int m0 = mt * 16;
    const u32x4_t a0 = ld16(dh + (size_t)(m0 + lr) * kHid + lg * 8);
    const u32x4_t a1 = ld16(dh + (size_t)(m0 + lr) * kHid + 32 + lg * 8);
    for (int nt = blockIdx.y * 14; nt < blockIdx.y * 14 + 14; ++nt) {
        f32x4_t c = {0.f, 0.f, 0.f, 0.f};
        c = mfma32(a0, ld16(wf1t + ((nt * 2 + 0) * 64 + lane) * 8), c);
        c = mfma32(a1, ld16(wf1t + ((nt * 2 + 1) * 64 + lane) * 8), c);
        const int kf = nt * 16 + lr;
        uint16_t msk[4], o[4];
        ld8(a2T + (size_t)kf * B + m0 + 4 * lg, msk);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o[i] = f2bf(bf_pos(msk[i]) ? c[i] : 0.f);
            da2[(size_t)(m0 + 4 * lg + i) * kF + kf] = o[i];
        }
        st8(da2T + (size_t)kf * B + m0 + 4 * lg, o[0], o[1], o[2], o[3]);
    }
}

/* Weight gradient of a GEMM whose K runs over patches: G[m][n] += sum_k A(m,k) B(k,n),
 * A rows and B columns stored patch-fastest.  One wave per (M-tile, N-tile, K-slice). */
__device__ inline void kgemm_tile(const uint16_t *arow, const uint16_t *brow, int k0, int k1, f32x4_t &c) {
    for (int k = k0; k < k1; k += 32) c = mfma32(ld16(arow + k), ld16(brow + k), c);
}

/* dW_fc1[j][kf] = sum_r dh[r][j] a2[r][kf].  grid (98, ksplit), block 256 = 4 M-tiles. */
__global__ void __launch_bounds__(256) fc1_dw_kernel(const uint16_t *dhT, const uint16_t *a2T, int B, int kslice,
                                                     float *grad) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
    const int nt = blockIdx.x, mt = wave;
    const int k0 = blockIdx.y * kslice, k1 = min(B, k0 + kslice);
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
    kgemm_tile(dhT + (size_t)(mt * 16 + lr) * B + lg * 8, a2T + (size_t)(nt * 16 + lr) * B + lg * 8, k0, k1, c);
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&grad[OFF_WF1 + (mt * 16 + 4 * lg + i) * kF + nt * 16 + lr], c[i]);
}

/* d a1[r][p][c] = sum_t sum_o da2[r][p - off(t)][o] W2[o][c][t], masked by a1 > 0.
 * One wave per (16-patch M-tile, position p). */
__global__ void __launch_bounds__(256) conv2_dx_kernel(const uint16_t *da2, const uint16_t *w2t,
                                                       const uint16_t *a1T, int B, uint16_t *da1T) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
    const int p = blockIdx.y;
    const int mt = blockIdx.x * 4 + wave;
    if (mt * 16 >= B) return;
    const int m0 = mt * 16;
    const int py = p / kS1, px = p - py * kS1;
    f32x4_t c[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int qy = py - t / 3, qx = px - t % 3;
        if (qy < 0 || qy >= kS2 || qx < 0 || qx >= kS2) continue; /* uniform across the wave */
        const int q = qy * kS2 + qx;
        const u32x4_t a = ld16(da2 + (size_t)(m0 + lr) * kF + q * kC + lg * 8);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) c[nt] = mfma32(a, ld16(w2t + ((t * 2 + nt) * 64 + lane) * 8), c[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int ch = nt * 16 + lr;
        uint16_t msk[4], o[4];
        ld8(a1T + (size_t)(p * kC + ch) * B + m0 + 4 * lg, msk);
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = f2bf(bf_pos(msk[i]) ? c[nt][i] : 0.f);
        st8(da1T + (size_t)(p * kC + ch) * B + m0 + 4 * lg, o[0], o[1], o[2], o[3]);
    }
}

/* dW2[o][c][t] += sum_r sum_q a1[r][q + off(t)][c] da2[r][q][o].  One wave per (t, q). */
__global__ void __launch_bounds__(64) conv2_dw_kernel(const uint16_t *a1T, const uint16_t *da2T, int B,
                                                      float *grad) {
    const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4;
    const int t = blockIdx.x, q = blockIdx.y;
    const int qy = q / kS2, qx = q - qy * kS2;
    const int pos = (qy + t / 3) * kS1 + qx + t % 3;
    f32x4_t c[2][2] = {{{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}, {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}};
    const uint16_t *A0 = a1T + (size_t)(pos * kC + lr) * B + lg * 8;
    const uint16_t *A1 = a1T + (size_t)(pos * kC + 16 + lr) * B + lg * 8;
    const uint16_t *B0 = da2T + (size_t)(q * kC + lr) * B + lg * 8;
    const uint16_t *B1 = da2T + (size_t)(q * kC + 16 + lr) * B + lg * 8;
    for (int k = 0; k < B; k += 32) {
        const u32x4_t a0 = ld16(A0 + k), a1 = ld16(A1 + k), b0 = ld16(B0 + k), b1 = ld16(B1 + k);
        c[0][0] = mfma32(a0, b0, c[0][0]);
        c[0][1] = mfma32(a0, b1, c[0][1]);
        c[1][0] = mfma32(a1, b0, c[1][0]);
        c[1][1] = mfma32(a1, b1, c[1][1]);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int ch = mt * 16 + 4 * lg + i, o = nt * 16 + lr;
                atomicAdd(&grad[OFF_W2 + o * 288 + ch * 9 + t], c[mt][nt][i]);
            }
}

/* dW1[c][t] += sum_r sum_p x[r][p + off(t)] da1[r][p][c].  One wave per position p. */
__global__ void __launch_bounds__(64) conv1_dw_kernel(const uint16_t *xT, const uint16_t *da1T, int B,
                                                      float *grad) {
    const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4;
    const int p = blockIdx.x;
    const int py = p / kS1, px = p - py * kS1;
    const int tap = lr < 9 ? lr : 0;
    const int pix = (py + tap / 3) * kN + px + tap % 3;
    const uint16_t *A = xT + (size_t)pix * B + lg * 8;
    const uint16_t *B0 = da1T + (size_t)(p * kC + lr) * B + lg * 8;
    const uint16_t *B1 = da1T + (size_t)(p * kC + 16 + lr) * B + lg * 8;
    f32x4_t c0 = {0.f, 0.f, 0.f, 0.f}, c1 = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < B; k += 32) {
        u32x4_t a = ld16(A + k);
        if (lr >= 9) a = (u32x4_t){0u, 0u, 0u, 0u};
        c0 = mfma32(a, ld16(B0 + k), c0);
        c1 = mfma32(a, ld16(B1 + k), c1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int t = 4 * lg + i;
        if (t < 9) {
            atomicAdd(&grad[OFF_W1 + lr * 9 + t], c0[i]);
            atomicAdd(&grad[OFF_W1 + (16 + lr) * 9 + t], c1[i]);
        }
    }
}

/* Bias gradient: grad[off + c] += sum over positions and patches of g[(p*32 + c)*B + r].
 * One block per (p, c). */
__global__ void __launch_bounds__(256) bias_grad_kernel(const uint16_t *gT, int B, float *grad, int off) {
    __shared__ float s[256];
    const int row = blockIdx.x; /* p*32 + c */
    float acc = 0.f;
    for (int r = threadIdx.x; r < B; r += 256) acc += bf2f(gT[(size_t)row * B + r]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&grad[off + (row & (kC - 1))], s[0]);
}

/* ---- optimizer and packing ------------------------------------------------------------ */

__global__ void zero_kernel(float *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void adam_kernel(float *p, float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* Rebuilds every BF16 operand buffer from the FP32 master parameters. */
__global__ void pack_kernel(const float *P, uint16_t *pk, float *wf2f) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 5 * kHid) wf2f[e] = bf2f(f2bf(P[OFF_WF2 + e]));
    if (e >= PK_TOTAL) return;
    const int idx = e;
    float v = 0.f;
    if (e < PK_W1) {
        const int nt = e / 256, lane = (e / 4) % 64, i = e % 4;
        const int k = 4 * (lane >> 4) + i, c = nt * 16 + (lane & 15);
        v = k < 9 ? P[OFF_W1 + c * 9 + k] : 0.f;
    } else if ((e -= PK_W1) < PK_W2) {
        const int t = e / 1024, nt = (e / 512) % 2, lane = (e / 8) % 64, i = e % 8;
        const int c = 8 * (lane >> 4) + i, o = nt * 16 + (lane & 15);
        v = P[OFF_W2 + o * 288 + c * 9 + t];
    } else if ((e -= PK_W2) < PK_W2T) {
        const int t = e / 1024, nt = (e / 512) % 2, lane = (e / 8) % 64, i = e % 8;
        const int o = 8 * (lane >> 4) + i, c = nt * 16 + (lane & 15);
        v = P[OFF_W2 + o * 288 + c * 9 + t];
    } else if ((e -= PK_W2T) < PK_WF1) {
        /* A operand of the transposed fc1: row j, k-slot (g, i) -> channel o(g, i) of chunk q */
        const int q = e / 2048, mt = (e / 512) % 4, lane = (e / 8) % 64, i = e % 8;
        const int g = lane >> 4;
        const int o = i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4);
        const int k = q * 32 + o, j = mt * 16 + (lane & 15);
        v = P[OFF_WF1 + j * kF + k];
    } else if ((e -= PK_WF1) < PK_WF1T) {
        const int nt = e / 1024, kc = (e / 512) % 2, lane = (e / 8) % 64, i = e % 8;
        const int j = kc * 32 + 8 * (lane >> 4) + i, kf = nt * 16 + (lane & 15);
        v = P[OFF_WF1 + j * kF + kf];
    } else {
        e -= PK_WF1T;
        const int kc = e / 512, lane = (e / 8) % 64, i = e % 8;
        const int j = kc * 32 + 8 * (lane >> 4) + i, t = lane & 15;
        v = t < kOut ? P[OFF_WF2 + t * kHid + j] : 0.f;
    }
    pk[idx] = f2bf(v);
}

}  // namespace lzr_net

using namespace lzr_net;

/* ------------------------------------------------------------------------------------ */
/* C ABI                                                                                  */
/* ------------------------------------------------------------------------------------ */

struct LzrNet {
    int device = 0;
    int num_cus = 0;
    double w = 1.0;
    std::string err;
    std::mutex mu;
    float *d_params = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_pack = nullptr;
    float *d_wf2f = nullptr;
    float *d_loss = nullptr;
    unsigned *d_counter = nullptr;
    long adam_t = 0;
    /* per-batch buffers */
    size_t rows_cap = 0;
    uint16_t *d_xin = nullptr;
    PatchStat *d_stats = nullptr;
    float *d_out = nullptr, *d_tgt = nullptr, *d_wt = nullptr;
    size_t train_cap = 0;
    uint16_t *d_xT = nullptr, *d_a1T = nullptr, *d_a2T = nullptr, *d_h = nullptr;
    uint16_t *d_dh = nullptr, *d_dhT = nullptr, *d_da2 = nullptr, *d_da2T = nullptr, *d_da1T = nullptr;
    LzrStreamOrder order; /* calls on different streams share the buffers above */
};

static thread_local std::string g_net_create_error;

template <typename T>
static void net_alloc(T *&p, size_t n) {
    if (p) NET_CHECK(hipFree(p));
    p = nullptr;
    NET_CHECK(hipMalloc((void **)&p, std::max(n, (size_t)1) * sizeof(T)));
}

static Packed packed(LzrNet *net) {
    Packed W;
    W.w1 = net->d_pack;
    W.w2 = W.w1 + PK_W1;
    W.w2t = W.w2 + PK_W2;
    W.wf1 = W.w2t + PK_W2T;
    W.wf1t = W.wf1 + PK_WF1;
    W.wf2 = W.wf1t + PK_WF1T;
    W.wf2f = net->d_wf2f;
    W.params = net->d_params;
    return W;
}

static void repack(LzrNet *net, hipStream_t s) {
    const int n = std::max(PK_TOTAL, 5 * kHid);
    hipLaunchKernelGGL(pack_kernel, dim3((n + 255) / 256), dim3(256), 0, s, net->d_params, net->d_pack, net->d_wf2f);
}

static void ensure_rows(LzrNet *net, size_t rows) {
    if (rows <= net->rows_cap) return;
    net_alloc(net->d_xin, rows * kXRow);
    net_alloc(net->d_stats, rows);
    net_alloc(net->d_out, rows * kOut);
    net_alloc(net->d_tgt, rows * kOut);
    net_alloc(net->d_wt, rows);
    net->rows_cap = rows;
}

static void ensure_train(LzrNet *net, size_t B) {
    if (B <= net->train_cap) return;
    net_alloc(net->d_xT, B * kXRow);
    net_alloc(net->d_a1T, B * kP1 * kC);
    net_alloc(net->d_a2T, B * kF);
    net_alloc(net->d_h, B * kHid);
    net_alloc(net->d_dh, B * kHid);
    net_alloc(net->d_dhT, B * kHid);
    net_alloc(net->d_da2, B * kF);
    net_alloc(net->d_da2T, B * kF);
    net_alloc(net->d_da1T, B * kP1 * kC);
    net->train_cap = B;
}

static int net_fail(LzrNet *net, const std::string &m) {
    if (net) net->err = m;
    return -1;
}

/* splitmix64 + Box-Muller: deterministic He-normal initialization */
static uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double gauss(uint64_t &s) {
    const double u1 = ((splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
static void he_init(std::vector<float> &P, uint64_t seed) {
    P.assign(NPARAMS, 0.f);
    uint64_t s = seed;
    auto fill = [&](int off, int count, int fan_in) {
        const double sd = std::sqrt(2.0 / fan_in);
        for (int i = 0; i < count; ++i) P[off + i] = (float)(sd * gauss(s));
    };
    fill(OFF_W1, 288, 9);
    fill(OFF_W2, 9216, 288);
    fill(OFF_WF1, 64 * kF, kF);
    fill(OFF_WF2, 320, 64);
    for (int i = 0; i < 320; ++i) P[OFF_WF2 + i] *= 0.1f; /* small initial outputs */
}

static void upload_params(LzrNet *net, const float *P) {
    NET_CHECK(hipMemcpy(net->d_params, P, NPARAMS * sizeof(float), hipMemcpyHostToDevice));
    NET_CHECK(hipMemset(net->d_grad, 0, NPARAMS * sizeof(float)));
    NET_CHECK(hipMemset(net->d_m, 0, NPARAMS * sizeof(float)));
    NET_CHECK(hipMemset(net->d_v, 0, NPARAMS * sizeof(float)));
    net->adam_t = 0;
    repack(net, nullptr);
    NET_CHECK(hipGetLastError());
    NET_CHECK(hipDeviceSynchronize());
}

LZR_EXPORT int lzr_net_create(LzrNet **out, double initial_psf_width, unsigned long long seed) {
    if (!out) {
        g_net_create_error = "lzr_net_create: null output pointer";
        return -1;
    }
    *out = nullptr;
    if (!(initial_psf_width > 0.0)) {
        g_net_create_error = "lzr_net_create: initialPSFWidth must be positive";
        return -1;
    }
    LzrNet *net = new LzrNet();
    net->w = initial_psf_width;
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::string("no HIP device available");
        NET_CHECK(hipGetDevice(&net->device));
        hipDeviceProp_t prop;
        NET_CHECK(hipGetDeviceProperties(&prop, net->device));
        net->num_cus = prop.multiProcessorCount;
        net_alloc(net->d_params, NPARAMS);
        net_alloc(net->d_grad, NPARAMS);
        net_alloc(net->d_m, NPARAMS);
        net_alloc(net->d_v, NPARAMS);
        net_alloc(net->d_pack, PK_TOTAL);
        net_alloc(net->d_wf2f, 5 * kHid);
        net_alloc(net->d_loss, 1);
        net_alloc(net->d_counter, 1);
        std::vector<float> P;
        he_init(P, seed);
        upload_params(net, P.data());
    } catch (const std::string &e) {
        g_net_create_error = "lzr_net_create: " + e;
        lzr_net_destroy(net);
        return -1;
    }
    *out = net;
    return 0;
}

LZR_EXPORT int lzr_net_destroy(LzrNet *net) {
    if (!net) return -1;
    void *ptrs[] = {net->d_params, net->d_grad, net->d_m, net->d_v, net->d_pack, net->d_wf2f, net->d_loss,
                    net->d_counter, net->d_xin, net->d_stats, net->d_out, net->d_tgt, net->d_wt, net->d_xT,
                    net->d_a1T, net->d_a2T, net->d_h, net->d_dh, net->d_dhT, net->d_da2, net->d_da2T, net->d_da1T};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    net->order.destroy();
    delete net;
    return 0;
}

LZR_EXPORT int lzr_net_last_error(const LzrNet *net, char *buf, int len) {
    const std::string &e = net ? net->err : g_net_create_error;
    if (buf && len > 0) {
        const int n = std::min((int)e.size(), len - 1);
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}

LZR_EXPORT int lzr_net_set_params(LzrNet *net, const float *params, int count) {
    if (!net) return -1;
    std::lock_guard<std::mutex> lock(net->mu);
    if (!params || count != NPARAMS) return net_fail(net, "lzr_net_set_params: expected LZR_NET_NPARAMS floats");
    try {
        NET_CHECK(hipSetDevice(net->device));
        upload_params(net, params);
    } catch (const std::string &e) {
        return net_fail(net, e);
    }
    return 0;
}

LZR_EXPORT int lzr_net_get_params(LzrNet *net, float *params, int count) {
    if (!net) return -1;
    std::lock_guard<std::mutex> lock(net->mu);
    if (!params || count != NPARAMS) return net_fail(net, "lzr_net_get_params: expected LZR_NET_NPARAMS floats");
    try {
        NET_CHECK(hipSetDevice(net->device));
        NET_CHECK(hipDeviceSynchronize());
        NET_CHECK(hipMemcpy(params, net->d_params, NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
    } catch (const std::string &e) {
        return net_fail(net, e);
    }
    return 0;
}

static bool net_shape_ok(LzrNet *net, const void *frames, int nx, int ny, int nframes, const void *cands,
                         int ncands) {
    if (nx < 1 || ny < 1 || nframes < 0 || ncands < 0) {
        net->err = "invalid frame or candidate counts";
        return false;
    }
    if (ncands > 0 && (!frames || !cands)) {
        net->err = "null device pointer";
        return false;
    }
    return true;
}

/* prep + forward for rows [0, nrows) (rows >= ncands are padding) */
template <bool kTrain>
static void run_forward(LzrNet *net, const uint16_t *frames, int nx, int ny, int nframes, const LzrCandidate *cands,
                        int ncands, int nrows, hipStream_t s, double k_start = 0.0, double *start_sse = nullptr,
                        const int *sel_idx = nullptr, const unsigned *sel_count = nullptr) {
    ensure_rows(net, (size_t)nrows);
    const int pblocks = (nrows + 7) / 8;
    hipLaunchKernelGGL(prep_kernel, dim3(sel_count ? std::min(pblocks, net->num_cus * 32) : pblocks), dim3(256), 0, s,
                       frames, nx, ny, nframes, cands, ncands, nrows, net->d_xin, net->d_stats, k_start, start_sse,
                       sel_idx, sel_count);
    TrainStore T{};
    if (kTrain) {
        T.xT = net->d_xT;
        T.a1T = net->d_a1T;
        T.a2T = net->d_a2T;
        T.h = net->d_h;
        T.ldb = nrows;
    }
    const int ngroups = (nrows + kGroup - 1) / kGroup;
    const int grid = std::max(1, std::min(ngroups, net->num_cus));
    hipLaunchKernelGGL((forward_kernel<kTrain>), dim3(grid), dim3(kThreads), 0, s, net->d_xin, nrows, packed(net),
                       net->d_out, T, sel_count);
}

LZR_EXPORT int lzr_net_forward_device(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                                      const LzrCandidate *d_cands, int ncands, float *d_outputs, void *stream) {
    if (!net) return -1;
    std::lock_guard<std::mutex> lock(net->mu);
    if (!net_shape_ok(net, d_frames, nx, ny, nframes, d_cands, ncands)) return -1;
    if (ncands == 0) return 0;
    if (!d_outputs) return net_fail(net, "null device pointer");
    try {
        NET_CHECK(hipSetDevice(net->device));
        hipStream_t s = (hipStream_t)stream;
        NET_CHECK(net->order.begin(s));
        run_forward<false>(net, d_frames, nx, ny, nframes, d_cands, ncands, ncands, s);
        hipLaunchKernelGGL(decode_kernel, dim3((ncands + 255) / 256), dim3(256), 0, s, net->d_out, net->d_stats,
                           ncands, net->w, (LzrStart *)nullptr, d_outputs);
        NET_CHECK(hipGetLastError());
        NET_CHECK(net->order.end(s));
    } catch (const std::string &e) {
        return net_fail(net, e);
    }
    return 0;
}

int lzr_net_predict_start_sse(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                              const LzrCandidate *d_cands, int ncands, LzrStart *d_starts, double w_start,
                              double *d_start_sse, const int *d_sel, const unsigned *d_nsel, void *stream) {
    if (!net) return -1;
    std::lock_guard<std::mutex> lock(net->mu);
    if (!net_shape_ok(net, d_frames, nx, ny, nframes, d_cands, ncands)) return -1;
    if (ncands == 0) return 0;
    if (!d_starts) return net_fail(net, "null device pointer");
    try {
        NET_CHECK(hipSetDevice(net->device));
        hipStream_t s = (hipStream_t)stream;
        NET_CHECK(net->order.begin(s));
        run_forward<false>(net, d_frames, nx, ny, nframes, d_cands, ncands, ncands, s, 0.5 / (w_start * w_start),
                           d_start_sse, d_sel, d_nsel);
        hipLaunchKernelGGL(decode_kernel, dim3((ncands + 255) / 256), dim3(256), 0, s, net->d_out, net->d_stats,
                           ncands, net->w, d_starts, (float *)nullptr, d_sel, d_nsel);
        NET_CHECK(hipGetLastError());
        NET_CHECK(net->order.end(s));
    } catch (const std::string &e) {
        return net_fail(net, e);
    }
    return 0;
}

LZR_EXPORT int lzr_net_predict_device(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                                      const LzrCandidate *d_cands, int ncands, LzrStart *d_starts, void *stream) {
    return lzr_net_predict_start_sse(net, d_frames, nx, ny, nframes, d_cands, ncands, d_starts, 1.0, nullptr, nullptr,
                                     nullptr, stream);
}

LZR_EXPORT int lzr_net_train_step_device(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                                         const LzrCandidate *d_cands, const LzrFit *d_labels, int ncands,
                                         float learning_rate, int apply, float *loss, float *grads, void *stream) {
    if (!net) return -1;
    std::lock_guard<std::mutex> lock(net->mu);
    if (!net_shape_ok(net, d_frames, nx, ny, nframes, d_cands, ncands)) return -1;
    if (ncands == 0) return net_fail(net, "empty training batch");
    if (!d_labels) return net_fail(net, "null device pointer");
    if (ncands > (1 << 24)) return net_fail(net, "training batch too large");
    try {
        NET_CHECK(hipSetDevice(net->device));
        hipStream_t s = (hipStream_t)stream;
        const int B = (ncands + 63) / 64 * 64; /* padded rows carry weight 0 */
        ensure_rows(net, (size_t)B);
        ensure_train(net, (size_t)B);
        NET_CHECK(net->order.begin(s));
        /* fill kernels, not memsets: the step stays capturable in a HIP graph */
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, net->d_loss, 1);
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<float *>(net->d_counter), 1);
        hipLaunchKernelGGL(zero_kernel, dim3((NPARAMS + 255) / 256), dim3(256), 0, s, net->d_grad, NPARAMS);
        run_forward<true>(net, d_frames, nx, ny, nframes, d_cands, ncands, B, s);
        hipLaunchKernelGGL(targets_kernel, dim3((B + 255) / 256), dim3(256), 0, s, d_labels, net->d_stats, ncands, B,
                           net->w, net->d_tgt, net->d_wt, net->d_counter);
        const Packed W = packed(net);
        hipLaunchKernelGGL(head_backward_kernel, dim3((B + 255) / 256), dim3(256), 0, s, net->d_out, net->d_tgt,
                           net->d_wt, net->d_counter, net->d_h, W.wf2f, B, net->d_grad, net->d_loss, net->d_dh,
                           net->d_dhT);
        const int mtiles = B / 16;
        hipLaunchKernelGGL(fc1_dx_kernel, dim3((mtiles + 3) / 4, 7), dim3(256), 0, s, net->d_dh, W.wf1t, net->d_a2T, B,
                           net->d_da2, net->d_da2T);
        const int kslice = 1024;
        hipLaunchKernelGGL(fc1_dw_kernel, dim3(98, (B + kslice - 1) / kslice), dim3(256), 0, s, net->d_dhT,
                           net->d_a2T, B, kslice, net->d_grad);
        hipLaunchKernelGGL(bias_grad_kernel, dim3(kP2 * kC), dim3(256), 0, s, net->d_da2T, B, net->d_grad, OFF_B2);
        hipLaunchKernelGGL(conv2_dx_kernel, dim3((mtiles + 3) / 4, kP1), dim3(256), 0, s, net->d_da2, W.w2t,
                           net->d_a1T, B, net->d_da1T);
        hipLaunchKernelGGL(conv2_dw_kernel, dim3(9, kP2), dim3(64), 0, s, net->d_a1T, net->d_da2T, B, net->d_grad);
        hipLaunchKernelGGL(conv1_dw_kernel, dim3(kP1), dim3(64), 0, s, net->d_xT, net->d_da1T, B, net->d_grad);
        hipLaunchKernelGGL(bias_grad_kernel, dim3(kP1 * kC), dim3(256), 0, s, net->d_da1T, B, net->d_grad, OFF_B1);
        if (apply) {
            net->adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - (float)std::pow(0.9, (double)net->adam_t);
            const float c2 = 1.f - (float)std::pow(0.999, (double)net->adam_t);
            hipLaunchKernelGGL(adam_kernel, dim3((NPARAMS + 255) / 256), dim3(256), 0, s, net->d_params, net->d_grad,
                               net->d_m, net->d_v, NPARAMS, learning_rate, b1, b2, 1e-8f, c1, c2);
            repack(net, s);
        }
        NET_CHECK(hipGetLastError());
        if (loss) {
            NET_CHECK(hipMemcpyAsync(loss, net->d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
        }
        if (grads) {
            NET_CHECK(hipMemcpyAsync(grads, net->d_grad, NPARAMS * sizeof(float), hipMemcpyDeviceToHost, s));
        }
        NET_CHECK(net->order.end(s));
        if (loss || grads) NET_CHECK(hipStreamSynchronize(s));
    } catch (const std::string &e) {
        return net_fail(net, e);
    }
    return 0;
}
