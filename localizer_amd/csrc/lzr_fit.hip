/*
 * lzr_fit.hip -- persistent LM fitting kernels for gfx950 and the C ABI of
 * include/localizer_hip.h.
 *
 * Layout and schedule:
 *   - one ROI per lane; a workgroup is one wave64 that owns a private LDS tile of
 *     (2 h_max + 1)^2 x 64 uint16 pixels, pixel k of lane L at tile[k*64 + L]
 *     (consecutive lanes read consecutive halfwords: conflict-free);
 *   - the grid is exactly the resident capacity (occupancy query), each wave pulls
 *     candidates from a device work queue (one atomicAdd per refill, shared by all
 *     refilling lanes) and refills lanes as soon as their fit ends, so lanes with
 *     long LM runs do not stall the others;
 *   - a refilled lane's window is gathered cooperatively by the whole wave (rows of
 *     n contiguous uint16 -> coalesced global loads) into its LDS column;
 *   - each loop iteration is one pixel pass for all active lanes: residuals at the
 *     current and trial points, the residual change, the Jacobian at the trial point
 *     and its normal equations, all in registers (lzr_lane.h);
 *   - the per-lane LM logic (lmpar, ratio test, status) runs in FP64 between passes.
 * Candidates are split by parameter count (5: symmetric, 7: ellipsoidal) on the device
 * and each class runs its own kernel instantiation.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lzr_lane.h"

#define LZR_CHECK(call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) throw std::string(#call ": ") + hipGetErrorString(e_);         \
    } while (0)

namespace lzr {

struct KernelArgs {
    const uint16_t *frames;
    int nx, ny, nframes;
    const LzrCandidate *cands;
    const int *order;      /* queue position -> candidate index   */
    const unsigned *count; /* number of queue entries (device)    */
    unsigned *head;        /* queue head (device, zeroed per call) */
    LzrFit *out;
    LaneConfig cfg;
};

__device__ __forceinline__ int model_params(int model) { return model == LZR_MODEL_SYMMETRIC ? 5 : 7; }

/* Splits candidates into the 5- and 7-parameter queues. counters[0..1] = sizes,
 * counters[2..3] = queue heads (zeroed here for the fitting kernels). */
__global__ void classify_kernel(const LzrCandidate *cands, int ncands, int default_model, int *order,
                                unsigned *counters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int cls = -1;
    if (i < ncands) {
        int model = cands[i].model;
        if (model < 0) model = default_model;
        cls = model_params(model) == 5 ? 0 : 1;
    }
    /* one atomic per wave and class */
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const unsigned long long m = __ballot(cls == c);
        if (!m) continue;
        const int leader = __ffsll((long long)m) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&counters[c], (unsigned)__popcll(m));
        base = __shfl(base, leader, 64);
        if (cls == c) {
            const unsigned slot = base + (unsigned)__popcll(m & below);
            order[c == 0 ? slot : ncands - 1 - slot] = i;
        }
    }
}

/* Turns the raw end states into the reference records (post-checks, covariance,
 * 2x2 eigen-decomposition), one thread per candidate. */
__global__ void epilogue_kernel(const LzrCandidate *cands, int ncands, LzrFit *out, LaneConfig cfg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncands) return;
    finish_any(out[i], cfg, cands[i]);
}

__global__ void reset_counters(unsigned *counters) {
    if (threadIdx.x < 4) counters[threadIdx.x] = 0u;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

/* W = minimum waves per SIMD the register allocation must allow (occupancy target). */
template <int P, typename Acc, bool F64, int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) lm_persistent(KernelArgs args, int cls, int ncands_total) {
    extern __shared__ uint16_t tile[];
    const int lane = threadIdx.x;
    const unsigned count = args.count[cls];
    const int *order = args.order + (cls == 0 ? 0 : ncands_total - (int)count);

    LaneState<P> s;
    s.phase = kIdle;
    s.n = 0;
    int model = 0;
    bool open = true; /* wave-uniform: the queue may still hold work */

    for (;;) {
        /* ---- refill idle lanes from the queue ---- */
        while (open) {
            const unsigned long long want = __ballot(s.phase == kIdle);
            if (!want) break;
            const int cnt = __popcll(want);
            const int leader = __ffsll((long long)want) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(args.head + cls, (unsigned)cnt);
            base = __shfl(base, leader, 64);
            if (base + (unsigned)cnt >= count) open = false;
            if (base >= count) break;
            if (s.phase == kIdle) {
                const unsigned rank = __popcll(want & ((1ull << lane) - 1ull));
                const unsigned q = base + rank;
                if (q < count) {
                    const int idx = order[q];
                    const LzrCandidate c = args.cands[idx];
                    model = c.model >= 0 ? c.model : args.cfg.model;
                    if (!lane_start<P>(s, args.cfg, c, idx, args.nx, args.ny, args.nframes, model, args.out[idx]))
                        s.phase = kIdle;
                } else {
                    s.phase = kFinished;
                }
            }
        }
        if (!open && s.phase == kIdle) s.phase = kFinished;

        /* ---- gather the windows of freshly started lanes ---- */
        unsigned long long fresh = __ballot(s.phase == kInit);
        while (fresh) {
            const int L = __ffsll((long long)fresh) - 1;
            fresh &= fresh - 1ull;
            const int fL = __shfl(s.frame, L, 64);
            const int xL = __shfl(s.xoff, L, 64);
            const int yL = __shfl(s.yoff, L, 64);
            const int nL = __shfl(s.n, L, 64);
            const uint16_t *src = args.frames + ((size_t)fL * args.ny + yL) * args.nx + xL;
            for (int k = lane; k < nL * nL; k += 64) {
                const int j = k / nL, i = k - j * nL;
                tile[k * 64 + L] = src[(size_t)j * args.nx + i];
            }
        }
        __syncthreads();

        const bool active = s.phase == kInit || s.phase == kTrial;
        if (!__any(active)) break;

        /* ---- one pixel pass ---- */
        PassIn<P> q;
        PassAcc<P, Acc> acc;
        acc.zero();
        const int n = active ? s.n : 0;
        if (active) lane_pass_in<P>(s, args.cfg.sigma, q);
        const int nw = wave_max(n);
        for (int j = 0; j < nw; ++j) {
            const float fj = (float)j;
            for (int i = 0; i < nw; ++i) {
                if (i < n && j < n) {
                    const float D = (float)tile[(i + j * n) * 64 + lane];
                    pass_pixel<F64>(q, (float)i, fj, D, acc);
                }
            }
        }

        /* ---- LM logic ---- */
        if (active) {
            if (!lane_after_pass<P, Acc>(s, args.cfg, acc, model, *reinterpret_cast<LzrRawFit *>(args.out + s.roi)))
                s.phase = open ? kIdle : kFinished;
        }
        __syncthreads();
    }
}

}  // namespace lzr

using namespace lzr;

/* ------------------------------------------------------------------------------ */
/* C ABI                                                                            */
/* ------------------------------------------------------------------------------ */

struct LzrContext {
    int device = 0;
    LaneConfig cfg{};
    int waves_per_cu = 0;
    int occupancy5 = 2; /* register budget of the 5-parameter kernel: 1 or 2 waves/SIMD */
    int num_cus = 0;
    std::string err;
    std::mutex mu;
    /* device buffers */
    uint16_t *d_frames = nullptr;
    size_t frames_cap = 0;
    LzrCandidate *d_cands = nullptr;
    size_t cands_cap = 0;
    LzrFit *d_out = nullptr;
    size_t out_cap = 0;
    int *d_order = nullptr;
    size_t order_cap = 0;
    unsigned *d_counters = nullptr;
    hipStream_t stream = nullptr;
};

static thread_local std::string g_create_error;

template <typename T>
static void ensure(T *&ptr, size_t &cap, size_t need) {
    if (need <= cap) return;
    if (ptr) LZR_CHECK(hipFree(ptr));
    ptr = nullptr;
    cap = 0;
    const size_t n = std::max(need, (size_t)1);
    LZR_CHECK(hipMalloc((void **)&ptr, n * sizeof(T)));
    cap = n;
}

template <int P, typename Acc, bool F64, int W>
static int grid_for(LzrContext *ctx, size_t lds) {
    int per_cu = 0;
    LZR_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lm_persistent<P, Acc, F64, W>, 64, lds));
    if (per_cu < 1) throw std::string("fitting kernel cannot be resident with this window size");
    if (ctx->waves_per_cu > 0) per_cu = std::min(per_cu, ctx->waves_per_cu);
    return per_cu * ctx->num_cus;
}

typedef double AccT;           /* normal-equation accumulation type      */
constexpr bool kF64Residual = true; /* current-point residual in FP64 */

static void launch(LzrContext *ctx, const uint16_t *frames, int nx, int ny, int nframes,
                   const LzrCandidate *cands, int ncands, LzrFit *out, hipStream_t stream) {
    ensure(ctx->d_order, ctx->order_cap, (size_t)ncands);
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, stream, ctx->d_counters);
    const int tpb = 256;
    hipLaunchKernelGGL(classify_kernel, dim3((ncands + tpb - 1) / tpb), dim3(tpb), 0, stream, cands, ncands,
                       ctx->cfg.model, ctx->d_order, ctx->d_counters);
    KernelArgs a;
    a.frames = frames;
    a.nx = nx;
    a.ny = ny;
    a.nframes = nframes;
    a.cands = cands;
    a.order = ctx->d_order;
    a.count = ctx->d_counters;
    a.head = ctx->d_counters + 2;
    a.out = out;
    a.cfg = ctx->cfg;
    const int nmax = 2 * ctx->cfg.max_half_width + 1;
    const size_t lds = (size_t)nmax * nmax * 64 * sizeof(uint16_t);
    if (ctx->occupancy5 == 1) {
        const int g5 = grid_for<5, AccT, kF64Residual, 1>(ctx, lds);
        hipLaunchKernelGGL((lm_persistent<5, AccT, kF64Residual, 1>), dim3(g5), dim3(64), lds, stream, a, 0, ncands);
    } else {
        const int g5 = grid_for<5, AccT, kF64Residual, 2>(ctx, lds);
        hipLaunchKernelGGL((lm_persistent<5, AccT, kF64Residual, 2>), dim3(g5), dim3(64), lds, stream, a, 0, ncands);
    }
    const int g7 = grid_for<7, AccT, kF64Residual, 1>(ctx, lds);
    hipLaunchKernelGGL((lm_persistent<7, AccT, kF64Residual, 1>), dim3(g7), dim3(64), lds, stream, a, 1, ncands);
    hipLaunchKernelGGL(epilogue_kernel, dim3((ncands + tpb - 1) / tpb), dim3(tpb), 0, stream, cands, ncands, out,
                       ctx->cfg);
    LZR_CHECK(hipGetLastError());
}

static int fail(LzrContext *ctx, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return -1;
}

LZR_EXPORT int lzr_abi_version(void) { return LZR_ABI_VERSION; }

LZR_EXPORT int lzr_create(LzrContext **out, int model, double initial_psf_width, double sigma,
                          int default_half_width) {
    if (!out) {
        g_create_error = "lzr_create: null output pointer";
        return -1;
    }
    *out = nullptr;
    if (model != LZR_MODEL_SYMMETRIC && model != LZR_MODEL_ELLIPSOIDAL &&
        model != LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF) {
        g_create_error = "lzr_create: unknown model";
        return -1;
    }
    if (!(initial_psf_width > 0.0) || !(sigma > 0.0)) {
        g_create_error = "lzr_create: initialPSFWidth and sigma must be positive";
        return -1;
    }
    const int h = default_half_width > 0 ? default_half_width : (int)std::ceil(initial_psf_width * 4.0);
    if (h > LZR_MAX_HALF_WIDTH) {
        g_create_error = "lzr_create: window half-width exceeds LZR_MAX_HALF_WIDTH";
        return -1;
    }
    LzrContext *ctx = new LzrContext();
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::string("no HIP device available");
        LZR_CHECK(hipGetDevice(&ctx->device));
        hipDeviceProp_t prop;
        LZR_CHECK(hipGetDeviceProperties(&prop, ctx->device));
        ctx->num_cus = prop.multiProcessorCount;
        LZR_CHECK(hipMalloc((void **)&ctx->d_counters, 4 * sizeof(unsigned)));
    } catch (const std::string &e) {
        g_create_error = "lzr_create: " + e;
        delete ctx;
        return -1;
    }
    ctx->cfg.initial_psf_width = initial_psf_width;
    ctx->cfg.sigma = sigma;
    ctx->cfg.epsabs = 1e-4;
    ctx->cfg.epsrel = 1e-4;
    ctx->cfg.max_iterations = 200;
    ctx->cfg.default_half_width = h;
    ctx->cfg.max_half_width = h;
    ctx->cfg.diag_update_rows = -1;
    ctx->cfg.model = model;
    *out = ctx;
    return 0;
}

LZR_EXPORT int lzr_destroy(LzrContext *ctx) {
    if (!ctx) return -1;
    (void)hipFree(ctx->d_frames);
    (void)hipFree(ctx->d_cands);
    (void)hipFree(ctx->d_out);
    (void)hipFree(ctx->d_order);
    (void)hipFree(ctx->d_counters);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

LZR_EXPORT int lzr_last_error(const LzrContext *ctx, char *buf, int len) {
    const std::string &e = ctx ? ctx->err : g_create_error;
    if (buf && len > 0) {
        const int n = std::min((int)e.size(), len - 1);
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}

LZR_EXPORT int lzr_set_option(LzrContext *ctx, int option, double value) {
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mu);
    switch (option) {
    case LZR_OPT_MAX_ITERATIONS:
        if (!(value >= 1 && value <= 32767)) return fail(ctx, "max iterations out of range");
        ctx->cfg.max_iterations = (int)value;
        return 0;
    case LZR_OPT_EPSABS:
        if (!(value >= 0)) return fail(ctx, "epsabs must be >= 0");
        ctx->cfg.epsabs = value;
        return 0;
    case LZR_OPT_EPSREL:
        if (!(value >= 0)) return fail(ctx, "epsrel must be >= 0");
        ctx->cfg.epsrel = value;
        return 0;
    case LZR_OPT_DIAG_UPDATE_ROWS:
        if (!(value >= -1 && value <= 1 << 20)) return fail(ctx, "diag update rows out of range");
        ctx->cfg.diag_update_rows = (int)value;
        return 0;
    case LZR_OPT_WAVES_PER_CU:
        if (!(value >= 0 && value <= 64)) return fail(ctx, "waves per CU out of range");
        ctx->waves_per_cu = (int)value;
        return 0;
    case LZR_OPT_OCCUPANCY:
        if (!(value == 1 || value == 2)) return fail(ctx, "occupancy must be 1 or 2");
        ctx->occupancy5 = (int)value;
        return 0;
    case LZR_OPT_MAX_HALF_WIDTH:
        if (!(value >= 1 && value <= LZR_MAX_HALF_WIDTH)) return fail(ctx, "max half-width out of range");
        ctx->cfg.max_half_width = std::max((int)value, ctx->cfg.default_half_width);
        return 0;
    default:
        return fail(ctx, "unknown option");
    }
}

static bool check_shape(LzrContext *ctx, int nx, int ny, int nframes, int ncands) {
    if (nx < 1 || ny < 1 || nframes < 0 || ncands < 0) {
        ctx->err = "invalid frame or candidate counts";
        return false;
    }
    return true;
}

LZR_EXPORT int lzr_fit_batch_device(LzrContext *ctx, const uint16_t *d_frames, int nx, int ny, int nframes,
                                    const LzrCandidate *d_cands, int ncands, LzrFit *d_out, void *stream) {
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!check_shape(ctx, nx, ny, nframes, ncands)) return -1;
    if (ncands == 0) return 0;
    if (!d_frames || !d_cands || !d_out) return fail(ctx, "null device pointer");
    try {
        LZR_CHECK(hipSetDevice(ctx->device));
        launch(ctx, d_frames, nx, ny, nframes, d_cands, ncands, d_out, (hipStream_t)stream);
    } catch (const std::string &e) {
        return fail(ctx, e);
    }
    return 0;
}

LZR_EXPORT int lzr_fit_batch(LzrContext *ctx, const uint16_t *frames, int nx, int ny, int nframes,
                             const LzrCandidate *cands, int ncands, LzrFit *results) {
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!check_shape(ctx, nx, ny, nframes, ncands)) return -1;
    if (ncands == 0) return 0;
    if (!frames || !cands || !results) return fail(ctx, "null host pointer");
    try {
        LZR_CHECK(hipSetDevice(ctx->device));
        if (!ctx->stream) LZR_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        const size_t npx = (size_t)nx * ny * nframes;
        ensure(ctx->d_frames, ctx->frames_cap, npx);
        ensure(ctx->d_cands, ctx->cands_cap, (size_t)ncands);
        ensure(ctx->d_out, ctx->out_cap, (size_t)ncands);
        LZR_CHECK(hipMemcpyAsync(ctx->d_frames, frames, npx * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
        LZR_CHECK(hipMemcpyAsync(ctx->d_cands, cands, (size_t)ncands * sizeof(LzrCandidate), hipMemcpyHostToDevice,
                                 ctx->stream));
        launch(ctx, ctx->d_frames, nx, ny, nframes, ctx->d_cands, ncands, ctx->d_out, ctx->stream);
        LZR_CHECK(hipMemcpyAsync(results, ctx->d_out, (size_t)ncands * sizeof(LzrFit), hipMemcpyDeviceToHost,
                                 ctx->stream));
        LZR_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (const std::string &e) {
        return fail(ctx, e);
    }
    int accepted = 0;
    for (int i = 0; i < ncands; ++i) accepted += results[i].status == LZR_ACCEPTED;
    return accepted;
}
