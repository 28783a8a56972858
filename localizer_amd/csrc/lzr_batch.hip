/*
 * Host batcher: concurrent per-frame fit calls gathered into one device launch.
 *
 * The reference's analysis controller shares one fitter between hardware_concurrency()
 * worker threads, each calling fit_positions(image, particles) once per frame
 * (source/PALMAnalysis.cpp:96-130, :204-252, the call at :234).  A device launch per
 * frame would run one sort / persistent / straggler / epilogue chain for ~200 fits; the
 * batcher instead lets every caller deposit its frame's fitting windows into an open
 * batch and block, while launcher threads launch whole batches.
 *
 * Only the pixels a fit reads travel: each candidate's window (the reference rule
 * floor(x - h) .. + 2h, source/Localization.cpp:112-124) is packed on the host into
 * pinned memory, uploaded, and placed at its original position in a device frame stack;
 * pixels outside every window are never read.  Candidates keep their own coordinates
 * and frames keep their size, so every window offset -- and with it every absolute
 * coordinate the LM's convergence test sees -- is that of a one-frame call: each record
 * equals the record lzr_fit_batch returns for that frame alone, bit for bit (the fits of
 * one launch do not depend on each other or on the batch composition).
 *
 * Up to `inflight` batches run at once, each launched by its own thread on its own clone
 * of the context (identical fits, separate work memory and streams), while callers fill
 * the next one.  A batch is launched when it is full, when every caller that could still
 * join it has joined, or max_wait_us after its first caller arrived.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/localizer_hip.h"

/* lzr_fit_batch_device launching only the model classes of h_cands, the host copy of the
 * candidates (lzr_fit.hip; internal to the library). */
int lzr_fit_batch_device_hostcands(LzrContext *ctx, const uint16_t *d_frames, int nx, int ny, int nframes,
                                   const LzrCandidate *d_cands, const LzrCandidate *h_cands, int ncands, LzrFit *d_out,
                                   void *stream);

namespace lzr_batch {

/* One packed window: its frame slot, offsets, side and first pixel in the pack. */
struct Window {
    int32_t slot, xoff, yoff, n;
    int64_t pix;
};

__global__ void __launch_bounds__(256) place_windows(const uint16_t *pack, const Window *win, int nwin,
                                                     uint16_t *frames, int nx, int ny) {
    const int w = blockIdx.x;
    if (w >= nwin) return;
    const Window W = win[w];
    const int npx = W.n * W.n;
    uint16_t *dst = frames + ((size_t)W.slot * ny + W.yoff) * nx + W.xoff;
    const uint16_t *src = pack + W.pix;
    for (int k = threadIdx.x; k < npx; k += blockDim.x) {
        const int j = k / W.n, i = k - j * W.n;
        dst[(size_t)j * nx + i] = src[k];
    }
}

enum class State { Free, Open, Inflight, Done };

struct Buf {
    /* pinned host staging */
    uint16_t *h_pix = nullptr;
    Window *h_win = nullptr;
    LzrCandidate *h_cand = nullptr;
    LzrFit *h_out = nullptr;
    /* device */
    uint16_t *d_pix = nullptr;
    Window *d_win = nullptr;
    LzrCandidate *d_cand = nullptr;
    LzrFit *d_out = nullptr;
    uint16_t *d_frames = nullptr;
    /* contents */
    State state = State::Free;
    int nframes = 0, nrois = 0, nwin = 0;
    int64_t npix = 0;
    int writers = 0;  /* callers still copying their inputs in          */
    int members = 0;  /* callers assigned to this batch, not yet returned */
    bool full = false; /* a caller found no room: launch without waiting  */
    bool owned = false; /* a worker is about to launch it                  */
    unsigned long long gen = 0;
    int rc = 0;
    std::string err;
    std::chrono::steady_clock::time_point first;
};

/* One launcher: its own context (a clone: identical fits) and stream. */
struct Slot {
    LzrContext *ctx = nullptr;
    hipStream_t stream = nullptr;
    std::thread thread;
};

}  // namespace lzr_batch

using lzr_batch::Buf;
using lzr_batch::Slot;
using lzr_batch::State;
using lzr_batch::Window;

struct LzrBatcher {
    int device = 0;
    int nx = 0, ny = 0;
    int max_frames = 0, max_rois = 0;
    int64_t max_pix = 0;
    int wait_us = 0;
    int default_h = 0, max_h = 0;
    std::vector<Slot> slots;
    std::vector<Buf> buf;
    int open = -1;  /* the batch accepting callers, -1 when none is free */
    int active = 0; /* callers inside lzr_batcher_fit_frame* */
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv;
    std::string err; /* last error */
    unsigned long long batches = 0, frames = 0;
};

namespace {

thread_local std::string g_batcher_create_error;

bool hip_ok(hipError_t e, std::string &err, const char *what) {
    if (e == hipSuccess) return true;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

/* Window rule of a one-frame call (the checks a lane makes before fitting, in the same
 * order and arithmetic): returns the window side n > 0 and its offsets when the candidate
 * is fitted, 0 when it is rejected before fitting (invalid request, window outside). */
int window_of(const LzrBatcher *b, const LzrCandidate &c, int &xoff, int &yoff) {
    const int h = c.half_width > 0 ? c.half_width : b->default_h;
    if (h < 1 || h > b->max_h || c.frame != 0) return 0;
    const int n = 2 * h + 1;
    const double fx = c.x - (double)h, fy = c.y - (double)h;
    const bool inside = (fx >= 0.0) && (fy >= 0.0) && (fx < (double)(b->nx - n + 1)) && (fy < (double)(b->ny - n + 1));
    if (!inside) return 0;
    xoff = (int)fx;
    yoff = (int)fy;
    return n;
}

template <typename Pixel>
bool pixel_ok(Pixel v, uint16_t &out);

template <>
bool pixel_ok<uint16_t>(uint16_t v, uint16_t &out) {
    out = v;
    return true;
}

template <>
bool pixel_ok<double>(double v, uint16_t &out) {
    if (!(v >= 0.0 && v <= 65535.0) || v != std::floor(v)) return false;
    out = (uint16_t)v;
    return true;
}

void release(Buf &b) {
    (void)hipHostFree(b.h_pix);
    (void)hipHostFree(b.h_win);
    (void)hipHostFree(b.h_cand);
    (void)hipHostFree(b.h_out);
    (void)hipFree(b.d_pix);
    (void)hipFree(b.d_win);
    (void)hipFree(b.d_cand);
    (void)hipFree(b.d_out);
    (void)hipFree(b.d_frames);
}

/* Uploads one batch, fits it and copies the records back (on a launcher thread). */
void run_batch(LzrBatcher *B, Slot &slot, Buf &b) {
    std::string err;
    bool ok = hip_ok(hipSetDevice(B->device), err, "hipSetDevice");
    hipStream_t s = slot.stream;
    if (ok && b.nwin > 0) {
        ok = hip_ok(hipMemcpyAsync(b.d_pix, b.h_pix, (size_t)b.npix * sizeof(uint16_t), hipMemcpyHostToDevice, s),
                    err, "upload windows") &&
             hip_ok(hipMemcpyAsync(b.d_win, b.h_win, (size_t)b.nwin * sizeof(Window), hipMemcpyHostToDevice, s), err,
                    "upload window table");
        if (ok) {
            hipLaunchKernelGGL(lzr_batch::place_windows, dim3(b.nwin), dim3(256), 0, s, b.d_pix, b.d_win, b.nwin,
                               b.d_frames, B->nx, B->ny);
            ok = hip_ok(hipGetLastError(), err, "place_windows");
        }
    }
    if (ok)
        ok = hip_ok(hipMemcpyAsync(b.d_cand, b.h_cand, (size_t)b.nrois * sizeof(LzrCandidate), hipMemcpyHostToDevice,
                                   s),
                    err, "upload candidates");
    if (ok && lzr_fit_batch_device_hostcands(slot.ctx, b.d_frames, B->nx, B->ny, b.nframes, b.d_cand, b.h_cand, b.nrois,
                                             b.d_out, s) != 0) {
        char buf[512];
        lzr_last_error(slot.ctx, buf, (int)sizeof buf);
        err = std::string("fit: ") + buf;
        ok = false;
    }
    if (ok)
        ok = hip_ok(hipMemcpyAsync(b.h_out, b.d_out, (size_t)b.nrois * sizeof(LzrFit), hipMemcpyDeviceToHost, s), err,
                    "download records");
    if (ok) ok = hip_ok(hipStreamSynchronize(s), err, "hipStreamSynchronize");
    b.rc = ok ? 0 : -1;
    b.err = err;
}

int members_elsewhere(const LzrBatcher *B, int k) {
    int m = 0;
    for (int i = 0; i < (int)B->buf.size(); ++i)
        if (i != k) m += B->buf[i].members;
    return m;
}

/* A launcher: takes the open batch once it is full, once every caller that could still
 * join has joined, or max_wait_us after its first caller; launches it; repeats. */
void launcher_loop(LzrBatcher *B, int si) {
    Slot &slot = B->slots[si];
    std::unique_lock<std::mutex> lock(B->mu);
    for (;;) {
        B->cv.wait(lock, [&] {
            return B->stop || (B->open >= 0 && B->buf[B->open].members > 0 && !B->buf[B->open].owned);
        });
        if (B->open < 0 || B->buf[B->open].members == 0 || B->buf[B->open].owned) {
            if (B->stop) return;
            continue;
        }
        const int k = B->open;
        Buf &b = B->buf[k];
        b.owned = true;
        const auto deadline = b.first + std::chrono::microseconds(B->wait_us);
        B->cv.wait_until(lock, deadline, [&] {
            return B->stop || b.full || b.members >= B->active - members_elsewhere(B, k);
        });
        b.state = State::Inflight;
        b.owned = false;
        /* the next open batch: a free one, if any (else the first batch to drain) */
        B->open = -1;
        for (int i = 0; i < (int)B->buf.size(); ++i)
            if (B->buf[i].state == State::Free) {
                B->buf[i].state = State::Open;
                B->open = i;
                break;
            }
        B->cv.notify_all();
        B->cv.wait(lock, [&] { return b.writers == 0; });
        lock.unlock();
        run_batch(B, slot, b);
        lock.lock();
        B->batches++;
        B->frames += (unsigned long long)b.nframes;
        b.state = State::Done;
        B->cv.notify_all();
    }
}

template <typename Pixel>
int fit_frame(LzrBatcher *B, const Pixel *frame, int nx, int ny, const LzrCandidate *cands, int ncands,
              LzrFit *out) {
    if (!B) return -1;
    if (nx != B->nx || ny != B->ny) {
        std::lock_guard<std::mutex> g(B->mu);
        B->err = "frame size differs from the batcher's";
        return -1;
    }
    if (ncands < 0 || (ncands > 0 && (!frame || !cands || !out))) {
        std::lock_guard<std::mutex> g(B->mu);
        B->err = "invalid frame or candidate arguments";
        return -1;
    }
    if (ncands == 0) return 0;
    /* the windows this frame contributes */
    int nwin = 0;
    int64_t npix = 0;
    std::vector<int> wn((size_t)ncands), wx((size_t)ncands), wy((size_t)ncands);
    for (int i = 0; i < ncands; ++i) {
        wn[i] = window_of(B, cands[i], wx[i], wy[i]);
        if (wn[i] > 0) {
            ++nwin;
            npix += (int64_t)wn[i] * wn[i];
        }
    }
    if (ncands > B->max_rois || npix > B->max_pix) {
        std::lock_guard<std::mutex> g(B->mu);
        B->err = "frame exceeds the batcher's capacity";
        return -1;
    }
    std::unique_lock<std::mutex> lock(B->mu);
    B->active++;
    B->cv.notify_all();
    /* reserve room in the open batch */
    int bi = -1, slot = 0, r0 = 0, w0 = 0;
    int64_t p0 = 0;
    unsigned long long gen = 0;
    for (;;) {
        if (B->open >= 0) {
            Buf &b = B->buf[B->open];
            const bool fits =
                b.nframes < B->max_frames && b.nrois + ncands <= B->max_rois && b.npix + npix <= B->max_pix;
            if (fits && !b.full) {
                bi = B->open;
                slot = b.nframes++;
                r0 = b.nrois;
                b.nrois += ncands;
                w0 = b.nwin;
                b.nwin += nwin;
                p0 = b.npix;
                b.npix += npix;
                if (b.members == 0) b.first = std::chrono::steady_clock::now();
                b.members++;
                b.writers++;
                gen = b.gen;
                break;
            }
            b.full = true; /* no room: its launcher goes without waiting */
        }
        B->cv.notify_all();
        B->cv.wait(lock);
    }
    B->cv.notify_all();
    lock.unlock();
    Buf &b = B->buf[bi];
    /* copy the inputs in (outside the lock; the launcher waits for writers == 0) */
    bool pixels_ok = true;
    int w = w0;
    int64_t p = p0;
    for (int i = 0; i < ncands; ++i) {
        LzrCandidate c = cands[i];
        c.frame = c.frame == 0 ? slot : -1; /* a one-frame call knows frame 0 only */
        b.h_cand[r0 + i] = c;
        const int n = wn[i];
        if (n == 0) continue;
        Window &W = b.h_win[w++];
        W.slot = slot;
        W.xoff = wx[i];
        W.yoff = wy[i];
        W.n = n;
        W.pix = p;
        for (int j = 0; j < n; ++j) {
            const Pixel *src = frame + (size_t)(wy[i] + j) * nx + wx[i];
            uint16_t *dst = b.h_pix + p + (int64_t)j * n;
            for (int k = 0; k < n; ++k) pixels_ok &= pixel_ok<Pixel>(src[k], dst[k]);
        }
        p += (int64_t)n * n;
    }
    lock.lock();
    b.writers--;
    B->cv.notify_all();
    B->cv.wait(lock, [&] { return b.state == State::Done && b.gen == gen; });
    /* a frame with unusable pixels fails alone: the other frames' fits do not read its windows */
    const int rc = pixels_ok ? b.rc : -1;
    if (!pixels_ok)
        B->err = "frame pixels inside a fitting window must be integer counts in [0, 65535]";
    else if (rc != 0)
        B->err = b.err;
    lock.unlock();
    int accepted = 0;
    if (rc == 0) {
        std::memcpy(out, b.h_out + r0, (size_t)ncands * sizeof(LzrFit));
        for (int i = 0; i < ncands; ++i) accepted += out[i].status == LZR_ACCEPTED;
    }
    lock.lock();
    B->active--;
    if (--b.members == 0) {
        /* the last caller of this batch frees it for the next one */
        b.nframes = b.nrois = b.nwin = 0;
        b.npix = 0;
        b.full = false;
        b.rc = 0;
        b.err.clear();
        b.gen++;
        b.state = State::Free;
        if (B->open < 0) {
            b.state = State::Open;
            B->open = bi;
        }
    }
    B->cv.notify_all();
    return rc == 0 ? accepted : -1;
}

void destroy(LzrBatcher *B) {
    {
        std::lock_guard<std::mutex> g(B->mu);
        B->stop = true;
    }
    B->cv.notify_all();
    for (Slot &s : B->slots)
        if (s.thread.joinable()) s.thread.join();
    (void)hipSetDevice(B->device);
    for (Buf &b : B->buf) release(b);
    for (Slot &s : B->slots) {
        if (s.stream) (void)hipStreamDestroy(s.stream);
        if (s.ctx) lzr_destroy(s.ctx);
    }
    delete B;
}

}  // namespace

LZR_EXPORT int lzr_batcher_create(LzrBatcher **out, LzrContext *ctx, int nx, int ny, int max_frames, int max_rois,
                                  int max_wait_us, int inflight) {
    if (!out) {
        g_batcher_create_error = "lzr_batcher_create: null output pointer";
        return -1;
    }
    *out = nullptr;
    if (!ctx || nx < 1 || ny < 1 || max_frames < 1 || max_rois < 1 || max_wait_us < 0 || inflight < 1 ||
        inflight > 64) {
        g_batcher_create_error = "lzr_batcher_create: invalid arguments";
        return -1;
    }
    double dh = 0, mh = 0;
    if (lzr_get_option(ctx, LZR_OPT_DEFAULT_HALF_WIDTH, &dh) != 0 || lzr_get_option(ctx, LZR_OPT_MAX_HALF_WIDTH, &mh) != 0) {
        g_batcher_create_error = "lzr_batcher_create: cannot read the context's window options";
        return -1;
    }
    LzrBatcher *B = new LzrBatcher();
    B->nx = nx;
    B->ny = ny;
    B->max_frames = max_frames;
    B->max_rois = max_rois;
    B->default_h = (int)dh;
    B->max_h = (int)mh;
    const int64_t n = 2 * (int64_t)B->default_h + 1;
    B->max_pix = (int64_t)max_rois * n * n; /* room for every candidate at the default window */
    B->wait_us = max_wait_us;
    B->slots.resize((size_t)inflight);
    B->buf.resize((size_t)inflight + 1);
    std::string err;
    bool ok = hip_ok(hipGetDevice(&B->device), err, "hipGetDevice");
    for (Slot &s : B->slots) {
        if (!ok) break;
        if (lzr_clone(ctx, &s.ctx) != 0) {
            char buf[512];
            lzr_last_error(nullptr, buf, (int)sizeof buf);
            err = buf;
            ok = false;
            break;
        }
        ok = hip_ok(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), err, "hipStreamCreate");
    }
    const size_t frame_px = (size_t)nx * ny;
    for (Buf &b : B->buf) {
        if (!ok) break;
        ok = hip_ok(hipHostMalloc((void **)&b.h_pix, (size_t)B->max_pix * sizeof(uint16_t)), err, "hipHostMalloc") &&
             hip_ok(hipHostMalloc((void **)&b.h_win, (size_t)max_rois * sizeof(Window)), err, "hipHostMalloc") &&
             hip_ok(hipHostMalloc((void **)&b.h_cand, (size_t)max_rois * sizeof(LzrCandidate)), err,
                    "hipHostMalloc") &&
             hip_ok(hipHostMalloc((void **)&b.h_out, (size_t)max_rois * sizeof(LzrFit)), err, "hipHostMalloc") &&
             hip_ok(hipMalloc((void **)&b.d_pix, (size_t)B->max_pix * sizeof(uint16_t)), err, "hipMalloc") &&
             hip_ok(hipMalloc((void **)&b.d_win, (size_t)max_rois * sizeof(Window)), err, "hipMalloc") &&
             hip_ok(hipMalloc((void **)&b.d_cand, (size_t)max_rois * sizeof(LzrCandidate)), err, "hipMalloc") &&
             hip_ok(hipMalloc((void **)&b.d_out, (size_t)max_rois * sizeof(LzrFit)), err, "hipMalloc") &&
             hip_ok(hipMalloc((void **)&b.d_frames, frame_px * (size_t)max_frames * sizeof(uint16_t)), err,
                    "hipMalloc");
    }
    if (!ok) {
        destroy(B);
        g_batcher_create_error = "lzr_batcher_create: " + err;
        return -1;
    }
    B->buf[0].state = State::Open;
    B->open = 0;
    for (int i = 0; i < inflight; ++i) B->slots[i].thread = std::thread(launcher_loop, B, i);
    *out = B;
    return 0;
}

LZR_EXPORT int lzr_batcher_fit_frame(LzrBatcher *batcher, const uint16_t *frame, int nx, int ny,
                                     const LzrCandidate *candidates, int ncandidates, LzrFit *results) {
    return fit_frame<uint16_t>(batcher, frame, nx, ny, candidates, ncandidates, results);
}

LZR_EXPORT int lzr_batcher_fit_frame_f64(LzrBatcher *batcher, const double *frame, int nx, int ny,
                                         const LzrCandidate *candidates, int ncandidates, LzrFit *results) {
    return fit_frame<double>(batcher, frame, nx, ny, candidates, ncandidates, results);
}

LZR_EXPORT int lzr_batcher_stats(LzrBatcher *batcher, unsigned long long *batches, unsigned long long *frames) {
    if (!batcher) return -1;
    std::lock_guard<std::mutex> g(batcher->mu);
    if (batches) *batches = batcher->batches;
    if (frames) *frames = batcher->frames;
    return 0;
}

LZR_EXPORT int lzr_batcher_last_error(const LzrBatcher *batcher, char *buf, int len) {
    std::string e;
    if (batcher) {
        LzrBatcher *B = const_cast<LzrBatcher *>(batcher);
        std::lock_guard<std::mutex> g(B->mu);
        e = B->err;
    } else {
        e = g_batcher_create_error;
    }
    if (buf && len > 0) {
        const int n = (int)std::min(e.size(), (size_t)(len - 1));
        std::memcpy(buf, e.data(), (size_t)n);
        buf[n] = 0;
    }
    return (int)e.size();
}

LZR_EXPORT int lzr_batcher_destroy(LzrBatcher *batcher) {
    if (!batcher) return -1;
    destroy(batcher);
    return 0;
}
