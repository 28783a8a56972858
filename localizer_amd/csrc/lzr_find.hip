/*
 * lzr_find.hip -- candidate generation on gfx950: GLRT segmentation and 8-connected
 * particle finding, and the C ABI lzr_finder_* / lzr_find_candidates_device of
 * include/localizer_hip.h.
 *
 *   ThresholdImage_GLRT_FFT::do_thresholding   (source/Segmentation.cpp:241-430)
 *   ParticleFinder_adjacent8::findPositions     (source/ParticleFinding.cpp:283-427)
 *
 * GLRT: for every tested pixel the window sums S1 = sum x, S2 = sum x^2 and the
 * correlation with the Gaussian, G = sum gc x = c * sum_j E(j) sum_i E(i) x - mean(g) S1
 * (g = c E(i) E(j) is separable; gc = g - mean(g)), all in FP64 (S1, S2 are exact
 * integers); then the reference's formula order m = S1/N, V = S2 - m*m*N,
 * alpha = G/sum(gc^2), T = alpha*alpha/V*sum(gc^2), segmented iff thr < T < 1 with
 * thr = 1 - exp(-PFA/N), over x in [h+1, X-h), y in [h+1, Y-h), X, Y rounded down to
 * even.  One 256-thread block per 64 x 64 tile: the tile and its halo are staged in LDS,
 * each thread walks a column keeping the window's row sums in a register ring.
 * Segmented pixels are appended to a per-frame list.
 *
 * Particles: one 1024-thread block per frame sorts the frame's segmented pixels (bitonic,
 * LDS) -- raster order -- and joins 8-neighbours with a lock-free union-find whose roots
 * are the smallest index of each component, i.e. the component's first pixel in the
 * reference's scan order (y outer, x inner), so numbering the roots in sorted order
 * reproduces the reference's particle order.  x, y = mean of the component's pixel
 * coordinates, intensity = max pixel - background, background = mean of all
 * unsegmented pixels of the frame (from the frame sum and the segmented pixels' sum).
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

#define FIND_CHECK(call)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) throw std::string(#call ": ") + hipGetErrorString(e_);        \
    } while (0)

namespace lzr_find {

constexpr int kMaxWin = 15;         /* largest GLRT window side accepted */

struct GlrtParams {
    int ws, h;
    double n;           /* ws * ws */
    double c;           /* 1 / (1.77245385 w) */
    double gmean;       /* mean of g over the window */
    double ssg;         /* sum of gc^2 */
    double thr;         /* 1 - exp(-PFA / N) */
    double e[kMaxWin];  /* E(d) = exp(-d^2 / (2 w^2)), d = -h..h */
};

struct FrameInfo {
    unsigned long long frame_sum; /* sum of every pixel of the frame */
    unsigned fg_count;            /* segmented pixels found (may exceed the capacity) */
    unsigned particles;
};

/* The test in the reference's operation order, without contractions into FMAs. */
__device__ inline bool glrt_test(const GlrtParams &P, double S1, double S2, double Sg) {
#pragma clang fp contract(off)
    const double G = P.c * Sg - P.gmean * S1;
    const double m = S1 / P.n;
    const double V = S2 - m * m * P.n;
    const double alpha = G / P.ssg;
    const double T = alpha * alpha / V * P.ssg;
    return T < 1.0 && T > P.thr;
}

/* One 256-thread block per 64 x 64 tile of one frame; the tile and its halo are staged
 * in LDS as floats (exact for uint16).  Thread (column c, row group q) walks down its
 * column over 16 output rows: the row sums of each input row over the window's columns
 * (sum x, sum x^2, sum E(d) x) go into a ring of WS registers per sum (WS is a template
 * parameter, so the ring is unrolled); S1 and S2 (integers below 2^53, exact) are running
 * sums over the ring, Sg = sum_d E(d) ring_g[d] is added per output pixel.  No LDS traffic beyond one read per
 * (input row, window column). */
constexpr int kGTX = 64, kGTY = 64, kGRows = kGTY / 4;

template <int WS>
__global__ void __launch_bounds__(256) glrt_kernel(const uint16_t *frames, int nx, int ny, GlrtParams P, int cap,
                                                   int *fg_list, FrameInfo *info) {
    constexpr int H = WS / 2;
    constexpr int IX = kGTX + 2 * H, IY = kGTY + 2 * H;
    __shared__ float vin[IY][IX];
    const int f = blockIdx.z;
    const int x0 = blockIdx.x * kGTX, y0 = blockIdx.y * kGTY;
    const int X = nx - (nx % 2), Y = ny - (ny % 2);
    const uint16_t *src = frames + (size_t)f * nx * ny;
    const int tid = threadIdx.x, lane = tid & 63;
    /* stage the tile with its halo; pixels outside the frame read as 0 (never tested) */
    unsigned long long own = 0;
    for (int k = tid; k < IY * IX; k += 256) {
        const int r = k / IX, c = k - r * IX;
        const int gx = x0 - H + c, gy = y0 - H + r;
        uint16_t v = 0;
        if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) v = src[(size_t)gy * nx + gx];
        vin[r][c] = (float)v;
        /* the tile's own pixels feed the frame sum (each pixel once) */
        if (r >= H && r < H + kGTY && c >= H && c < H + kGTX && gx < nx && gy < ny) own += v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) own += __shfl_xor(own, off, 64);
    if (lane == 0 && own) atomicAdd(&info[f].frame_sum, own);
    __syncthreads();
    double e[WS];
#pragma unroll
    for (int d = 0; d < WS; ++d) e[d] = P.e[d];
    const int c = lane, q = tid >> 6;
    const int rb = q * kGRows; /* first output row of this thread (tile-relative) */
    const unsigned long long below = (1ull << lane) - 1ull;
    double r1[WS], r2[WS], rg[WS];
    auto row_sums = [&](int ir, double &a, double &b, double &g) {
        a = 0.0;
        b = 0.0;
        g = 0.0;
#pragma unroll
        for (int d = 0; d < WS; ++d) {
            const double v = (double)vin[ir][c + d];
            a += v;
            b += v * v;
            g += e[d] * v;
        }
    };
    /* S1 and S2 are sums of integers below 2^53, exact in FP64 in any order: running sums
     * over the window's rows (add the entering row, subtract the leaving one) */
    double S1r = 0.0, S2r = 0.0;
#pragma unroll
    for (int d = 0; d < WS - 1; ++d) {
        row_sums(rb + d, r1[d], r2[d], rg[d]);
        S1r += r1[d];
        S2r += r2[d];
    }
#pragma unroll
    for (int k = 0; k < kGRows; ++k) {
        const int slot = (k + WS - 1) % WS;
        row_sums(rb + k + WS - 1, r1[slot], r2[slot], rg[slot]);
        S1r += r1[slot];
        S2r += r2[slot];
        const int x = x0 + c, y = y0 + rb + k;
        bool seg = false;
        if (x >= H + 1 && x < X - H && y >= H + 1 && y < Y - H) {
            double Sg = 0.0;
#pragma unroll
            for (int d = 0; d < WS; ++d) Sg += e[d] * rg[(k + d) % WS]; /* input row rb + k + d */
            seg = glrt_test(P, S1r, S2r, Sg);
        }
        S1r -= r1[k % WS]; /* input row rb + k leaves the window */
        S2r -= r2[k % WS];
        /* wave-aggregated append to the frame's list (the wave walks one row at a time) */
        const unsigned long long m = __ballot(seg);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&info[f].fg_count, (unsigned)__popcll(m));
            base = __shfl(base, leader, 64);
            if (seg) {
                const unsigned slot2 = base + (unsigned)__popcll(m & below);
                if (slot2 < (unsigned)cap) fg_list[(size_t)f * cap + slot2] = y * nx + x;
            }
        }
    }
}

/* ---- per-frame components ------------------------------------------------------------ */

__device__ inline int lds_load(int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

__device__ inline int uf_find(int *lab, int a) {
    int p = lds_load(&lab[a]);
    while (p != a) {
        a = p;
        p = lds_load(&lab[a]);
    }
    return a;
}

/* joins the components of a and b; the root is always the smaller index */
__device__ inline void uf_union(int *lab, int a, int b) {
    for (;;) {
        a = uf_find(lab, a);
        b = uf_find(lab, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(&lab[b], a);
        if (old == b) return;
        b = old;
    }
}

/* position of key q in sorted keys[0, n), or -1 */
__device__ inline int bsearch_key(const int *keys, int n, int q) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = keys[mid];
        if (v == q) return mid;
        if (v < q)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

struct Particle {
    long long sx, sy; /* coordinate sums */
    int count, maxv;
};

constexpr int kCCThreads = 1024;

/* One block per frame.  keys/labels live in dynamic LDS (2 * cap ints); cap is a power
 * of two.  Writes the frame's particles to out[offset_f ...] (offset_f = scan of the
 * segmented counts) and their number to info[f].particles; info[f].fg_count > cap marks
 * an overflow (no particles written). */
__global__ void __launch_bounds__(kCCThreads) cc_kernel(const uint16_t *frames, int nx, int ny, int cap,
                                                        const int *fg_list, FrameInfo *info,
                                                        const unsigned *fg_offset, Particle *scratch,
                                                        LzrCandidate *out, int frame0) {
    extern __shared__ int lds[];
    int *keys = lds;
    int *lab = lds + cap;
    __shared__ unsigned long long fg_sum;
    __shared__ unsigned roots_total;
    __shared__ unsigned warp_counts[kCCThreads / 64];
    const int f = blockIdx.x, tid = threadIdx.x;
    const unsigned nfg = info[f].fg_count;
    if (nfg > (unsigned)cap) { /* overflow: reported by the host wrapper */
        if (tid == 0) info[f].particles = 0;
        return;
    }
    const int n = (int)nfg;
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    const uint16_t *src = frames + (size_t)f * nx * ny;
    if (tid == 0) fg_sum = 0ull;
    for (int i = tid; i < np2; i += kCCThreads) keys[i] = i < n ? fg_list[(size_t)f * cap + i] : 0x7fffffff;
    __syncthreads();
    /* bitonic sort of keys[0, np2) */
    for (int k = 2; k <= np2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < np2; i += kCCThreads) {
                const int p = i ^ j;
                if (p > i) {
                    const int a = keys[i], b = keys[p];
                    const bool asc = (i & k) == 0;
                    if ((a > b) == asc) {
                        keys[i] = b;
                        keys[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = tid; i < n; i += kCCThreads) lab[i] = i;
    __syncthreads();
    /* union with the neighbours earlier in raster order: W, NW, N, NE */
    unsigned long long vs = 0;
    for (int i = tid; i < n; i += kCCThreads) {
        const int p = keys[i];
        const int x = p % nx, y = p / nx;
        vs += src[p];
        const int nb[4][2] = {{x - 1, y}, {x - 1, y - 1}, {x, y - 1}, {x + 1, y - 1}};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ux = nb[q][0], uy = nb[q][1];
            if (ux < 0 || uy < 0 || ux >= nx) continue;
            const int j = bsearch_key(keys, i, uy * nx + ux);
            if (j >= 0) uf_union(lab, i, j);
        }
    }
    atomicAdd(&fg_sum, vs);
    __syncthreads();
    for (int i = tid; i < n; i += kCCThreads) lab[i] = uf_find(lab, i);
    __syncthreads();
    /* number the roots in sorted (raster) order: per-thread contiguous chunks */
    const int chunk = (n + kCCThreads - 1) / kCCThreads;
    const int c0 = tid * chunk, c1 = min(n, c0 + chunk);
    unsigned mine = 0;
    for (int i = c0; i < c1; ++i) mine += lab[i] == i;
    /* block exclusive scan of `mine` */
    const int lane = tid & 63, wv = tid >> 6;
    unsigned incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) warp_counts[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        unsigned acc = 0;
        for (int w = 0; w < kCCThreads / 64; ++w) {
            const unsigned t = warp_counts[w];
            warp_counts[w] = acc;
            acc += t;
        }
        roots_total = acc;
    }
    __syncthreads();
    unsigned id = warp_counts[wv] + incl - mine;
    Particle *P = scratch + fg_offset[f];
    /* roots get their particle number in lab (as -(id + 1)), their record is zeroed */
    for (int i = c0; i < c1; ++i)
        if (lab[i] == i) {
            P[id].sx = 0;
            P[id].sy = 0;
            P[id].count = 0;
            P[id].maxv = 0;
            lab[i] = -(int)(id + 1);
            ++id;
        }
    __syncthreads();
    for (int i = tid; i < n; i += kCCThreads) {
        int r = lab[i];
        if (r >= 0) r = lab[r]; /* member: its root holds -(id + 1) */
        const int pid = -r - 1;
        const int p = keys[i];
        atomicAdd((unsigned long long *)&P[pid].sx, (unsigned long long)(p % nx));
        atomicAdd((unsigned long long *)&P[pid].sy, (unsigned long long)(p / nx));
        atomicAdd(&P[pid].count, 1);
        atomicMax(&P[pid].maxv, (int)src[p]);
    }
    __syncthreads();
    const unsigned npart = roots_total;
    const double npx = (double)nx * (double)ny;
    const double bg = (double)(info[f].frame_sum - fg_sum) / (npx - (double)n);
    for (unsigned k = tid; k < npart; k += kCCThreads) {
        const Particle q = P[k];
        LzrCandidate c;
        c.x = (double)q.sx / (double)q.count;
        c.y = (double)q.sy / (double)q.count;
        c.intensity = (double)q.maxv - bg;
        c.background = bg;
        c.frame = frame0 + f;
        c.half_width = 0;
        c.model = -1;
        out[fg_offset[f] + k] = c;
    }
    if (tid == 0) info[f].particles = npart;
}

/* exclusive scan of per-frame counts (one block): mode 0 scans fg_count, mode 1 scans
 * particles; total -> *total (mode 1: -1 if any frame overflowed or total > capacity) */
__global__ void __launch_bounds__(1024) scan_frames_kernel(FrameInfo *info, int nframes, int mode, int cap,
                                                           unsigned *offset, int *total, int capacity) {
    __shared__ unsigned long long part[1024];
    __shared__ int overflow;
    const int tid = threadIdx.x;
    if (tid == 0) overflow = 0;
    __syncthreads();
    const int chunk = (nframes + 1023) / 1024;
    const int f0 = tid * chunk, f1 = min(nframes, f0 + chunk);
    unsigned long long s = 0;
    for (int f = f0; f < f1; ++f) {
        const unsigned v = mode == 0 ? min(info[f].fg_count, (unsigned)cap) : info[f].particles;
        if (mode == 1 && info[f].fg_count > (unsigned)cap) overflow = 1;
        s += v;
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned long long t = tid >= off ? part[tid - off] : 0ull;
        __syncthreads();
        part[tid] += t;
        __syncthreads();
    }
    unsigned long long acc = part[tid] - s;
    for (int f = f0; f < f1; ++f) {
        offset[f] = (unsigned)acc;
        acc += mode == 0 ? min(info[f].fg_count, (unsigned)cap) : info[f].particles;
    }
    if (tid == 1023 && total) {
        const unsigned long long t = part[1023];
        *total = (mode == 1 && (overflow || t > (unsigned long long)capacity)) ? -1 : (int)t;
    }
}

/* gathers every frame's particles (scratch, frame-major by segmented-pixel offsets) into
 * the caller's array in frame order */
__global__ void compact_kernel(const LzrCandidate *staged, const unsigned *fg_offset, const unsigned *out_offset,
                               const FrameInfo *info, int nframes, const int *total, LzrCandidate *out) {
    if (*total < 0) return;
    const int f = blockIdx.y;
    if (f >= nframes) return;
    const unsigned np = info[f].particles;
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < np; k += gridDim.x * blockDim.x)
        out[out_offset[f] + k] = staged[fg_offset[f] + k];
}

__global__ void reset_info(FrameInfo *info, int nframes) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nframes) {
        info[f].frame_sum = 0ull;
        info[f].fg_count = 0u;
        info[f].particles = 0u;
    }
}

}  // namespace lzr_find

using namespace lzr_find;

struct LzrFinder {
    int device = 0;
    double width = 1.0, pfa = 20.0;
    int cap = 16384;
    GlrtParams P{};
    std::string err;
    std::mutex mu;
    size_t frames_cap = 0;
    int *d_fg = nullptr;
    FrameInfo *d_info = nullptr;
    unsigned *d_fg_off = nullptr, *d_out_off = nullptr;
    Particle *d_scratch = nullptr;
    LzrCandidate *d_staged = nullptr;
    int *d_fg_total = nullptr;
    LzrStreamOrder order; /* calls on different streams share the buffers above */
};

static thread_local std::string g_finder_error;

template <typename T>
static void falloc(T *&p, size_t n) {
    if (p) FIND_CHECK(hipFree(p));
    p = nullptr;
    FIND_CHECK(hipMalloc((void **)&p, std::max(n, (size_t)1) * sizeof(T)));
}

LZR_EXPORT int lzr_finder_create(LzrFinder **out, double psf_width, double pfa, int max_segmented_per_frame) {
    if (!out) {
        g_finder_error = "lzr_finder_create: null output pointer";
        return -1;
    }
    *out = nullptr;
    if (!(psf_width > 0.0) || !(pfa > 0.0)) {
        g_finder_error = "lzr_finder_create: PSF width and PFA must be positive";
        return -1;
    }
    int ws = (int)std::ceil(4.0 * psf_width);
    if (ws % 2 == 0) ws += 1;
    if (ws > kMaxWin) {
        g_finder_error = "lzr_finder_create: GLRT window (ceil(4 w), odd) larger than 15";
        return -1;
    }
    int cap = max_segmented_per_frame > 0 ? max_segmented_per_frame : 16384;
    if (cap > 16384) {
        g_finder_error = "lzr_finder_create: at most 16384 segmented pixels per frame";
        return -1;
    }
    int p2 = 1;
    while (p2 < cap) p2 <<= 1;
    LzrFinder *f = new LzrFinder();
    f->width = psf_width;
    f->pfa = pfa;
    f->cap = p2;
    /* kernel constants, the reference's formulas (source/Segmentation.cpp:283-300) */
    GlrtParams &P = f->P;
    P.ws = ws;
    P.h = ws / 2;
    P.n = (double)ws * (double)ws;
    P.c = 1.0 / (1.77245385 * psf_width);
    double sum = 0.0;
    std::vector<double> g((size_t)ws * ws);
    for (int j = -P.h; j <= P.h; ++j)
        for (int i = -P.h; i <= P.h; ++i) {
            const double v = 1.0 / (1.77245385 * psf_width) *
                             std::exp(-1.0 / (2.0 * psf_width * psf_width) * ((double)i * i + (double)j * j));
            g[(size_t)(i + P.h) + (size_t)ws * (j + P.h)] = v;
            sum += v;
        }
    sum /= P.n;
    double ssg = 0.0;
    for (double &v : g) {
        v -= sum;
        ssg += v * v;
    }
    P.gmean = sum;
    P.ssg = ssg;
    P.thr = 1.0 - std::exp(-pfa / P.n);
    for (int d = 0; d < ws; ++d) P.e[d] = std::exp(-((double)(d - P.h) * (d - P.h)) / (2.0 * psf_width * psf_width));
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::string("no HIP device available");
        FIND_CHECK(hipGetDevice(&f->device));
        /* the largest dynamic LDS any finder uses (the attribute is per kernel, not per finder) */
        FIND_CHECK(hipFuncSetAttribute((const void *)cc_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(2 * 16384 * sizeof(int))));
    } catch (const std::string &e) {
        g_finder_error = "lzr_finder_create: " + e;
        delete f;
        return -1;
    }
    *out = f;
    return 0;
}

LZR_EXPORT int lzr_finder_destroy(LzrFinder *f) {
    if (!f) return -1;
    void *ptrs[] = {f->d_fg, f->d_info, f->d_fg_off, f->d_out_off, f->d_scratch, f->d_staged, f->d_fg_total};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    f->order.destroy();
    delete f;
    return 0;
}

LZR_EXPORT int lzr_finder_last_error(const LzrFinder *f, char *buf, int len) {
    const std::string &e = f ? f->err : g_finder_error;
    if (buf && len > 0) {
        const int n = std::min((int)e.size(), len - 1);
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}

LZR_EXPORT int lzr_find_candidates_device(LzrFinder *f, const uint16_t *d_frames, int nx, int ny, int nframes,
                                          LzrCandidate *d_candidates, int max_candidates, int *d_count,
                                          void *stream) {
    if (!f) return -1;
    std::lock_guard<std::mutex> lock(f->mu);
    if (nx < 1 || ny < 1 || nframes < 0 || max_candidates < 0) {
        f->err = "invalid frame or candidate counts";
        return -1;
    }
    if (nframes > 65535) {
        f->err = "at most 65535 frames per call";
        return -1;
    }
    if ((size_t)nx * ny > (size_t)0x7fffffff) {
        f->err = "frame too large";
        return -1;
    }
    if (!d_count || (nframes > 0 && (!d_frames || !d_candidates))) {
        f->err = "null device pointer";
        return -1;
    }
    try {
        FIND_CHECK(hipSetDevice(f->device));
        hipStream_t s = (hipStream_t)stream;
        const int nf = std::max(nframes, 1);
        if ((size_t)nf > f->frames_cap) {
            falloc(f->d_fg, (size_t)nf * f->cap);
            falloc(f->d_info, (size_t)nf);
            falloc(f->d_fg_off, (size_t)nf);
            falloc(f->d_out_off, (size_t)nf);
            falloc(f->d_scratch, (size_t)nf * f->cap);
            falloc(f->d_staged, (size_t)nf * f->cap);
            falloc(f->d_fg_total, 1);
            f->frames_cap = (size_t)nf;
        }
        FIND_CHECK(f->order.begin(s));
        hipLaunchKernelGGL(reset_info, dim3((nf + 255) / 256), dim3(256), 0, s, f->d_info, nf);
        if (nframes > 0) {
            const dim3 grid((nx + kGTX - 1) / kGTX, (ny + kGTY - 1) / kGTY, nframes);
            switch (f->P.ws) {
#define LZR_GLRT_CASE(W)                                                                                     \
    case W:                                                                                                  \
        hipLaunchKernelGGL(glrt_kernel<W>, grid, dim3(256), 0, s, d_frames, nx, ny, f->P, f->cap, f->d_fg, f->d_info); \
        break;
                LZR_GLRT_CASE(1)
                LZR_GLRT_CASE(3)
                LZR_GLRT_CASE(5)
                LZR_GLRT_CASE(7)
                LZR_GLRT_CASE(9)
                LZR_GLRT_CASE(11)
                LZR_GLRT_CASE(13)
                LZR_GLRT_CASE(15)
#undef LZR_GLRT_CASE
            default:
                throw std::string("GLRT window side must be odd and at most 15");
            }
            hipLaunchKernelGGL(scan_frames_kernel, dim3(1), dim3(1024), 0, s, f->d_info, nframes, 0, f->cap,
                               f->d_fg_off, f->d_fg_total, 0);
            hipLaunchKernelGGL(cc_kernel, dim3(nframes), dim3(kCCThreads), 2 * f->cap * sizeof(int), s, d_frames, nx,
                               ny, f->cap, f->d_fg, f->d_info, f->d_fg_off, f->d_scratch, f->d_staged, 0);
        }
        hipLaunchKernelGGL(scan_frames_kernel, dim3(1), dim3(1024), 0, s, f->d_info, nf, 1, f->cap, f->d_out_off,
                           d_count, max_candidates);
        if (nframes > 0)
            hipLaunchKernelGGL(compact_kernel, dim3(4, nframes), dim3(256), 0, s, f->d_staged, f->d_fg_off,
                               f->d_out_off, f->d_info, nframes, d_count, d_candidates);
        FIND_CHECK(hipGetLastError());
        FIND_CHECK(f->order.end(s));
    } catch (const std::string &e) {
        f->err = e;
        return -1;
    }
    return 0;
}
