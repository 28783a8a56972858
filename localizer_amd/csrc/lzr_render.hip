/*
 * lzr_render.hip -- the localization bitmap on the GPU and its C ABI
 * (include/localizer_hip.h, "Localization bitmap").
 *
 * PALMBitmapImageCalculator::CalculateImage (source/PALMImages.cpp:68-157) adds, for
 * every localization in container order, a Gaussian integrated over each output pixel
 * of a box of +-4 deviations around it.  The sum at a pixel is a sequence of FP64
 * additions in position order; the GPU keeps exactly that order, so the image is the
 * reference computation's bit for bit (given the same erf table):
 *   1. prep: one thread per position computes the deviation (the three
 *      PALMBitmapImageDeviationCalculator kinds, source/PALMImages.h:106-160,
 *      source/PALMImages.cpp:159-177), the weight and the clamped pixel box, and counts
 *      the 16x16 output tiles the box touches;
 *   2. (tile, position) pairs are written in position order and radix-sorted by tile
 *      (stable: each tile's list stays in position order);
 *   3. one 256-thread block per tile, one thread per pixel: the tile's positions are
 *      staged through LDS 256 at a time and each thread accumulates its pixel in order.
 * Each pixel is written once (no atomics, no prior clear).  The erf table is the
 * reference's ERFLookupTable (erf at -4 + i*1e-4, linear interpolation), built on the
 * host with the C library's erf and kept on the device.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/localizer_hip.h"
#include "lzr_order.h"

#define RND_CHECK(call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) throw std::string(#call ": ") + hipGetErrorString(e_);        \
    } while (0)

namespace lzr_rnd {

constexpr double kErfLower = -4.0; /* source/PALMImages.cpp:35-37 */
constexpr double kErfUpper = 4.0;
constexpr double kErfStride = 0.0001;
constexpr int kTile = 16;

int erf_table_size() { return (int)((kErfUpper - kErfLower) / kErfStride + 1); }

/* One position prepared for the tile kernel. */
struct Item {
    double x, y, s, w;
    int r0, r1, c0, c1; /* inclusive pixel box; empty when r0 > r1 or c0 > c1 */
};

/* double -> int as the reference's conversion behaves on x86-64: INT_MIN when the
 * value is not representable (NaN, +-inf, out of range). */
__device__ __forceinline__ int x86_int(double v) {
    return (v >= -2147483648.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

struct Geometry {
    int width, height, ntr, ntc;
    double sf, sps, half;
};

__global__ void prep_kernel(const LzrPosition *pos, int n, Geometry g, int method, double p0, double p1, double p2,
                            int weighing, Item *items, long long *tiles_per_pos, int *tile_count, int *error) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const LzrPosition P = pos[k];
    Item it;
    it.x = P.x;
    it.y = P.y;
    it.r0 = 1;
    it.r1 = 0;
    it.c0 = 1;
    it.c1 = 0;
    it.s = 1.0;
    it.w = 0.0;
    int ntiles = 0;
    if (!(P.integral == 0.0 && weighing == LZR_WEIGHING_INTEGRAL)) {
        double dev;
        if (method == LZR_DEVIATION_CONSTANT) {
            dev = p0;
        } else if (method == LZR_DEVIATION_FIT_UNCERTAINTY) {
            dev = P.x_deviation < p1 ? P.x_deviation * p0 : -1.0;
        } else {
            if (P.integral == -1.0 || P.background == -1.0) atomicOr(error, 1);
            double I = floor(P.integral / p2 + 0.5);
            I = I < 0.0 ? 0.0 : I;
            const double w2 = p0 * p0;
            dev = sqrt((w2 + 1.0 / 12.0) / I);
        }
        it.s = sqrt(2.0 * dev * dev);
        it.w = weighing == LZR_WEIGHING_SAME ? 1.0 : P.integral;
        const double cr = P.x * g.sf, cc = P.y * g.sf;
        int r0 = x86_int(floor(cr - 4.0 * dev * g.sf));
        int c0 = x86_int(floor(cc - 4.0 * dev * g.sf));
        int r1 = x86_int(ceil(cr + 4.0 * dev * g.sf));
        int c1 = x86_int(ceil(cc + 4.0 * dev * g.sf));
        if (r0 < 0) r0 = 0;
        if (r1 >= g.width) r1 = g.width - 1;
        if (c0 < 0) c0 = 0;
        if (c1 >= g.height) c1 = g.height - 1;
        it.r0 = r0;
        it.r1 = r1;
        it.c0 = c0;
        it.c1 = c1;
        if (r0 <= r1 && c0 <= c1) {
            const int tr0 = r0 / kTile, tr1 = r1 / kTile, tc0 = c0 / kTile, tc1 = c1 / kTile;
            ntiles = (tr1 - tr0 + 1) * (tc1 - tc0 + 1);
            for (int tc = tc0; tc <= tc1; ++tc)
                for (int tr = tr0; tr <= tr1; ++tr) atomicAdd(&tile_count[tc * g.ntr + tr], 1);
        }
    }
    items[k] = it;
    tiles_per_pos[k] = ntiles;
}

/* (tile, position) pairs in position order. */
__global__ void pairs_kernel(const Item *items, int n, const long long *offsets, Geometry g, int *keys, int *vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Item it = items[k];
    if (!(it.r0 <= it.r1 && it.c0 <= it.c1)) return;
    long long o = offsets[k];
    for (int tc = it.c0 / kTile; tc <= it.c1 / kTile; ++tc)
        for (int tr = it.r0 / kTile; tr <= it.r1 / kTile; ++tr) {
            keys[o] = tc * g.ntr + tr;
            vals[o] = k;
            ++o;
        }
}

__device__ __forceinline__ double erf_lookup(const double *table, double x) {
#pragma clang fp contract(off)
    if (x != x) return x; /* NaN (a zero deviation): no table index */
    if (x < kErfLower) return -1.0;
    if (x > kErfUpper) return 1.0;
    const double fi = (x - kErfLower) / kErfStride;
    const int index = (int)floor(fi);
    const double fraction = fi - (double)index;
    const double a = table[index] * (1.0 - fraction);
    const double b = table[index + 1] * fraction;
    return a + b;
}

__global__ void __launch_bounds__(256) render_kernel(const Item *items, const int *order, const int *tile_start,
                                                     Geometry g, const double *table, double *image) {
#pragma clang fp contract(off)
    __shared__ Item sh[256];
    const int t = blockIdx.x;
    const int tr = t % g.ntr, tc = t / g.ntr;
    const int i = tr * kTile + (int)(threadIdx.x % kTile);
    const int j = tc * kTile + (int)(threadIdx.x / kTile);
    const double minX = (double)i * g.sps - g.half;
    const double maxX = minX + g.sps;
    const double minY = (double)j * g.sps - g.half;
    const double maxY = minY + g.sps;
    const int b = tile_start[t], e = tile_start[t + 1];
    double acc = 0.0;
    for (int k0 = b; k0 < e; k0 += 256) {
        const int k = k0 + (int)threadIdx.x;
        if (k < e) sh[threadIdx.x] = items[order[k]];
        __syncthreads();
        const int m = min(256, e - k0);
        for (int u = 0; u < m; ++u) {
            const Item it = sh[u];
            if (i >= it.r0 && i <= it.r1 && j >= it.c0 && j <= it.c1) {
                const double eminx = erf_lookup(table, (minX - it.x) / it.s);
                const double emaxx = erf_lookup(table, (maxX - it.x) / it.s);
                const double eminy = erf_lookup(table, (minY - it.y) / it.s);
                const double emaxy = erf_lookup(table, (maxY - it.y) / it.s);
                const double t1 = eminx * eminy, t2 = emaxx * eminy, t3 = eminx * emaxy, t4 = emaxx * emaxy;
                const double v = 0.25 * (((t1 - t2) - t3) + t4);
                const double add = v * it.w;
                acc = acc + add;
            }
        }
        __syncthreads();
    }
    if (i < g.width && j < g.height) image[(size_t)j * g.width + i] = acc;
}

template <typename T>
__global__ void fill(T *p, int n, T v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

/* LzrFit -> LzrPosition (the container getters of source/Storage.h: x, y, integral,
 * x deviation, background) and the accepted flag. */
__global__ void fits_to_positions(const LzrFit *fits, int n, LzrPosition *pos, char *flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const LzrFit f = fits[k];
    const int xi = f.model == LZR_MODEL_SYMMETRIC ? 2 : (f.model == LZR_MODEL_FIXED_WIDTH ? 1 : 3);
    const int bi = f.model == LZR_MODEL_SYMMETRIC ? 4 : (f.model == LZR_MODEL_FIXED_WIDTH ? 3 : 6);
    LzrPosition p;
    p.x = f.params[xi];
    p.y = f.params[xi + 1];
    p.integral = f.integral;
    p.x_deviation = f.deviations[xi];
    p.background = f.params[bi];
    pos[k] = p;
    flag[k] = f.status == LZR_ACCEPTED ? 1 : 0;
}

template <typename T>
void grow(T *&p, size_t &cap, size_t need) {
    if (need <= cap) return;
    if (p) RND_CHECK(hipFree(p));
    p = nullptr;
    cap = 0;
    RND_CHECK(hipMalloc((void **)&p, std::max<size_t>(need, 1) * sizeof(T)));
    cap = need;
}

}  // namespace lzr_rnd

using namespace lzr_rnd;

struct LzrRenderer {
    std::string err;
    double *d_table = nullptr;
    Item *d_items = nullptr;
    size_t items_cap = 0;
    long long *d_counts = nullptr, *d_offsets = nullptr;
    size_t counts_cap = 0, offsets_cap = 0;
    int *d_tile_count = nullptr, *d_tile_start = nullptr;
    size_t tile_cap = 0, tile_start_cap = 0;
    int *d_keys = nullptr, *d_vals = nullptr, *d_keys2 = nullptr, *d_vals2 = nullptr;
    size_t keys_cap = 0, vals_cap = 0, keys2_cap = 0, vals2_cap = 0;
    void *d_temp = nullptr;
    size_t temp_cap = 0;
    int *d_error = nullptr;
    LzrPosition *d_all = nullptr;
    size_t all_cap = 0;
    char *d_flag = nullptr;
    size_t flag_cap = 0;
    LzrStreamOrder order; /* calls on different streams share the buffers above */
};

static thread_local std::string g_render_create_error;

static int rfail(LzrRenderer *r, const std::string &m) {
    if (r) r->err = m;
    return -1;
}

static void ensure_temp(LzrRenderer *r, size_t bytes) {
    if (bytes <= r->temp_cap) return;
    if (r->d_temp) RND_CHECK(hipFree(r->d_temp));
    r->d_temp = nullptr;
    r->temp_cap = 0;
    RND_CHECK(hipMalloc(&r->d_temp, bytes));
    r->temp_cap = bytes;
}

LZR_EXPORT int lzr_renderer_create(LzrRenderer **out) {
    if (!out) return -1;
    *out = nullptr;
    LzrRenderer *r = new LzrRenderer();
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::string("no HIP device available");
        const int n = erf_table_size();
        std::vector<double> table((size_t)n + 1);
        for (int i = 0; i < n; ++i) table[i] = std::erf(kErfLower + (double)i * kErfStride);
        table[n] = table[n - 1]; /* read with weight 0 at x = 4 exactly */
        RND_CHECK(hipMalloc((void **)&r->d_table, table.size() * sizeof(double)));
        RND_CHECK(hipMemcpy(r->d_table, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice));
        RND_CHECK(hipMalloc((void **)&r->d_error, sizeof(int)));
    } catch (const std::string &e) {
        g_render_create_error = "lzr_renderer_create: " + e;
        lzr_renderer_destroy(r);
        return -1;
    }
    *out = r;
    return 0;
}

LZR_EXPORT int lzr_renderer_destroy(LzrRenderer *r) {
    if (!r) return 0;
    void *bufs[] = {r->d_table, r->d_items, r->d_counts, r->d_offsets, r->d_tile_count, r->d_tile_start, r->d_keys,
                    r->d_vals, r->d_keys2, r->d_vals2, r->d_temp, r->d_error, r->d_all, r->d_flag};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    r->order.destroy();
    delete r;
    return 0;
}

LZR_EXPORT int lzr_renderer_last_error(const LzrRenderer *r, char *buf, int len) {
    const std::string &e = r ? r->err : g_render_create_error;
    if (buf && len > 0) {
        const int n = std::min((int)e.size(), len - 1);
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}

LZR_EXPORT int lzr_render_image_size(int x_size, int y_size, double output_scale, int *width, int *height) {
    if (!width || !height) return -1;
    if (!(output_scale > 0.0) || x_size < 2 || y_size < 2) return -1;
    /* int imageWidth = xSize * outputScaleFactor (source/PALMImages.cpp:74-75) */
    const double w = (double)x_size * output_scale, h = (double)y_size * output_scale;
    if (!(w < 2147483648.0 && h < 2147483648.0)) return -1;
    *width = (int)w;
    *height = (int)h;
    return (*width >= 2 && *height >= 2) ? 0 : -1;
}

LZR_EXPORT int lzr_render_device(LzrRenderer *r, const LzrPosition *d_positions, int n_positions, int x_size,
                                 int y_size, double output_scale, int deviation_method,
                                 const double *deviation_params, int weighing, double *d_image, void *stream) {
    if (!r) return -1;
    int W = 0, H = 0;
    if (lzr_render_image_size(x_size, y_size, output_scale, &W, &H) != 0)
        return rfail(r, "lzr_render_device: output_scale must be > 0 and the image at least 2x2 pixels");
    if (deviation_method < LZR_DEVIATION_CONSTANT || deviation_method > LZR_DEVIATION_GAUSSIAN_MASK)
        return rfail(r, "lzr_render_device: unknown deviation method");
    if (weighing != LZR_WEIGHING_SAME && weighing != LZR_WEIGHING_INTEGRAL)
        return rfail(r, "Unrecognized emitter weighing method while calculating a PALM bitmap");
    if (n_positions < 0 || (n_positions > 0 && !d_positions) || !d_image || !deviation_params)
        return rfail(r, "lzr_render_device: null pointer or negative count");
    const hipStream_t s = (hipStream_t)stream;
    try {
        Geometry g;
        g.width = W;
        g.height = H;
        g.ntr = (W + kTile - 1) / kTile;
        g.ntc = (H + kTile - 1) / kTile;
        g.sf = (double)(W - 1) / (double)(x_size - 1);
        g.sps = 1.0 / g.sf;
        g.half = 0.5 * g.sps;
        const size_t ntiles = (size_t)g.ntr * g.ntc;
        if (ntiles >= (size_t)INT_MAX) throw std::string("image too large");
        const int n = n_positions;
        RND_CHECK(r->order.begin(s));
        grow(r->d_items, r->items_cap, (size_t)n);
        grow(r->d_counts, r->counts_cap, (size_t)n + 1);
        grow(r->d_offsets, r->offsets_cap, (size_t)n + 1);
        grow(r->d_tile_count, r->tile_cap, ntiles + 1);
        grow(r->d_tile_start, r->tile_start_cap, ntiles + 1);
        const int tpb = 256;
        hipLaunchKernelGGL(fill<int>, dim3((unsigned)((ntiles + 1 + tpb - 1) / tpb)), dim3(tpb), 0, s,
                           r->d_tile_count, (int)ntiles + 1, 0);
        hipLaunchKernelGGL(fill<long long>, dim3((n + 1 + tpb - 1) / tpb), dim3(tpb), 0, s, r->d_counts, n + 1, 0ll);
        hipLaunchKernelGGL(fill<int>, dim3(1), dim3(64), 0, s, r->d_error, 1, 0);
        if (n > 0)
            hipLaunchKernelGGL(prep_kernel, dim3((n + tpb - 1) / tpb), dim3(tpb), 0, s, d_positions, n, g,
                               deviation_method, deviation_params[0], deviation_params[1], deviation_params[2],
                               weighing, r->d_items, r->d_counts, r->d_tile_count, r->d_error);
        /* offsets of the pairs (the total lands in offsets[n]) and of the tiles */
        size_t t1 = 0, t2 = 0;
        RND_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, r->d_counts, r->d_offsets, n + 1, s));
        RND_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, r->d_tile_count, r->d_tile_start, (int)ntiles + 1, s));
        ensure_temp(r, std::max(t1, t2));
        RND_CHECK(hipcub::DeviceScan::ExclusiveSum(r->d_temp, t1, r->d_counts, r->d_offsets, n + 1, s));
        RND_CHECK(hipcub::DeviceScan::ExclusiveSum(r->d_temp, t2, r->d_tile_count, r->d_tile_start, (int)ntiles + 1,
                                                   s));
        long long total = 0;
        int error = 0;
        RND_CHECK(hipMemcpyAsync(&total, r->d_offsets + n, sizeof(long long), hipMemcpyDeviceToHost, s));
        RND_CHECK(hipMemcpyAsync(&error, r->d_error, sizeof(int), hipMemcpyDeviceToHost, s));
        RND_CHECK(hipStreamSynchronize(s));
        if (error) /* the reference's text (source/PALMImages.cpp:164) */
            return rfail(r, "the used positions do not provide a PSF width and/or background estimate "
                            "(use a different localization algorithm)");
        if (total < 0 || total >= (long long)INT_MAX) throw std::string("too many (tile, position) pairs");
        const int M = (int)total;
        if (M > 0) {
            grow(r->d_keys, r->keys_cap, (size_t)M);
            grow(r->d_vals, r->vals_cap, (size_t)M);
            grow(r->d_keys2, r->keys2_cap, (size_t)M);
            grow(r->d_vals2, r->vals2_cap, (size_t)M);
            hipLaunchKernelGGL(pairs_kernel, dim3((n + tpb - 1) / tpb), dim3(tpb), 0, s, r->d_items, n, r->d_offsets, g,
                               r->d_keys, r->d_vals);
            int bits = 1;
            while (bits < 31 && ((size_t)1 << bits) < ntiles) ++bits;
            size_t t3 = 0;
            RND_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t3, r->d_keys, r->d_keys2, r->d_vals, r->d_vals2, M, 0,
                                                         bits, s));
            ensure_temp(r, t3);
            RND_CHECK(hipcub::DeviceRadixSort::SortPairs(r->d_temp, t3, r->d_keys, r->d_keys2, r->d_vals, r->d_vals2, M,
                                                         0, bits, s));
        }
        hipLaunchKernelGGL(render_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, r->d_items, r->d_vals2,
                           r->d_tile_start, g, r->d_table, d_image);
        RND_CHECK(hipGetLastError());
        RND_CHECK(r->order.end(s));
    } catch (const std::string &e) {
        return rfail(r, "lzr_render_device: " + e);
    }
    return 0;
}

LZR_EXPORT int lzr_render(LzrRenderer *r, const LzrPosition *positions, int n_positions, int x_size, int y_size,
                          double output_scale, int deviation_method, const double *deviation_params, int weighing,
                          double *image) {
    if (!r) return -1;
    int W = 0, H = 0;
    if (lzr_render_image_size(x_size, y_size, output_scale, &W, &H) != 0)
        return rfail(r, "lzr_render: output_scale must be > 0 and the image at least 2x2 pixels");
    if (n_positions < 0 || (n_positions > 0 && !positions) || !image)
        return rfail(r, "lzr_render: null pointer or negative count");
    LzrPosition *d_pos = nullptr;
    double *d_img = nullptr;
    hipStream_t s = nullptr;
    int rc = -1;
    try {
        RND_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        RND_CHECK(hipMalloc((void **)&d_pos, std::max(n_positions, 1) * sizeof(LzrPosition)));
        RND_CHECK(hipMalloc((void **)&d_img, (size_t)W * H * sizeof(double)));
        if (n_positions > 0)
            RND_CHECK(hipMemcpyAsync(d_pos, positions, n_positions * sizeof(LzrPosition), hipMemcpyHostToDevice, s));
        if (lzr_render_device(r, d_pos, n_positions, x_size, y_size, output_scale, deviation_method, deviation_params,
                              weighing, d_img, s) == 0) {
            RND_CHECK(hipMemcpyAsync(image, d_img, (size_t)W * H * sizeof(double), hipMemcpyDeviceToHost, s));
            RND_CHECK(hipStreamSynchronize(s));
            rc = 0;
        }
    } catch (const std::string &e) {
        rfail(r, "lzr_render: " + e);
    }
    if (d_pos) (void)hipFree(d_pos);
    if (d_img) (void)hipFree(d_img);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

LZR_EXPORT int lzr_positions_from_fits_device(LzrRenderer *r, const LzrFit *d_fits, int n_fits,
                                              LzrPosition *d_positions, int *d_count, void *stream) {
    if (!r) return -1;
    if (n_fits < 0 || (n_fits > 0 && (!d_fits || !d_positions)) || !d_count)
        return rfail(r, "lzr_positions_from_fits_device: null pointer or negative count");
    const hipStream_t s = (hipStream_t)stream;
    try {
        if (n_fits == 0) {
            hipLaunchKernelGGL(fill<int>, dim3(1), dim3(64), 0, s, d_count, 1, 0);
            return 0;
        }
        RND_CHECK(r->order.begin(s));
        grow(r->d_all, r->all_cap, (size_t)n_fits);
        grow(r->d_flag, r->flag_cap, (size_t)n_fits);
        const int tpb = 256;
        hipLaunchKernelGGL(fits_to_positions, dim3((n_fits + tpb - 1) / tpb), dim3(tpb), 0, s, d_fits, n_fits,
                           r->d_all, r->d_flag);
        size_t t = 0;
        RND_CHECK(hipcub::DeviceSelect::Flagged(nullptr, t, r->d_all, r->d_flag, d_positions, d_count, n_fits, s));
        ensure_temp(r, t);
        RND_CHECK(hipcub::DeviceSelect::Flagged(r->d_temp, t, r->d_all, r->d_flag, d_positions, d_count, n_fits, s));
        RND_CHECK(hipGetLastError());
        RND_CHECK(r->order.end(s));
    } catch (const std::string &e) {
        return rfail(r, "lzr_positions_from_fits_device: " + e);
    }
    return 0;
}
