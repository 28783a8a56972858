"""ctypes binding of the C ABI (include/localizer_hip.h, localizer_amd/lib/liblocalizer_hip.so).

Every call goes through the HIP library; there is no CPU path.  Loading fails loudly
when the library has not been built.
"""
import ctypes
import os

import numpy as np

from . import abi

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "liblocalizer_hip.so")

# Every function declared in include/localizer_hip.h.
EXPORTED = (
    "lzr_create",
    "lzr_clone",
    "lzr_destroy",
    "lzr_last_error",
    "lzr_set_option",
    "lzr_fit_batch",
    "lzr_fit_batch_device",
    "lzr_fit_batch_f32",
    "lzr_fit_batch_device_f32",
    "lzr_abi_version",
    "lzr_fit_batch_device_start",
    "lzr_fit_batch_device_netinit",
    "lzr_fit_batch_device_noalloc",
    "lzr_net_create",
    "lzr_net_destroy",
    "lzr_net_last_error",
    "lzr_net_set_params",
    "lzr_net_get_params",
    "lzr_net_forward_device",
    "lzr_net_predict_device",
    "lzr_net_train_step_device",
    "lzr_get_option",
    "lzr_graph_create",
    "lzr_graph_launch",
    "lzr_graph_destroy",
    "lzr_autotune_device",
    "lzr_finder_create",
    "lzr_finder_destroy",
    "lzr_finder_last_error",
    "lzr_find_candidates_device",
    "lzr_fit_emitters",
    "lzr_localize",
    "lzr_free",
    "lzr_pipeline_last_error",
    "lzr_renderer_create",
    "lzr_renderer_destroy",
    "lzr_renderer_last_error",
    "lzr_render_image_size",
    "lzr_render_device",
    "lzr_render",
    "lzr_positions_from_fits_device",
    "lzr_batcher_create",
    "lzr_batcher_fit_frame",
    "lzr_batcher_fit_frame_f64",
    "lzr_batcher_stats",
    "lzr_batcher_last_error",
    "lzr_batcher_destroy",
)


class LocalizerError(RuntimeError):
    """Raised when the C ABI returns -1 (the reference raises std::runtime_error)."""


def lib():
    global _LIB
    if _LIB is None:
        # torch ships its own libamdhip64 (same SONAME).  Loading it first makes this
        # library bind to that runtime instead of starting a second HIP runtime in the
        # process, which would leave torch without a device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise LocalizerError(f"{LIB_PATH} is missing: build it with `python -m localizer_amd.build hip`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.lzr_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.lzr_create.restype = ctypes.c_int
        L.lzr_destroy.argtypes = [vp]
        L.lzr_destroy.restype = ctypes.c_int
        L.lzr_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.lzr_last_error.restype = ctypes.c_int
        L.lzr_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lzr_set_option.restype = ctypes.c_int
        L.lzr_fit_batch.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp]
        L.lzr_fit_batch.restype = ctypes.c_int
        L.lzr_fit_batch_device.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp]
        L.lzr_fit_batch_device.restype = ctypes.c_int
        L.lzr_fit_batch_f32.argtypes = L.lzr_fit_batch.argtypes
        L.lzr_fit_batch_f32.restype = ctypes.c_int
        L.lzr_fit_batch_device_f32.argtypes = L.lzr_fit_batch_device.argtypes
        L.lzr_fit_batch_device_f32.restype = ctypes.c_int
        L.lzr_abi_version.argtypes = []
        L.lzr_abi_version.restype = ctypes.c_int
        i, d, f = ctypes.c_int, ctypes.c_double, ctypes.c_float
        L.lzr_fit_batch_device_start.argtypes = [vp, vp, i, i, i, vp, vp, i, vp, vp]
        L.lzr_fit_batch_device_netinit.argtypes = [vp, vp, vp, i, i, i, vp, i, vp, vp, d, i, vp, vp]
        L.lzr_net_create.argtypes = [ctypes.POINTER(vp), d, ctypes.c_ulonglong]
        L.lzr_net_destroy.argtypes = [vp]
        L.lzr_net_last_error.argtypes = [vp, ctypes.c_char_p, i]
        L.lzr_net_set_params.argtypes = [vp, vp, i]
        L.lzr_net_get_params.argtypes = [vp, vp, i]
        L.lzr_net_forward_device.argtypes = [vp, vp, i, i, i, vp, i, vp, vp]
        L.lzr_net_predict_device.argtypes = [vp, vp, i, i, i, vp, i, vp, vp]
        L.lzr_net_train_step_device.argtypes = [vp, vp, i, i, i, vp, vp, i, f, i, ctypes.POINTER(f), vp, vp]
        L.lzr_get_option.argtypes = [vp, i, ctypes.POINTER(d)]
        L.lzr_fit_batch_device_noalloc.argtypes = [vp, vp, i, i, i, vp, i, vp, vp]
        L.lzr_graph_create.argtypes = [vp, ctypes.POINTER(vp), vp, i, i, i, vp, vp, i, vp]
        L.lzr_graph_launch.argtypes = [vp, vp]
        L.lzr_graph_destroy.argtypes = [vp]
        L.lzr_autotune_device.argtypes = [vp, vp, i, i, i, vp, i, vp, vp, ctypes.POINTER(d)]
        L.lzr_finder_create.argtypes = [ctypes.POINTER(vp), d, d, i]
        L.lzr_finder_destroy.argtypes = [vp]
        L.lzr_finder_last_error.argtypes = [vp, ctypes.c_char_p, i]
        L.lzr_find_candidates_device.argtypes = [vp, vp, i, i, i, vp, i, vp, vp]
        L.lzr_fit_emitters.argtypes = [vp, i, i, i, d, d, ctypes.POINTER(i), ctypes.POINTER(ctypes.POINTER(d))]
        L.lzr_fit_emitters.restype = ctypes.c_int
        L.lzr_localize.argtypes = [vp, i, i, i, d, d, i, ctypes.POINTER(i), ctypes.POINTER(i),
                                   ctypes.POINTER(ctypes.POINTER(d))]
        L.lzr_localize.restype = ctypes.c_int
        L.lzr_free.argtypes = [ctypes.POINTER(d)]
        L.lzr_free.restype = None
        L.lzr_pipeline_last_error.argtypes = [ctypes.c_char_p, i]
        L.lzr_pipeline_last_error.restype = ctypes.c_int
        L.lzr_batcher_create.argtypes = [ctypes.POINTER(vp), vp, i, i, i, i, i, i]
        L.lzr_clone.argtypes = [vp, ctypes.POINTER(vp)]
        L.lzr_clone.restype = ctypes.c_int
        L.lzr_batcher_fit_frame.argtypes = [vp, vp, i, i, vp, i, vp]
        L.lzr_batcher_fit_frame_f64.argtypes = [vp, vp, i, i, vp, i, vp]
        L.lzr_batcher_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]
        L.lzr_batcher_last_error.argtypes = [vp, ctypes.c_char_p, i]
        L.lzr_batcher_destroy.argtypes = [vp]
        for name in EXPORTED:
            if name.startswith(("lzr_net", "lzr_fit_batch_device_", "lzr_graph", "lzr_get", "lzr_autotune", "lzr_find",
                                "lzr_render", "lzr_positions", "lzr_batcher")):
                getattr(L, name).restype = ctypes.c_int
        if L.lzr_abi_version() != 1:
            raise LocalizerError("ABI version mismatch")
        _LIB = L
    return _LIB


def _error(ctx):
    buf = ctypes.create_string_buffer(1024)
    lib().lzr_last_error(ctx, buf, len(buf))
    return buf.value.decode(errors="replace")


class Context:
    """One fitting context (lzr_create); the Python face of a FitPositions_* instance."""

    def __init__(self, model, initial_psf_width, sigma=1.0, default_half_width=0):
        h = ctypes.c_void_p()
        if lib().lzr_create(ctypes.byref(h), int(model), float(initial_psf_width), float(sigma),
                            int(default_half_width)) != 0:
            raise LocalizerError(_error(None))
        self._h = h
        self._destroy = lib().lzr_destroy  # held so that closing works during interpreter shutdown
        self.model = int(model)
        self.initial_psf_width = float(initial_psf_width)
        self.sigma = float(sigma)

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, option, value):
        if lib().lzr_set_option(self._h, int(option), float(value)) != 0:
            raise LocalizerError(_error(self._h))

    def fit(self, frames, candidates):
        """Host arrays in, FIT_DTYPE array out: uint16 frames through lzr_fit_batch, floating-
        point frames (any pixel values) through lzr_fit_batch_f32."""
        flt = np.issubdtype(np.asarray(frames).dtype, np.floating)
        frames = np.ascontiguousarray(frames, np.float32 if flt else np.uint16)
        if frames.ndim == 2:
            frames = frames[None]
        nf, ny, nx = frames.shape
        cands = np.ascontiguousarray(candidates, abi.CANDIDATE_DTYPE)
        out = np.zeros(cands.size, abi.FIT_DTYPE)
        fn = lib().lzr_fit_batch_f32 if flt else lib().lzr_fit_batch
        rc = fn(self._h, frames.ctypes.data, nx, ny, nf, cands.ctypes.data, cands.size, out.ctypes.data)
        if rc < 0:
            raise LocalizerError(_error(self._h))
        return out

    def fit_device_f32(self, d_frames, nx, ny, nframes, d_cands, ncands, d_out, stream=0):
        """lzr_fit_batch_device_f32: float frames in device memory."""
        rc = lib().lzr_fit_batch_device_f32(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                            ctypes.c_void_p(d_cands), int(ncands), ctypes.c_void_p(d_out),
                                            ctypes.c_void_p(stream))
        if rc < 0:
            raise LocalizerError(_error(self._h))
        return rc

    def fit_device_start(self, d_frames, nx, ny, nframes, d_cands, d_starts, ncands, d_out, stream=0):
        """lzr_fit_batch_device_start: per-candidate LM start points (d_starts may be 0)."""
        rc = lib().lzr_fit_batch_device_start(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                              ctypes.c_void_p(d_cands), ctypes.c_void_p(d_starts), int(ncands),
                                              ctypes.c_void_p(d_out), ctypes.c_void_p(stream))
        if rc < 0:
            raise LocalizerError(_error(self._h))
        return rc

    def fit_device_netinit(self, net, d_frames, nx, ny, nframes, d_cands, ncands, d_out, d_work,
                           residual_factor=3.0, refit_rejected=False, d_nrefit=0, stream=0):
        """lzr_fit_batch_device_netinit: network start, fallback to the reference start for
        fits whose residual exceeds residual_factor times the residual at the reference start
        and, with refit_rejected, for rejected or non-converged fits; d_nrefit: optional device
        uint32 that receives the number of refits."""
        rc = lib().lzr_fit_batch_device_netinit(self._h, net._h, ctypes.c_void_p(d_frames), int(nx), int(ny),
                                                int(nframes), ctypes.c_void_p(d_cands), int(ncands),
                                                ctypes.c_void_p(d_out), ctypes.c_void_p(d_work),
                                                float(residual_factor), int(bool(refit_rejected)),
                                                ctypes.c_void_p(d_nrefit), ctypes.c_void_p(stream))
        if rc < 0:
            raise LocalizerError(_error(self._h))
        return rc

    def get_option(self, option):
        v = ctypes.c_double()
        if lib().lzr_get_option(self._h, int(option), ctypes.byref(v)) != 0:
            raise LocalizerError(_error(self._h))
        return v.value

    def autotune_device(self, d_frames, nx, ny, nframes, d_cands, ncands, d_out, stream=0):
        """lzr_autotune_device: returns the best time (ms); the context keeps that configuration."""
        best = ctypes.c_double()
        rc = lib().lzr_autotune_device(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                       ctypes.c_void_p(d_cands), int(ncands), ctypes.c_void_p(d_out),
                                       ctypes.c_void_p(stream), ctypes.byref(best))
        if rc != 0:
            raise LocalizerError(_error(self._h))
        return best.value

    def graph(self, d_frames, nx, ny, nframes, d_cands, d_count, max_candidates, d_out):
        """A frame-block graph (lzr_graph_create) on fixed device buffers."""
        return Graph(self, d_frames, nx, ny, nframes, d_cands, d_count, max_candidates, d_out)

    def fit_device(self, d_frames, nx, ny, nframes, d_cands, ncands, d_out, stream=0):
        """Device pointers in (ints), enqueued on `stream` (lzr_fit_batch_device)."""
        rc = lib().lzr_fit_batch_device(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                        ctypes.c_void_p(d_cands), int(ncands), ctypes.c_void_p(d_out),
                                        ctypes.c_void_p(stream))
        if rc < 0:
            raise LocalizerError(_error(self._h))
        return rc


class Graph:
    """lzr_graph_*: one captured fit replayed per frame block."""

    def __init__(self, ctx, d_frames, nx, ny, nframes, d_cands, d_count, max_candidates, d_out):
        h = ctypes.c_void_p()
        if lib().lzr_graph_create(ctx._h, ctypes.byref(h), ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                  ctypes.c_void_p(d_cands), ctypes.c_void_p(d_count), int(max_candidates),
                                  ctypes.c_void_p(d_out)) != 0:
            raise LocalizerError(_error(ctx._h))
        self._h = h
        self._ctx = ctx
        self._destroy = lib().lzr_graph_destroy

    def launch(self, stream=0):
        if lib().lzr_graph_launch(self._h, ctypes.c_void_p(stream)) != 0:
            raise LocalizerError(_error(self._ctx._h))

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    __del__ = close


def _net_error(h):
    buf = ctypes.create_string_buffer(1024)
    lib().lzr_net_last_error(h, buf, len(buf))
    return buf.value.decode(errors="replace")


class Net:
    """The learned initializer (lzr_net_*): a CNN on 11 x 11 patches whose forward and
    backward passes run on the GPU."""

    def __init__(self, initial_psf_width, seed=0x10CA112E):
        h = ctypes.c_void_p()
        if lib().lzr_net_create(ctypes.byref(h), float(initial_psf_width), int(seed)) != 0:
            raise LocalizerError(_net_error(None))
        self._h = h
        self._destroy = lib().lzr_net_destroy
        self.initial_psf_width = float(initial_psf_width)

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise LocalizerError(_net_error(self._h))

    def get_params(self):
        p = np.zeros(abi.NET_NPARAMS, np.float32)
        self._check(lib().lzr_net_get_params(self._h, p.ctypes.data, p.size))
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._check(lib().lzr_net_set_params(self._h, p.ctypes.data, p.size))

    def forward_device(self, d_frames, nx, ny, nframes, d_cands, ncands, d_out, stream=0):
        self._check(lib().lzr_net_forward_device(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                                 ctypes.c_void_p(d_cands), int(ncands), ctypes.c_void_p(d_out),
                                                 ctypes.c_void_p(stream)))

    def predict_device(self, d_frames, nx, ny, nframes, d_cands, ncands, d_starts, stream=0):
        self._check(lib().lzr_net_predict_device(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                                 ctypes.c_void_p(d_cands), int(ncands), ctypes.c_void_p(d_starts),
                                                 ctypes.c_void_p(stream)))

    def train_step_device(self, d_frames, nx, ny, nframes, d_cands, d_labels, ncands, lr, apply=True,
                          want_loss=True, want_grads=False, stream=0):
        """One step; returns (loss or None, grads or None)."""
        loss = ctypes.c_float(0.0)
        grads = np.zeros(abi.NET_NPARAMS, np.float32) if want_grads else None
        self._check(lib().lzr_net_train_step_device(
            self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes), ctypes.c_void_p(d_cands),
            ctypes.c_void_p(d_labels), int(ncands), float(lr), int(bool(apply)),
            ctypes.byref(loss) if want_loss else None, ctypes.c_void_p(grads.ctypes.data) if want_grads else None,
            ctypes.c_void_p(stream)))
        return (loss.value if want_loss else None), grads


def _finder_error(h):
    buf = ctypes.create_string_buffer(1024)
    lib().lzr_finder_last_error(h, buf, len(buf))
    return buf.value.decode(errors="replace")


class Finder:
    """GLRT segmentation + 8-way particle finding on the device (lzr_finder_*)."""

    def __init__(self, psf_width, pfa, max_segmented_per_frame=0):
        h = ctypes.c_void_p()
        if lib().lzr_finder_create(ctypes.byref(h), float(psf_width), float(pfa), int(max_segmented_per_frame)) != 0:
            raise LocalizerError(_finder_error(None))
        self._h = h
        self._destroy = lib().lzr_finder_destroy

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def find_device(self, d_frames, nx, ny, nframes, d_cands, max_candidates, d_count, stream=0):
        if lib().lzr_find_candidates_device(self._h, ctypes.c_void_p(d_frames), int(nx), int(ny), int(nframes),
                                            ctypes.c_void_p(d_cands), int(max_candidates), ctypes.c_void_p(d_count),
                                            ctypes.c_void_p(stream)) != 0:
            raise LocalizerError(_finder_error(self._h))


def localize(frames, psf_width, glrt_insensitivity, model=abi.MODEL_ELLIPSOIDAL):
    """Whole pipeline on host frames [n, ny, nx] (lzr_localize): the reference's positions
    matrix (rows sorted by frame)."""
    frames = np.ascontiguousarray(frames, np.uint16)
    if frames.ndim == 2:
        frames = frames[None]
    nf, ny, nx = frames.shape
    n, cols = ctypes.c_int(), ctypes.c_int()
    ptr = ctypes.POINTER(ctypes.c_double)()
    if lib().lzr_localize(frames.ctypes.data, nx, ny, nf, float(psf_width), float(glrt_insensitivity), int(model),
                          ctypes.byref(n), ctypes.byref(cols), ctypes.byref(ptr)) != 0:
        buf = ctypes.create_string_buffer(1024)
        lib().lzr_pipeline_last_error(buf, len(buf))
        raise LocalizerError(buf.value.decode(errors="replace"))
    try:
        flat = np.ctypeslib.as_array(ptr, shape=(max(n.value, 1) * cols.value,)).copy()
    finally:
        lib().lzr_free(ptr)
    return flat[: n.value * cols.value].reshape(cols.value, n.value).T.copy()


def _render_error(h):
    buf = ctypes.create_string_buffer(1024)
    lib().lzr_renderer_last_error(h, buf, len(buf))
    return buf.value.decode(errors="replace")


def render_image_size(x_size, y_size, output_scale):
    """(width, height) of the localization bitmap (lzr_render_image_size)."""
    w, h = ctypes.c_int(), ctypes.c_int()
    if lib().lzr_render_image_size(int(x_size), int(y_size), ctypes.c_double(output_scale), ctypes.byref(w),
                                   ctypes.byref(h)) != 0:
        raise LocalizerError("output scale must be > 0 and the image at least 2x2 pixels")
    return w.value, h.value


class Renderer:
    """Localization bitmap on the device (lzr_renderer_*, PALMBitmapImageCalculator)."""

    def __init__(self):
        h = ctypes.c_void_p()
        if lib().lzr_renderer_create(ctypes.byref(h)) != 0:
            raise LocalizerError(_render_error(None))
        self._h = h
        self._destroy = lib().lzr_renderer_destroy

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def render_device(self, d_positions, n, x_size, y_size, output_scale, method, params, weighing, d_image, stream=0):
        p = (ctypes.c_double * 3)(*(list(params) + [0.0, 0.0, 0.0])[:3])
        if lib().lzr_render_device(self._h, ctypes.c_void_p(d_positions), int(n), int(x_size), int(y_size),
                                   ctypes.c_double(output_scale), int(method), p, int(weighing),
                                   ctypes.c_void_p(d_image), ctypes.c_void_p(stream)) != 0:
            raise LocalizerError(_render_error(self._h))

    def render(self, positions, x_size, y_size, output_scale, method, params, weighing):
        """Host positions (abi.POSITION_DTYPE) -> image [height, width] (lzr_render)."""
        pos = np.ascontiguousarray(positions, abi.POSITION_DTYPE)
        w, h = render_image_size(x_size, y_size, output_scale)
        img = np.zeros((h, w))
        p = (ctypes.c_double * 3)(*(list(params) + [0.0, 0.0, 0.0])[:3])
        if lib().lzr_render(self._h, pos.ctypes.data if pos.size else None, int(pos.size), int(x_size), int(y_size),
                            ctypes.c_double(output_scale), int(method), p, int(weighing), img.ctypes.data) != 0:
            raise LocalizerError(_render_error(self._h))
        return img

    def positions_from_fits_device(self, d_fits, n, d_positions, d_count, stream=0):
        if lib().lzr_positions_from_fits_device(self._h, ctypes.c_void_p(d_fits), int(n), ctypes.c_void_p(d_positions),
                                                ctypes.c_void_p(d_count), ctypes.c_void_p(stream)) != 0:
            raise LocalizerError(_render_error(self._h))


def _batcher_error(h):
    buf = ctypes.create_string_buffer(1024)
    lib().lzr_batcher_last_error(h, buf, len(buf))
    return buf.value.decode(errors="replace")


class Batcher:
    """lzr_batcher_*: concurrent per-frame fit calls (the reference controller's worker
    threads) gathered into one device launch.  fit_frame may be called from any number
    of threads (ctypes releases the GIL during the call)."""

    def __init__(self, ctx, nx, ny, max_frames=64, max_rois=1 << 16, max_wait_us=500, inflight=4):
        h = ctypes.c_void_p()
        if lib().lzr_batcher_create(ctypes.byref(h), ctx._h, int(nx), int(ny), int(max_frames), int(max_rois),
                                    int(max_wait_us), int(inflight)) != 0:
            raise LocalizerError(_batcher_error(None))
        self._h = h
        self._destroy = lib().lzr_batcher_destroy
        self.nx, self.ny = int(nx), int(ny)

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def fit_frame(self, frame, candidates):
        """One frame (ny x nx, uint16 or float64) and its candidates (frame field 0) ->
        FIT_DTYPE records, equal to a one-frame lzr_fit_batch call."""
        frame = np.asarray(frame)
        f64 = frame.dtype == np.float64
        frame = np.ascontiguousarray(frame, np.float64 if f64 else np.uint16)
        ny, nx = frame.shape
        cands = np.ascontiguousarray(candidates, abi.CANDIDATE_DTYPE)
        out = np.zeros(cands.size, abi.FIT_DTYPE)
        fn = lib().lzr_batcher_fit_frame_f64 if f64 else lib().lzr_batcher_fit_frame
        if fn(self._h, frame.ctypes.data, nx, ny, cands.ctypes.data, cands.size, out.ctypes.data) < 0:
            raise LocalizerError(_batcher_error(self._h))
        return out

    def stats(self):
        b, f = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        lib().lzr_batcher_stats(self._h, ctypes.byref(b), ctypes.byref(f))
        return {"batches": b.value, "frames": f.value}
