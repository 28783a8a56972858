"""ctypes front of tests/emu/liblzr_emu.so (host build of the GPU lane logic).
TEST INFRASTRUCTURE ONLY."""
import ctypes
import os

import numpy as np

from localizer_amd import abi

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.environ.get("LZR_EMU_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblzr_emu.so")
        L = ctypes.CDLL(path)
        L.emu_fit_batch.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p]
        L.emu_fit_batch.restype = ctypes.c_int
        L.emu_sep_pass.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.emu_sep_pass.restype = ctypes.c_int
        L.emu_ell_pass.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p]
        L.emu_ell_pass.restype = ctypes.c_int
        L.emu_sep_pass_group.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.emu_sep_pass_group.restype = ctypes.c_int
        L.emu_ell_pass_group.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.emu_ell_pass_group.restype = ctypes.c_int
        L.emu_start_residual.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_void_p]
        L.emu_start_residual.restype = None
        _LIB = L
    return _LIB


def fit_batch(frames, cands, model, w, sigma=1.0, default_half_width=0, max_half_width=0, diag_rows=-1,
              max_iter=200, epsabs=1e-4, epsrel=1e-4):
    frames = np.ascontiguousarray(frames, np.uint16)
    if frames.ndim == 2:
        frames = frames[None]
    nf, ny, nx = frames.shape
    cands = np.ascontiguousarray(cands, abi.CANDIDATE_DTYPE)
    h = default_half_width if default_half_width > 0 else abi.default_half_width(w)
    out = np.zeros(cands.size, abi.FIT_DTYPE)
    lib().emu_fit_batch(int(model), float(w), float(sigma), int(h), int(max(max_half_width, h)), int(diag_rows),
                        int(max_iter), float(epsabs), float(epsrel), frames.ctypes.data,
                        nx, ny, nf, cands.ctypes.data, cands.size, out.ctypes.data)
    return out


def start_residual(frames, cands, model, w, sigma=1.0, default_half_width=0):
    """|f| at each candidate's reference start point (lzr_lane.h start_residual)."""
    frames = np.ascontiguousarray(frames, np.uint16)
    if frames.ndim == 2:
        frames = frames[None]
    nf, ny, nx = frames.shape
    cands = np.ascontiguousarray(cands, abi.CANDIDATE_DTYPE)
    h = default_half_width if default_half_width > 0 else abi.default_half_width(w)
    out = np.zeros(cands.size)
    lib().emu_start_residual(int(model), float(w), float(sigma), int(h), 15, frames.ctypes.data, nx, ny, nf,
                             cands.ctypes.data, cands.size, out.ctypes.data)
    return out


def sep_pass(theta, dx, sigma, w, window, nloop=None):
    """Normal-equation sums of one separable pass (lzr_lane.h sep_pass + sep_finish) at
    theta + dx over `window` (n x n, [j][i]); returns (hq, sj, gq, sf, sse)."""
    theta = np.ascontiguousarray(theta, np.float64)
    dx = np.ascontiguousarray(dx, np.float64)
    p = theta.size
    window = np.ascontiguousarray(window, np.uint16)
    n = window.shape[0]
    sq = (p - 1) * p // 2
    out = np.zeros(sq + 2 * (p - 1) + 2)
    rc = lib().emu_sep_pass(p, theta.ctypes.data, dx.ctypes.data, float(sigma), float(w), n, int(nloop or n),
                            window.ctypes.data, out.ctypes.data)
    assert rc == 0
    return out[:sq], out[sq:sq + p - 1], out[sq + p - 1:sq + 2 * (p - 1)], out[-2], out[-1]


def ell_pass(theta, dx, sigma, window, nloop=None):
    """Normal-equation sums of one ellipsoidal moment pass (lzr_lane.h ell_pass +
    ell_finish) at theta + dx over `window`; returns (hq, sj, gq, sf, sse)."""
    theta = np.ascontiguousarray(theta, np.float64)
    dx = np.ascontiguousarray(dx, np.float64)
    window = np.ascontiguousarray(window, np.uint16)
    n = window.shape[0]
    out = np.zeros(21 + 6 + 6 + 2)
    rc = lib().emu_ell_pass(theta.ctypes.data, dx.ctypes.data, float(sigma), n, int(nloop or n), window.ctypes.data,
                            out.ctypes.data)
    assert rc == 0
    return out[:21], out[21:27], out[27:33], out[-2], out[-1]


def sep_pass_group(theta, dx, sigma, w, window, k):
    """The separable pass as a straggler group of k lanes computes it (every lane's copy must
    agree); returns the flat output of sep_pass."""
    theta = np.ascontiguousarray(theta, np.float64)
    dx = np.ascontiguousarray(dx, np.float64)
    p = theta.size
    window = np.ascontiguousarray(window, np.uint16)
    n = window.shape[0]
    out = np.zeros((p - 1) * p // 2 + 2 * (p - 1) + 2)
    assert lib().emu_sep_pass_group(p, theta.ctypes.data, dx.ctypes.data, float(sigma), float(w), n, int(k),
                                    window.ctypes.data, out.ctypes.data) == 0
    return out


def sep_pass_flat(theta, dx, sigma, w, window):
    return np.concatenate([np.atleast_1d(a) for a in sep_pass(theta, dx, sigma, w, window)])


def ell_pass_group(theta, dx, sigma, window, k):
    """The ellipsoidal moment pass as a straggler group of k lanes computes it."""
    theta = np.ascontiguousarray(theta, np.float64)
    dx = np.ascontiguousarray(dx, np.float64)
    window = np.ascontiguousarray(window, np.uint16)
    n = window.shape[0]
    out = np.zeros(21 + 6 + 6 + 2)
    assert lib().emu_ell_pass_group(theta.ctypes.data, dx.ctypes.data, float(sigma), n, int(k), window.ctypes.data,
                                    out.ctypes.data) == 0
    return out


def ell_pass_flat(theta, dx, sigma, window):
    return np.concatenate([np.atleast_1d(a) for a in ell_pass(theta, dx, sigma, window)])
