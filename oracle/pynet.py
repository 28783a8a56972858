"""CPU oracle of the learned initializer's network (TEST INFRASTRUCTURE ONLY).

The network is the one of include/localizer_hip.h (LZR_NET_*): 11 x 11 patch ->
conv 3x3 1->32 + ReLU -> conv 3x3 32->32 + ReLU -> flatten (index (y*7+x)*32+c) ->
FC 64 + ReLU -> FC 5.  The reference has no such network (it always starts the LM at
the particle, source/Localization.cpp:138-142); this module defines its arithmetic:

  * rounding=True reproduces the GPU's rounding points: the input, the weights and
    every stored activation / activation gradient rounded to BF16 (round to nearest
    even), sums taken exactly (FP64 here, FP32 on the GPU), biases in FP32;
  * rounding=False is the same network in plain FP64 (for finite differences and the
    comparison with torch autograd that pin the backward pass).

Forward and backward are written with numpy im2col; the gradient layout is the
canonical parameter vector (LZR_NET_NPARAMS floats).
"""
import numpy as np

from localizer_amd import abi

H = abi.NET_HALF_WIDTH
N = 2 * H + 1
OFF = abi.NET_OFF
NPARAMS = abi.NET_NPARAMS


def bf16(a):
    """Round to the nearest BF16 value (ties to even) via float32, returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def unpack(params):
    p = np.asarray(params, np.float64)
    return {k: p[o:o + int(np.prod(shape))].reshape(shape) for k, (o, shape) in OFF.items()}


def pack(grads):
    out = np.zeros(NPARAMS)
    for k, (o, shape) in OFF.items():
        out[o:o + int(np.prod(shape))] = np.asarray(grads[k]).reshape(-1)
    return out


def prep(patches):
    """patches uint16 [n, 11, 11] -> (x [n, 11, 11] float64 (BF16 values), median, scale):
    x = bf16(float32(v - median) / float32(scale)), scale = max - median, or 1 if < 1."""
    v = np.asarray(patches, np.float32).reshape(len(patches), -1)
    med = np.median(v, axis=1).astype(np.float32)  # odd count: the middle value
    scale = v.max(axis=1) - med
    scale = np.where(scale >= 1.0, scale, np.float32(1.0)).astype(np.float32)
    x = (v - med[:, None]) / scale[:, None]
    return bf16(x).reshape(-1, N, N), med.astype(np.float64), scale.astype(np.float64)


def _windows(a, k=3):
    """a [n, S, S, ...] -> [n, S-k+1, S-k+1, k*k, ...] with tap t = dy*k + dx."""
    n, S = a.shape[0], a.shape[1]
    T = S - k + 1
    return np.stack([a[:, dy:dy + T, dx:dx + T] for dy in range(k) for dx in range(k)], axis=3)


def forward(params, x, rounding=True):
    """x [n, 11, 11] -> cache dict with the output 't' [n, 5]."""
    P = unpack(params)
    R = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W1 = R(P["conv1_w"].reshape(32, 9))
    W2 = R(P["conv2_w"])  # [o][c][dy][dx]
    Wf1 = R(P["fc1_w"])
    Wf2 = R(P["fc2_w"])
    b1, b2, bf1, bf2 = P["conv1_b"], P["conv2_b"], P["fc1_b"], P["fc2_b"]
    n = x.shape[0]
    X1 = _windows(x)                                     # [n, 9, 9, 9]
    a1 = R(np.maximum(X1 @ W1.T + b1, 0.0))              # [n, 9, 9, 32]
    X2 = _windows(a1)                                    # [n, 7, 7, 9, 32]
    W2r = W2.reshape(32, 32, 9).transpose(0, 2, 1)       # [o][t][c]
    a2 = R(np.maximum(np.einsum("nyxtc,otc->nyxo", X2, W2r) + b2, 0.0))
    f = a2.reshape(n, -1)                                # (y*7 + x)*32 + o
    h = R(np.maximum(f @ Wf1.T + bf1, 0.0))
    t = h @ Wf2.T + bf2
    return dict(x=x, X1=X1, a1=a1, X2=X2, a2=a2, f=f, h=h, t=t, W1=W1, W2r=W2r, Wf1=Wf1, Wf2=Wf2, rounding=rounding)


def backward(params, cache, targets, weights):
    """Loss = sum_i w_i sum_k (t_ik - target_ik)^2 / sum_i w_i and its gradient
    (canonical vector), with the GPU's rounding points when the cache was rounded."""
    R = bf16 if cache["rounding"] else (lambda a: np.asarray(a, np.float64))
    t = cache["t"]
    w = np.asarray(weights, np.float64)
    nv = w.sum()
    inv = 1.0 / nv if nv > 0 else 0.0
    e = (t - targets) * (w[:, None] > 0)
    loss = float((e * e).sum() * inv)
    d = 2.0 * e * inv                                    # [n, 5]
    h, f, a1, a2, X1, X2 = cache["h"], cache["f"], cache["a1"], cache["a2"], cache["X1"], cache["X2"]
    g = {}
    g["fc2_w"] = d.T @ h
    g["fc2_b"] = d.sum(0)
    dh = R((d @ cache["Wf2"]) * (h > 0))
    g["fc1_w"] = dh.T @ f
    g["fc1_b"] = dh.sum(0)
    df = R((dh @ cache["Wf1"]) * (f > 0))
    n = t.shape[0]
    da2 = df.reshape(n, 7, 7, 32)
    g["conv2_b"] = da2.sum((0, 1, 2))
    gW2r = np.einsum("nyxtc,nyxo->otc", X2, da2)         # [o][t][c]
    g["conv2_w"] = gW2r.transpose(0, 2, 1).reshape(32, 32, 3, 3)
    dX2 = np.einsum("nyxo,otc->nyxtc", da2, cache["W2r"])
    da1 = np.zeros_like(a1)
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        da1[:, dy:dy + 7, dx:dx + 7] += dX2[:, :, :, tap]
    da1 = R(da1 * (a1 > 0))
    g["conv1_b"] = da1.sum((0, 1, 2))
    g["conv1_w"] = np.einsum("nyxt,nyxc->ct", X1, da1).reshape(32, 3, 3)
    return loss, pack(g)


def targets_from_fits(fits, stats, w):
    """Normalized targets and weights from LM fits (accepted symmetric fits only);
    stats = (median, scale, xoff, yoff, valid)."""
    med, scale, xoff, yoff, valid = stats
    use = (fits["status"] == abi.ACCEPTED) & (fits["model"] == abi.MODEL_SYMMETRIC) & valid & (fits["params"][:, 1] > 0)
    p = fits["params"]
    with np.errstate(all="ignore"):
        t = np.stack([p[:, 0] / scale, np.log(p[:, 1] / w), p[:, 2] - (xoff + H), p[:, 3] - (yoff + H),
                      (p[:, 4] - med) / scale], 1)
    t = np.where(use[:, None], t, 0.0).astype(np.float32).astype(np.float64)
    return t, use.astype(np.float64)


def patches_for(frames, cands):
    """The 11 x 11 patches the GPU reads (x_offset = floor(x - 5)) and their stats."""
    n = len(cands)
    out = np.zeros((n, N, N), np.uint16)
    xoff = np.floor(cands["x"] - H).astype(np.int64)
    yoff = np.floor(cands["y"] - H).astype(np.int64)
    nf, ny, nx = frames.shape
    valid = (xoff >= 0) & (yoff >= 0) & (xoff + N <= nx) & (yoff + N <= ny) & (cands["frame"] >= 0) & (cands["frame"] < nf)
    for i in np.nonzero(valid)[0]:
        out[i] = frames[cands["frame"][i], yoff[i]:yoff[i] + N, xoff[i]:xoff[i] + N]
    return out, xoff, yoff, valid
