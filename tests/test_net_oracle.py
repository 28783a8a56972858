"""Pins the learned initializer's CPU oracle (oracle/pynet.py): BF16 rounding against
torch's bfloat16 conversion, forward and gradient against torch autograd (FP64, an
independent implementation of the same network), gradient against central finite
differences, and the patch normalization against its definition."""
import numpy as np
import pytest
import torch

from localizer_amd import abi
from oracle import pynet


def _rand_params(rng, scale=1.0):
    p = np.zeros(abi.NET_NPARAMS)
    for k, (o, shape) in abi.NET_OFF.items():
        fan = {"conv1_w": 9, "conv2_w": 288, "fc1_w": 1568, "fc2_w": 64}.get(k, 4)
        p[o:o + int(np.prod(shape))] = rng.normal(0, scale * np.sqrt(2.0 / fan), int(np.prod(shape)))
    return p


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-3, 2000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.0078125, 3.0e38])]).astype(np.float32)
    # ties: values exactly half-way between two BF16 numbers
    ties = (np.arange(1, 2001, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ours = pynet.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(ours, ref)


def _torch_net(params, x):
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in pynet.unpack(params).items()}
    xt = torch.tensor(x[:, None], dtype=torch.float64)
    a1 = torch.relu(torch.nn.functional.conv2d(xt, P["conv1_w"][:, None], P["conv1_b"]))
    a2 = torch.relu(torch.nn.functional.conv2d(a1, P["conv2_w"], P["conv2_b"]))
    f = a2.permute(0, 2, 3, 1).reshape(len(x), -1)  # index (y*7 + x)*32 + c
    h = torch.relu(f @ P["fc1_w"].T + P["fc1_b"])
    return h @ P["fc2_w"].T + P["fc2_b"], P


def test_forward_and_gradient_match_torch_autograd():
    rng = np.random.default_rng(2)
    params = _rand_params(rng)
    x = rng.normal(0, 1, (6, 11, 11))
    tgt = rng.normal(0, 1, (6, 5))
    w = np.array([1, 1, 0, 1, 1, 1.0])
    cache = pynet.forward(params, x, rounding=False)
    loss, g = pynet.backward(params, cache, tgt, w)
    out, P = _torch_net(params, x)
    np.testing.assert_allclose(cache["t"], out.detach().numpy(), rtol=1e-12, atol=1e-12)
    wt = torch.tensor(w)
    L = (((out - torch.tensor(tgt)) ** 2).sum(1) * wt).sum() / wt.sum()
    L.backward()
    assert abs(L.item() - loss) < 1e-12 * max(1.0, loss)
    gt = pynet.pack({k: v.grad.numpy() for k, v in P.items()})
    np.testing.assert_allclose(g, gt, rtol=1e-9, atol=1e-12 * np.abs(gt).max())


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = _rand_params(rng)
    x = rng.normal(0, 1, (3, 11, 11))
    tgt = rng.normal(0, 1, (3, 5))
    w = np.ones(3)
    _, g = pynet.backward(params, pynet.forward(params, x, rounding=False), tgt, w)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), min(4, int(np.prod(s))), replace=False)
                          for o, s in abi.NET_OFF.values()])
    for i in idx:
        e = 1e-6 * max(1.0, abs(params[i]))
        pp, pm = params.copy(), params.copy()
        pp[i] += e
        pm[i] -= e
        lp, _ = pynet.backward(pp, pynet.forward(pp, x, rounding=False), tgt, w)
        lm, _ = pynet.backward(pm, pynet.forward(pm, x, rounding=False), tgt, w)
        fd = (lp - lm) / (2 * e)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(g[i])), (i, fd, g[i])


def test_patch_normalization():
    rng = np.random.default_rng(4)
    p = rng.integers(90, 140, (50, 11, 11)).astype(np.uint16)
    p[0] = 100  # flat patch: scale falls back to 1
    p[1, 5, 5] = 1000
    x, med, scale = pynet.prep(p)
    for i in range(50):
        v = np.sort(p[i].reshape(-1).astype(np.float64))
        assert med[i] == v[60]
        s = max(v[-1] - v[60], 1.0) if v[-1] - v[60] >= 1.0 else 1.0
        assert scale[i] == s
    assert np.all(x[0] == 0.0) and scale[0] == 1.0
    assert x[1, 5, 5] == 1.0  # the maximum maps to 1 exactly


@pytest.mark.parametrize("rounding", [True, False])
def test_rounded_forward_stays_close_to_fp64(rounding):
    """BF16 rounding changes outputs by ~1e-2 relative at most on He-scale weights."""
    rng = np.random.default_rng(5)
    params = _rand_params(rng)
    x = pynet.bf16(rng.normal(0, 0.5, (20, 11, 11)))
    t64 = pynet.forward(params, x, rounding=False)["t"]
    t = pynet.forward(params, x, rounding=rounding)["t"]
    assert np.abs(t - t64).max() <= (5e-2 if rounding else 0.0) * np.abs(t64).max()


def test_net_selection_threshold_choice():
    """netinit.best_threshold: the candidate threshold with the most counted localizations
    per second wins; equal rates go to the lower threshold (fewer network starts)."""
    from localizer_amd import netinit
    thr = [-np.inf, 10.0, 20.0, np.inf]
    assert netinit.best_threshold(thr, [100, 110, 120, 125], [1.0, 1.0, 1.0, 1.1]) == (20.0, 2)
    assert netinit.best_threshold(thr, [100, 100, 100, 100], [1.0, 1.0, 1.0, 1.0]) == (-np.inf, 0)
    assert netinit.best_threshold(thr, [100, 100, 100, 100], [2.0, 1.5, 1.0, 0.5]) == (np.inf, 3)
