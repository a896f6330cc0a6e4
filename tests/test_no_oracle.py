"""Pins the field-operator oracle (oracle/no_ref.py) on the CPU: against
torch's conv3d and autograd (library routines) and against central finite
differences of the loss."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import no_ref


def torch_weight(w: np.ndarray) -> torch.Tensor:
    taps, co, ci = w.shape
    k = 3 if taps == 27 else 1
    return torch.from_numpy(w.reshape(k, k, k, co, ci).transpose(3, 4, 0, 1, 2).copy())


def torch_conv(x, w, b, dil):
    xt = torch.from_numpy(x).permute(0, 4, 1, 2, 3)
    k = 3 if w.shape[0] == 27 else 1
    y = torch.nn.functional.conv3d(xt, torch_weight(w), None if b is None else torch.from_numpy(b),
                                   padding=dil if k == 3 else 0, dilation=dil)
    return y.permute(0, 2, 3, 4, 1)


@pytest.mark.parametrize("taps,dil", [(27, 1), (27, 2), (27, 3), (1, 1)])
def test_conv_matches_torch(taps, dil):
    rng = np.random.default_rng(taps + dil)
    x = rng.standard_normal((2, 5, 4, 6, 3))
    w = rng.standard_normal((taps, 4, 3))
    b = rng.standard_normal(4)
    got = no_ref.conv(x, w, b, dil)
    want = torch_conv(x, w, b, dil).numpy()
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("taps,dil", [(27, 1), (27, 2), (1, 1)])
def test_conv_grads_match_autograd(taps, dil):
    rng = np.random.default_rng(7 + dil)
    x = rng.standard_normal((2, 4, 5, 3, 3))
    w = rng.standard_normal((taps, 4, 3))
    dy = rng.standard_normal((2, 4, 5, 3, 4))
    xt = torch.from_numpy(x).requires_grad_(True)
    wt = torch_weight(w).requires_grad_(True)
    k = 3 if taps == 27 else 1
    y = torch.nn.functional.conv3d(xt.permute(0, 4, 1, 2, 3), wt, padding=dil if k == 3 else 0, dilation=dil)
    (y.permute(0, 2, 3, 4, 1) * torch.from_numpy(dy)).sum().backward()
    gx = no_ref.conv_grad_x(dy, w, dil)
    gw = no_ref.conv_grad_w(x, dy, taps, dil)
    np.testing.assert_allclose(gx, xt.grad.numpy(), rtol=1e-10, atol=1e-10)
    gw_t = wt.grad.numpy().transpose(2, 3, 4, 0, 1).reshape(taps, 4, 3)
    np.testing.assert_allclose(gw, gw_t, rtol=1e-10, atol=1e-10)


def small_net(rng, cin=4, width=6, cout=5, dil=(1, 2)):
    p = {"w0": rng.standard_normal((27, width, cin)) * 0.3, "b0": rng.standard_normal(width) * 0.1,
         "wh": rng.standard_normal((1, cout, width)) * 0.3, "bh": rng.standard_normal(cout) * 0.1}
    for i in range(len(dil)):
        p[f"w{i + 1}"] = rng.standard_normal((27, width, width)) * 0.2
        p[f"b{i + 1}"] = rng.standard_normal(width) * 0.1
    return no_ref.Net(p, list(dil))


def test_net_gradient_finite_differences():
    rng = np.random.default_rng(3)
    net = small_net(rng)
    a0 = rng.standard_normal((1, 4, 3, 5, 4))
    target = rng.standard_normal((1, 4, 3, 5, 3))
    mask = (rng.random((1, 4, 3, 5)) > 0.3).astype(np.float64)

    def loss():
        pred, _ = net.forward(a0)
        return no_ref.Net.loss(pred, target, mask, 3)[0]

    pred, cache = net.forward(a0)
    _, dpred = no_ref.Net.loss(pred, target, mask, 3)
    g = net.backward(cache, dpred)
    h = 1e-6
    for name in ("w0", "b0", "w1", "b2", "w2", "wh", "bh"):
        arr = net.p[name]
        flat = arr.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = loss()
            flat[idx] = old - h
            lm = loss()
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(g[name].reshape(-1)[idx] - fd) <= 1e-6 + 1e-5 * abs(fd), (name, idx)


def test_net_gradient_matches_autograd():
    rng = np.random.default_rng(11)
    net = small_net(rng)
    a0 = rng.standard_normal((2, 3, 4, 5, 4))
    target = rng.standard_normal((2, 3, 4, 5, 3))
    mask = np.ones((2, 3, 4, 5))
    pred, cache = net.forward(a0)
    _, dpred = no_ref.Net.loss(pred, target, mask, 3)
    g = net.backward(cache, dpred)
    tp = {k: torch.from_numpy(v.copy()).requires_grad_(True) for k, v in net.p.items()}

    def tconv(x, w, b, dil):
        k = 3 if w.shape[0] == 27 else 1
        taps, co, ci = w.shape
        wt = w.reshape(k, k, k, co, ci).permute(3, 4, 0, 1, 2)
        y = torch.nn.functional.conv3d(x.permute(0, 4, 1, 2, 3), wt, b, padding=dil if k == 3 else 0,
                                       dilation=dil)
        return y.permute(0, 2, 3, 4, 1)

    x = torch.relu(tconv(torch.from_numpy(a0), tp["w0"], tp["b0"], 1))
    for i, d in enumerate(net.dil):
        x = x + torch.relu(tconv(x, tp[f"w{i + 1}"], tp[f"b{i + 1}"], d))
    pr = tconv(x, tp["wh"], tp["bh"], 1)
    l = ((pr[..., :3] - torch.from_numpy(target)) ** 2).mean()
    l.backward()
    for k in net.p:
        np.testing.assert_allclose(g[k], tp[k].grad.numpy(), rtol=1e-9, atol=1e-11, err_msg=k)


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e38, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9], np.float32)
    got = no_ref.bf16_round(x)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(got, want)
    rng = np.random.default_rng(0)
    y = (rng.standard_normal(10000) * 10.0 ** rng.uniform(-20, 20, 10000)).astype(np.float32)
    np.testing.assert_array_equal(no_ref.bf16_round(y), torch.from_numpy(y).to(torch.bfloat16).float().numpy())
