"""CPU oracle of the field operator -- TEST INFRASTRUCTURE.

A float64 numpy restatement of the operator's layers (3-D convolution with
zero padding and dilation, ReLU, residual add, 1x1 head, masked MSE) and of
their gradients, on channels-last arrays [B][Z][Y][X][C] with weights
[taps][Cout][Cin] (tap t = 9 (dz + 1) + 3 (dy + 1) + (dx + 1), offsets times
the dilation). The reference has no such operator; these functions are
pinned in tests against torch's CPU conv3d and autograd (library routines)
and against central finite differences. Only tests/ and the smoke check use
this module; the engine never does.
"""
from __future__ import annotations

import numpy as np


def tap_offsets(taps: int, dil: int) -> list[tuple[int, int, int]]:
    if taps == 1:
        return [(0, 0, 0)]
    return [((t // 9 - 1) * dil, ((t // 3) % 3 - 1) * dil, (t % 3 - 1) * dil) for t in range(27)]


def _shift(x: np.ndarray, dz: int, dy: int, dx: int) -> np.ndarray:
    """s[b, z, y, x] = x[b, z + dz, y + dy, x + dx], zero outside."""
    B, Z, Y, X, C = x.shape
    out = np.zeros_like(x)
    zs, ze = max(0, -dz), min(Z, Z - dz)
    ys, ye = max(0, -dy), min(Y, Y - dy)
    xs, xe = max(0, -dx), min(X, X - dx)
    if zs < ze and ys < ye and xs < xe:
        out[:, zs:ze, ys:ye, xs:xe] = x[:, zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv(x: np.ndarray, w: np.ndarray, bias: np.ndarray | None, dil: int) -> np.ndarray:
    """y[m, co] = bias[co] + sum_t sum_ci w[t, co, ci] x[m + off_t, ci]."""
    taps = w.shape[0]
    y = np.zeros(x.shape[:4] + (w.shape[1],), np.float64)
    for t, (dz, dy, dx) in enumerate(tap_offsets(taps, dil)):
        y += _shift(x.astype(np.float64), dz, dy, dx) @ w[t].astype(np.float64).T
    if bias is not None:
        y += bias
    return y


def conv_grad_w(x: np.ndarray, dy: np.ndarray, taps: int, dil: int) -> np.ndarray:
    g = np.zeros((taps, dy.shape[-1], x.shape[-1]), np.float64)
    dyf = dy.reshape(-1, dy.shape[-1]).astype(np.float64)
    for t, (dz, dy_, dx) in enumerate(tap_offsets(taps, dil)):
        xs = _shift(x.astype(np.float64), dz, dy_, dx).reshape(-1, x.shape[-1])
        g[t] = dyf.T @ xs
    return g


def conv_grad_x(dy: np.ndarray, w: np.ndarray, dil: int) -> np.ndarray:
    """Adjoint of conv: dx[q, ci] = sum_t sum_co w[t, co, ci] dy[q - off_t, co]."""
    taps = w.shape[0]
    dx = np.zeros(dy.shape[:4] + (w.shape[2],), np.float64)
    for t, (dz, dy_, dx_) in enumerate(tap_offsets(taps, dil)):
        dx += _shift(dy.astype(np.float64), -dz, -dy_, -dx_) @ w[t].astype(np.float64)
    return dx


class Net:
    """conv0 (ReLU) -> blocks of x + ReLU(conv(x)) -> 1x1 head; masked MSE."""

    def __init__(self, params: dict, dilations: list[int]):
        self.p = {k: np.asarray(v, np.float64) for k, v in params.items()}
        self.dil = list(dilations)

    def forward(self, a0: np.ndarray) -> tuple[np.ndarray, dict]:
        p = self.p
        cache = {"a0": a0}
        x = np.maximum(conv(a0, p["w0"], p["b0"], 1), 0.0)
        cache["x0"] = x
        for i, d in enumerate(self.dil):
            h = np.maximum(conv(x, p[f"w{i + 1}"], p[f"b{i + 1}"], d), 0.0)
            cache[f"h{i + 1}"] = h
            x = x + h
            cache[f"x{i + 1}"] = x
        L = len(self.dil)
        pred = conv(x, p["wh"], p["bh"], 1)
        return pred, cache

    @staticmethod
    def loss(pred, target, mask, nf):
        d = (pred[..., :nf] - target) * mask[..., None]
        cnt = mask.sum() * nf
        return float((d * d).sum() / cnt), np.concatenate(
            [2.0 * d / cnt, np.zeros(pred.shape[:-1] + (pred.shape[-1] - nf,))], axis=-1)

    def backward(self, cache: dict, dpred: np.ndarray) -> dict:
        p = self.p
        L = len(self.dil)
        g = {}
        xL = cache[f"x{L}"]
        g["wh"] = conv_grad_w(xL, dpred, 1, 1)
        g["bh"] = dpred.reshape(-1, dpred.shape[-1]).sum(0)
        dx = conv_grad_x(dpred, p["wh"], 1)
        for i in range(L, 0, -1):
            dz = dx * (cache[f"h{i}"] > 0)
            xin = cache[f"x{i - 1}"]
            g[f"w{i}"] = conv_grad_w(xin, dz, 27, self.dil[i - 1])
            g[f"b{i}"] = dz.reshape(-1, dz.shape[-1]).sum(0)
            dx = dx + conv_grad_x(dz, p[f"w{i}"], self.dil[i - 1])
        dz = dx * (cache["x0"] > 0)
        g["w0"] = conv_grad_w(cache["a0"], dz, 27, 1)
        g["b0"] = dz.reshape(-1, dz.shape[-1]).sum(0)
        return g


def inputs(p: np.ndarray, sphere: np.ndarray, box, shape, C: int) -> np.ndarray:
    """Operator input features (see dsmc_no_inputs), float32 before rounding."""
    B = p.shape[0]
    Z, Y, X = shape
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    fx, fy, fz = (x + 0.5) / X, (y + 0.5) / Y, (z + 0.5) / Z
    out = np.zeros((B, Z, Y, X, C), np.float32)
    for b in range(B):
        px = box[0] + fx * (box[3] - box[0])
        py = box[1] + fy * (box[4] - box[1])
        pz = box[2] + fz * (box[5] - box[2])
        r = np.sqrt((px - sphere[b, 0]) ** 2 + (py - sphere[b, 1]) ** 2 + (pz - sphere[b, 2]) ** 2)
        out[b, ..., 0] = 2 * fx - 1
        out[b, ..., 1] = 2 * fy - 1
        out[b, ..., 2] = 2 * fz - 1
        out[b, ..., 3] = (r - sphere[b, 3]) / (box[4] - box[1])
        out[b, ..., 4:9] = p[b]
    return out


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)
