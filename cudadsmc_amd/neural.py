"""Field operator: a 3-D convolutional network, trained on the GPU through
the engine's BF16-MFMA kernels, that predicts steady fields on the cell grid
from the case parameters, so that particles can be loaded close to the
steady state (warm start) instead of as a uniform freestream.

Architecture (about 2.1M parameters at width 96):

* inputs [B][Z][Y][X][32] BF16: cell-centre coordinates, signed distance to
  the sphere and the five normalised case parameters (U, n, T_inf, T_w, r),
  zero-padded to 32 channels;
* conv 3x3x3 32 -> W, ReLU;
* residual blocks x <- x + ReLU(conv 3x3x3 W -> W, dilation d_i);
* 1x1 head W -> 32, of which the first five channels are the fields
  (log n/n_inf, u/U, v/U, w/U, log T/T_inf).

Training minimises the masked mean-squared error with Adam on FP32 master
weights. Every tensor operation of the forward and backward passes is a
kernel of libdsmc_hip.so (dsmc_no_*); torch provides device memory and the
stream only.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._capi import lib

N_FIELDS = 5
C_IN = 32
C_OUT = 32
PARAM_NAMES = ("U", "n", "T_inf", "T_w", "r")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the field operator runs on the GPU only")
    return torch


def _bind() -> C.CDLL:
    L = lib()
    if getattr(L, "_no_bound", False):
        return L
    P, I, Lg, F = C.c_void_p, C.c_int, C.c_long, C.c_float
    sig = {
        "dsmc_no_conv_fwd": [P, P, P, P, P, P, P, I, I, I, I, I, I, I, I, I, P],
        "dsmc_no_weight_flip": [P, P, I, I, I, P],
        "dsmc_no_conv_bwd_w": [P, P, P, P, Lg, I, I, I, I, I, I, I, I, I, I, P],
        "dsmc_no_bias_grad": [P, Lg, I, P, P, Lg, I, P],
        "dsmc_no_relu_bwd": [P, P, P, Lg, P],
        "dsmc_no_add": [P, P, P, Lg, P],
        "dsmc_no_mse": [P, P, P, Lg, I, I, F, P, P, P],
        "dsmc_no_adam": [P, P, P, P, P, Lg, F, F, F, F, F, F, F, P],
        "dsmc_no_f32_to_bf16": [P, P, Lg, P],
        "dsmc_no_inputs": [P, P, P, I, I, I, I, I, P, P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L._no_bound = True
    return L


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# ----------------------------------------------------------- kernel wrappers
def conv_fwd(x, w, bias, B, shape, cin, cout, taps, dil, relu=False, res=None, y=None, y32=None, h=None):
    Z, Y, X = shape
    _chk(_bind().dsmc_no_conv_fwd(_p(x), _p(w), _p(bias), _p(res), _p(y), _p(y32), _p(h), B, Z, Y, X, cin, cout,
                                  taps, dil, int(relu), _stream()), "conv_fwd")


def weight_flip(w, wb, taps, cout, cin):
    _chk(_bind().dsmc_no_weight_flip(_p(w), _p(wb), taps, cout, cin, _stream()), "weight_flip")


def conv_bwd_w(x, dy, dw, part, B, shape, cin, cout, taps, dil, chunk=4096, accumulate=False):
    Z, Y, X = shape
    _chk(_bind().dsmc_no_conv_bwd_w(_p(x), _p(dy), _p(dw), _p(part), part.numel(), B, Z, Y, X, cin, cout, taps, dil,
                                    chunk, int(accumulate), _stream()), "conv_bwd_w")


def bias_grad(dy, M, cout, db, part, accumulate=False):
    _chk(_bind().dsmc_no_bias_grad(_p(dy), M, cout, _p(db), _p(part), part.numel(), int(accumulate), _stream()),
         "bias_grad")


def relu_bwd(dy, h, dz):
    _chk(_bind().dsmc_no_relu_bwd(_p(dy), _p(h), _p(dz), dy.numel(), _stream()), "relu_bwd")


def inputs(params, spheres, box, B, shape, out):
    Z, Y, X = shape
    _chk(_bind().dsmc_no_inputs(_p(params), _p(spheres), _p(box), B, Z, Y, X, C_IN, _p(out), _stream()), "inputs")


# ---------------------------------------------------------- normalisation
def normalise_params(U, n, T_inf, T_w, r) -> np.ndarray:
    """Case parameters -> operator inputs of order one."""
    return np.stack([(np.asarray(U) - 1800.0) / 1200.0,
                     (np.log10(np.asarray(n)) - 20.24) / 0.74,
                     (np.asarray(T_inf) - 225.0) / 75.0,
                     (np.asarray(T_w) - 400.0) / 200.0,
                     (np.asarray(r) - 0.025) / 0.01], axis=-1).astype(np.float32)


def fields_to_targets(n, u, v, w, T, n_inf, U, T_inf) -> np.ndarray:
    """Fields of one case ([Z][Y][X] each) -> the five operator outputs:
    log(n/n_inf), u/U, v/U, w/U, log(T/T_inf). Cells without samples (n or
    T not positive) get 0 and must be masked out of the loss."""
    n = np.asarray(n, np.float64)
    T = np.asarray(T, np.float64)
    ok_n, ok_T = n > 0, T > 0
    ln = np.where(ok_n, np.log(np.where(ok_n, n, 1.0) / n_inf), 0.0)
    lt = np.where(ok_T, np.log(np.where(ok_T, T, 1.0) / T_inf), 0.0)
    return np.stack([ln, np.asarray(u) / U, np.asarray(v) / U, np.asarray(w) / U, lt], axis=-1).astype(np.float32)


def targets_to_fields(t, n_inf, U, T_inf):
    """Inverse of fields_to_targets: returns n, u, v, w, T."""
    t = np.asarray(t, np.float64)
    return (n_inf * np.exp(t[..., 0]), U * t[..., 1], U * t[..., 2], U * t[..., 3], T_inf * np.exp(t[..., 4]))


# ------------------------------------------------------------------ network
class FieldOperator:
    def __init__(self, shape=(32, 32, 64), width: int = 96, dilations=(1, 2, 4, 8, 1, 2, 4, 8), seed: int = 0,
                 max_batch: int = 8):
        torch = _torch()
        self.torch = torch
        self.shape = tuple(int(s) for s in shape)
        self.width = int(width)
        self.dil = [int(d) for d in dilations]
        self.max_batch = int(max_batch)
        dev = torch.device("cuda")
        g = np.random.default_rng(seed)
        W = self.width
        spec = [("w0", (27, W, C_IN), math.sqrt(2.0 / (27 * C_IN))), ("b0", (W,), 0.0)]
        for i in range(len(self.dil)):
            spec += [(f"w{i + 1}", (27, W, W), 0.2 * math.sqrt(2.0 / (27 * W))), (f"b{i + 1}", (W,), 0.0)]
        spec += [("wh", (1, C_OUT, W), math.sqrt(1.0 / W)), ("bh", (C_OUT,), 0.0)]
        self.params = {}
        for name, shp, std in spec:
            a = (g.standard_normal(shp) * std).astype(np.float32) if std else np.zeros(shp, np.float32)
            self.params[name] = torch.from_numpy(a).to(dev)
        self.bf = {k: torch.zeros(v.shape, dtype=torch.bfloat16, device=dev) for k, v in self.params.items()
                   if k.startswith("w")}
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m1 = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m2 = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.flip = {k: torch.zeros(v.shape[0], v.shape[2], v.shape[1], dtype=torch.bfloat16, device=dev)
                     for k, v in self.params.items() if k.startswith("w") and k != "w0"}
        self.t = 0
        self._alloc(self.max_batch)
        self.sync_bf16()

    def n_params(self) -> int:
        return int(sum(v.numel() for v in self.params.values()))

    def _alloc(self, B: int) -> None:
        torch = self.torch
        Z, Y, X = self.shape
        M = B * Z * Y * X
        W = self.width
        L = len(self.dil)
        bf = dict(dtype=torch.bfloat16, device="cuda")
        self.M_cap = M
        self.a0 = torch.zeros(M, C_IN, **bf)
        self.xs = [torch.zeros(M, W, **bf) for _ in range(L + 1)]
        self.hs = [torch.zeros(M, W, **bf) for _ in range(L + 1)]
        self.pred = torch.zeros(M, C_OUT, dtype=torch.float32, device="cuda")
        self.dpred = torch.zeros(M, C_OUT, **bf)
        self.dx = [torch.zeros(M, W, **bf) for _ in range(2)]
        self.dz = torch.zeros(M, W, **bf)
        chunks = (M + 4095) // 4096
        self.part = torch.zeros(chunks * 27 * W * W, dtype=torch.float32, device="cuda")
        # bias-gradient partials: one row per 512 voxels
        self.bpart = torch.zeros(max((M + 511) // 512, 1) * max(W, C_OUT), dtype=torch.float32, device="cuda")
        self.loss_part = torch.zeros((M * C_OUT + 255) // 256, dtype=torch.float32, device="cuda")

    def sync_bf16(self) -> None:
        for k, v in self.bf.items():
            _chk(_bind().dsmc_no_f32_to_bf16(_p(self.params[k]), _p(v), v.numel(), _stream()), "f32_to_bf16")

    # -- passes
    def forward(self, a0, B: int):
        """a0: BF16 [B*Z*Y*X][32] device tensor; returns FP32 [M][32] (view)."""
        Z, Y, X = self.shape
        M = B * Z * Y * X
        if M > self.M_cap:
            self._alloc(B)
        W, p, bf = self.width, self.params, self.bf
        conv_fwd(a0, bf["w0"], p["b0"], B, self.shape, C_IN, W, 27, 1, relu=True, y=self.xs[0], h=self.hs[0])
        for i, d in enumerate(self.dil):
            conv_fwd(self.xs[i], bf[f"w{i + 1}"], p[f"b{i + 1}"], B, self.shape, W, W, 27, d, relu=True,
                     res=self.xs[i], y=self.xs[i + 1], h=self.hs[i + 1])
        conv_fwd(self.xs[len(self.dil)], bf["wh"], p["bh"], B, self.shape, W, C_OUT, 1, 1, y32=self.pred)
        return self.pred[:M]

    def backward(self, a0, B: int) -> None:
        """Gradients of all parameters from self.dpred (after forward)."""
        Z, Y, X = self.shape
        M = B * Z * Y * X
        W, p, bf, g = self.width, self.params, self.bf, self.grads
        L = len(self.dil)
        for k in self.flip:
            w = bf[k]
            weight_flip(w, self.flip[k], w.shape[0], w.shape[1], w.shape[2])
        conv_bwd_w(self.xs[L], self.dpred, g["wh"], self.part, B, self.shape, W, C_OUT, 1, 1)
        bias_grad(self.dpred, M, C_OUT, g["bh"], self.bpart)
        cur = 0
        conv_fwd(self.dpred, self.flip["wh"], None, B, self.shape, C_OUT, W, 1, 1, y=self.dx[cur])
        for i in range(L, 0, -1):
            relu_bwd(self.dx[cur][:M], self.hs[i][:M], self.dz[:M])
            conv_bwd_w(self.xs[i - 1], self.dz, g[f"w{i}"], self.part, B, self.shape, W, W, 27, self.dil[i - 1])
            bias_grad(self.dz, M, W, g[f"b{i}"], self.bpart)
            conv_fwd(self.dz, self.flip[f"w{i}"], None, B, self.shape, W, W, 27, self.dil[i - 1],
                     res=self.dx[cur], y=self.dx[1 - cur])
            cur = 1 - cur
        relu_bwd(self.dx[cur][:M], self.hs[0][:M], self.dz[:M])
        conv_bwd_w(a0, self.dz, g["w0"], self.part, B, self.shape, C_IN, W, 27, 1)
        bias_grad(self.dz, M, W, g["b0"], self.bpart)

    def loss_and_grad(self, target, mask, B: int) -> "object":
        """Masked MSE of self.pred against target (FP32 [M][5]); fills dpred.
        Returns the loss as a device scalar."""
        Z, Y, X = self.shape
        M = B * Z * Y * X
        cnt = float(mask[:M].sum().item()) * N_FIELDS
        inv = 1.0 / max(cnt, 1.0)
        nblk = (M * C_OUT + 255) // 256
        _chk(_bind().dsmc_no_mse(_p(self.pred), _p(target), _p(mask), M, C_OUT, N_FIELDS, inv, _p(self.dpred),
                                 _p(self.loss_part), _stream()), "mse")
        return self.loss_part[:nblk].sum()

    def adam(self, lr: float, b1=0.9, b2=0.999, eps=1e-8, wd=0.0) -> None:
        self.t += 1
        c1 = 1.0 - b1 ** self.t
        c2 = 1.0 - b2 ** self.t
        for k, w in self.params.items():
            _chk(_bind().dsmc_no_adam(_p(w), _p(self.grads[k]), _p(self.m1[k]), _p(self.m2[k]),
                                      _p(self.bf.get(k)), w.numel(), lr, b1, b2, eps, c1, c2,
                                      wd if k.startswith("w") else 0.0, _stream()), "adam")

    def train_step(self, a0, target, mask, B: int, lr: float):
        self.forward(a0, B)
        loss = self.loss_and_grad(target, mask, B)
        self.backward(a0, B)
        self.adam(lr)
        return loss

    def make_inputs(self, params: np.ndarray, spheres: np.ndarray, box) -> "object":
        """Operator inputs for B cases (params normalised, spheres (B, 4))."""
        torch = self.torch
        B = params.shape[0]
        Z, Y, X = self.shape
        out = torch.zeros(B * Z * Y * X, C_IN, dtype=torch.bfloat16, device="cuda")
        pt = torch.from_numpy(np.ascontiguousarray(params, np.float32)).cuda()
        st = torch.from_numpy(np.ascontiguousarray(spheres, np.float32)).cuda()
        bt = torch.from_numpy(np.asarray(box, np.float32)).cuda()
        inputs(pt, st, bt, B, self.shape, out)
        return out

    def predict(self, params: np.ndarray, spheres: np.ndarray, box) -> np.ndarray:
        """(B, Z, Y, X, 5) predicted operator outputs."""
        B = params.shape[0]
        out = []
        for b0 in range(0, B, self.max_batch):
            b1 = min(B, b0 + self.max_batch)
            a0 = self.make_inputs(params[b0:b1], spheres[b0:b1], box)
            pred = self.forward(a0, b1 - b0)
            out.append(pred[:, :N_FIELDS].float().cpu().numpy().reshape((b1 - b0,) + self.shape + (N_FIELDS,)))
        return np.concatenate(out)

    # -- persistence
    def state_dict(self) -> dict:
        d = {k: v.cpu().numpy() for k, v in self.params.items()}
        d["_meta"] = np.array([self.width, *self.shape, *self.dil], np.int64)
        return d

    def save(self, path) -> None:
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path, max_batch: int = 8) -> "FieldOperator":
        z = np.load(path, allow_pickle=False)
        meta = z["_meta"]
        op = cls(shape=tuple(meta[1:4]), width=int(meta[0]), dilations=tuple(meta[4:]), max_batch=max_batch)
        for k in op.params:
            op.params[k].copy_(op.torch.from_numpy(z[k]))
        op.sync_bf16()
        return op
