"""Field-operator kernels (BF16 MFMA, FP32 accumulation) against the float64
oracle in oracle/no_ref.py, on BF16-representable inputs (GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import no_ref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bf(a: np.ndarray):
    """BF16 device tensor holding exactly bf16_round(a)."""
    return torch.from_numpy(no_ref.bf16_round(np.asarray(a, np.float32))).cuda().to(torch.bfloat16)


def f32(a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t) -> np.ndarray:
    torch.cuda.synchronize()
    return t.float().cpu().numpy()


SHAPE = (6, 5, 9)  # Z, Y, X: 270 voxels per case, not a multiple of the 128-voxel block


def rel_err(got, want):
    return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-30))


@pytest.mark.parametrize("cin,cout,taps,dil", [(16, 32, 27, 1), (32, 96, 27, 2), (96, 96, 27, 3), (96, 32, 1, 1),
                                               (32, 128, 27, 1)])
def test_conv_fwd_matches_oracle(cin, cout, taps, dil):
    from cudadsmc_amd import neural

    print(f"conv_fwd cin={cin} cout={cout} taps={taps} dil={dil}", flush=True)
    rng = np.random.default_rng(cin + cout + taps + dil)
    B = 2
    x = no_ref.bf16_round(rng.standard_normal((B,) + SHAPE + (cin,)).astype(np.float32))
    w = no_ref.bf16_round((rng.standard_normal((taps, cout, cin)) / np.sqrt(taps * cin)).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    M = B * int(np.prod(SHAPE))
    y32 = torch.zeros(M, cout, dtype=torch.float32, device="cuda")
    neural.conv_fwd(bf(x).reshape(M, cin), bf(w), f32(b), B, SHAPE, cin, cout, taps, dil, y32=y32)
    want = no_ref.conv(x, w, b, dil).reshape(M, cout)
    assert rel_err(host(y32), want) < 2e-5


def test_conv_fwd_relu_residual_outputs():
    from cudadsmc_amd import neural

    rng = np.random.default_rng(5)
    B, cin, cout = 2, 32, 32
    M = B * int(np.prod(SHAPE))
    x = no_ref.bf16_round(rng.standard_normal((B,) + SHAPE + (cin,)).astype(np.float32))
    w = no_ref.bf16_round((rng.standard_normal((27, cout, cin)) * 0.05).astype(np.float32))
    b = (rng.standard_normal(cout) * 0.1).astype(np.float32)
    res = no_ref.bf16_round(rng.standard_normal((M, cout)).astype(np.float32))
    y = torch.zeros(M, cout, dtype=torch.bfloat16, device="cuda")
    h = torch.zeros_like(y)
    neural.conv_fwd(bf(x).reshape(M, cin), bf(w), f32(b), B, SHAPE, cin, cout, 27, 2, relu=True, res=bf(res), y=y,
                    h=h)
    hw = np.maximum(no_ref.conv(x, w, b, 2).reshape(M, cout), 0.0)
    # BF16 outputs: within one BF16 rounding (2^-8 relative) of the exact values
    assert np.all(np.abs(host(h) - hw) <= 2.0 ** -8 * np.abs(hw) + 1e-6)
    # y = BF16(v + res) with v the FP32 activation: one BF16 rounding of the
    # exact sum, plus the FP32 accumulation error of v
    yw = hw + res.astype(np.float64)
    assert np.all(np.abs(host(y) - yw) <= 2.0 ** -8 * np.abs(yw) + 1e-5 * (np.abs(hw) + np.abs(res)) + 1e-6)


@pytest.mark.parametrize("cin,cout,taps,dil,chunk", [(32, 96, 27, 1, 64), (96, 96, 27, 2, 128), (96, 32, 1, 1, 4096),
                                                     (32, 32, 27, 4, 192)])
def test_conv_bwd_w_matches_oracle(cin, cout, taps, dil, chunk):
    from cudadsmc_amd import neural

    print(f"conv_bwd_w cin={cin} cout={cout} taps={taps} dil={dil} chunk={chunk}", flush=True)
    rng = np.random.default_rng(100 + cin + dil)
    B = 2
    M = B * int(np.prod(SHAPE))
    x = no_ref.bf16_round(rng.standard_normal((B,) + SHAPE + (cin,)).astype(np.float32))
    dy = no_ref.bf16_round(rng.standard_normal((B,) + SHAPE + (cout,)).astype(np.float32))
    chunks = (M + chunk - 1) // chunk
    part = torch.zeros(chunks * taps * cout * cin, dtype=torch.float32, device="cuda")
    dw = torch.full((taps, cout, cin), 7.0, dtype=torch.float32, device="cuda")
    neural.conv_bwd_w(bf(x).reshape(M, cin), bf(dy).reshape(M, cout), dw, part, B, SHAPE, cin, cout, taps, dil,
                      chunk=chunk)
    want = no_ref.conv_grad_w(x, dy, taps, dil)
    got = host(dw)
    assert rel_err(got, want) < 2e-5
    # accumulate mode adds to what is there
    neural.conv_bwd_w(bf(x).reshape(M, cin), bf(dy).reshape(M, cout), dw, part, B, SHAPE, cin, cout, taps, dil,
                      chunk=chunk, accumulate=True)
    assert rel_err(host(dw), 2 * want) < 2e-5


@pytest.mark.parametrize("cin,cout,taps,dil", [(32, 96, 27, 1), (96, 96, 27, 2), (96, 32, 1, 1)])
def test_data_gradient_through_flipped_weights(cin, cout, taps, dil):
    from cudadsmc_amd import neural

    rng = np.random.default_rng(200 + cin + dil)
    B = 2
    M = B * int(np.prod(SHAPE))
    w = no_ref.bf16_round((rng.standard_normal((taps, cout, cin)) * 0.1).astype(np.float32))
    dy = no_ref.bf16_round(rng.standard_normal((B,) + SHAPE + (cout,)).astype(np.float32))
    wb = torch.zeros(taps, cin, cout, dtype=torch.bfloat16, device="cuda")
    neural.weight_flip(bf(w), wb, taps, cout, cin)
    dx = torch.zeros(M, cin, dtype=torch.float32, device="cuda")
    neural.conv_fwd(bf(dy).reshape(M, cout), wb, None, B, SHAPE, cout, cin, taps, dil, y32=dx)
    want = no_ref.conv_grad_x(dy, w, dil).reshape(M, cin)
    assert rel_err(host(dx), want) < 2e-5


def test_bias_grad_and_relu_bwd():
    from cudadsmc_amd import neural

    rng = np.random.default_rng(9)
    M, Cc = 9000, 96
    dy = no_ref.bf16_round(rng.standard_normal((M, Cc)).astype(np.float32))
    h = no_ref.bf16_round(rng.standard_normal((M, Cc)).astype(np.float32))
    part = torch.zeros(((M + 511) // 512) * Cc, dtype=torch.float32, device="cuda")  # one row per 512 voxels
    db = torch.zeros(Cc, dtype=torch.float32, device="cuda")
    neural.bias_grad(bf(dy), M, Cc, db, part)
    np.testing.assert_allclose(host(db), dy.astype(np.float64).sum(0), rtol=1e-5, atol=1e-3)
    dz = torch.zeros(M, Cc, dtype=torch.bfloat16, device="cuda")
    neural.relu_bwd(bf(dy), bf(h), dz)
    np.testing.assert_array_equal(host(dz), np.where(h > 0, dy, 0.0))


def test_inputs_kernel():
    from cudadsmc_amd import neural

    rng = np.random.default_rng(4)
    B = 3
    p = rng.standard_normal((B, 5)).astype(np.float32)
    sph = np.array([[0.1, 0.1, 0.1, 0.025], [0.12, 0.09, 0.1, 0.03], [0.08, 0.1, 0.11, 0.015]], np.float32)
    box = np.array([0.0, 0.0, 0.0, 0.4, 0.2, 0.2], np.float32)
    out = torch.zeros(B * int(np.prod(SHAPE)), 32, dtype=torch.bfloat16, device="cuda")
    neural.inputs(f32(p), f32(sph), f32(box), B, SHAPE, out)
    want = no_ref.inputs(p, sph, box, SHAPE, 32).reshape(-1, 32)
    got = host(out)
    assert np.all(np.abs(got - want) <= 2.0 ** -8 * np.abs(want) + 1e-6)


def test_network_gradients_match_oracle():
    """One training step's gradients of a small operator against the float64
    oracle network with the same (BF16) weights and inputs. The GPU keeps
    activations and gradients in BF16 between layers, so the comparison is
    by relative norm error and cosine per parameter."""
    from cudadsmc_amd import neural

    op = neural.FieldOperator(shape=SHAPE, width=32, dilations=(1, 2), seed=3, max_batch=2)
    rng = np.random.default_rng(1)
    B = 2
    M = B * int(np.prod(SHAPE))
    a0 = no_ref.bf16_round(rng.standard_normal((B,) + SHAPE + (32,)).astype(np.float32) * 0.5)
    target = rng.standard_normal((B,) + SHAPE + (5,)).astype(np.float32) * 0.3
    mask = (rng.random((B,) + SHAPE) > 0.2).astype(np.float32)
    params = {}
    for k, v in op.params.items():
        a = v.cpu().numpy()
        params[k] = no_ref.bf16_round(a) if k.startswith("w") else a
    net = no_ref.Net(params, [1, 2])
    pred, cache = net.forward(a0.astype(np.float64))
    loss_w, dpred = no_ref.Net.loss(pred, target, mask, 5)
    g = net.backward(cache, dpred)

    ta0 = bf(a0).reshape(M, 32)
    op.forward(ta0, B)
    got_pred = host(op.pred[:M]).reshape(pred.shape)
    assert rel_err(got_pred, pred) < 3e-2
    loss = op.loss_and_grad(f32(target.reshape(M, 5)), f32(mask.reshape(M)), B)
    assert abs(float(loss.item()) - loss_w) <= 3e-2 * loss_w
    op.backward(ta0, B)
    for k in params:
        gg = host(op.grads[k]).ravel().astype(np.float64)
        gw = g[k].ravel()
        err = np.linalg.norm(gg - gw) / max(np.linalg.norm(gw), 1e-30)
        cos = gg @ gw / max(np.linalg.norm(gg) * np.linalg.norm(gw), 1e-30)
        print(f"{k}: rel {err:.2e} cos {cos:.6f}", flush=True)
        assert err < 5e-2 and cos > 0.998, k


def test_adam_step():
    from cudadsmc_amd import neural

    op = neural.FieldOperator(shape=SHAPE, width=32, dilations=(1,), seed=0, max_batch=1)
    rng = np.random.default_rng(2)
    w0 = {k: v.cpu().numpy().astype(np.float64) for k, v in op.params.items()}
    gs = {k: rng.standard_normal(v.shape).astype(np.float32) for k, v in op.params.items()}
    for k, g in gs.items():
        op.grads[k].copy_(torch.from_numpy(g))
    op.adam(1e-2, wd=1e-4)
    op.adam(1e-2, wd=1e-4)
    for k in w0:
        w, m1, m2 = w0[k].copy(), 0.0, 0.0
        for t in (1, 2):
            gi = gs[k] + (1e-4 * w if k.startswith("w") else 0.0)
            m1 = 0.9 * m1 + 0.1 * gi
            m2 = 0.999 * m2 + 0.001 * gi * gi
            w = w - 1e-2 * (m1 / (1 - 0.9 ** t)) / (np.sqrt(m2 / (1 - 0.999 ** t)) + 1e-8)
        np.testing.assert_allclose(host(op.params[k]), w, rtol=1e-5, atol=1e-6, err_msg=k)
        if k in op.bf:
            np.testing.assert_array_equal(host(op.bf[k]), no_ref.bf16_round(host(op.params[k])))


def test_training_reduces_loss():
    from cudadsmc_amd import neural

    op = neural.FieldOperator(shape=(8, 8, 16), width=32, dilations=(1, 2), seed=0, max_batch=4)
    rng = np.random.default_rng(0)
    B = 4
    p = rng.uniform(-1, 1, (B, 5)).astype(np.float32)
    sph = np.tile(np.array([[0.1, 0.1, 0.1, 0.03]], np.float32), (B, 1))
    box = (0.0, 0.0, 0.0, 0.4, 0.2, 0.2)
    a0 = op.make_inputs(p, sph, box)
    feats = host(a0).reshape(B, 8, 8, 16, 32)
    # a smooth target that depends on position and the case parameters
    tgt = np.stack([np.tanh(feats[..., 0] * feats[..., 4]), feats[..., 1] * feats[..., 5], np.exp(-feats[..., 3]),
                    feats[..., 6] * 0.5, np.sin(feats[..., 2] * 2)], -1).astype(np.float32)
    M = B * 8 * 8 * 16
    t_tgt, t_mask = f32(tgt.reshape(M, 5)), f32(np.ones(M))
    first = None
    for it in range(150):
        loss = float(op.train_step(a0, t_tgt, t_mask, B, 3e-3).item())
        first = loss if first is None else first
    print(f"loss {first:.4f} -> {loss:.4f}", flush=True)
    assert np.isfinite(loss) and loss < 0.2 * first
