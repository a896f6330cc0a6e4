"""Accuracy of the deterministic math shared by kernels and oracle (CPU).

The engine evaluates sin/cos/log/exp2/pow with its own fixed-operation
implementations (Taylor/atanh series, Cody-Waite reductions) so GPU and CPU
agree bit for bit; here they are checked against float64 numpy (libm) with
error bounds in units of the last place, and on special values.
"""
from __future__ import annotations

import numpy as np

import oracle


def ulp_err(got: np.ndarray, ref: np.ndarray) -> np.ndarray:
    ref32 = ref.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_sincos_over_scattering_angles():
    x = np.linspace(1e-7, 2 * np.pi, 400001).astype(np.float32)
    s = oracle.detmath(0, x)
    c = oracle.detmath(1, x)
    xs = x.astype(np.float64)
    # absolute error relative to 1 (angles near zeros of sin/cos)
    assert np.max(np.abs(s - np.sin(xs))) < 2.5e-7
    assert np.max(np.abs(c - np.cos(xs))) < 2.5e-7
    big = np.abs(np.sin(xs)) > 0.1
    assert np.max(ulp_err(s[big], np.sin(xs[big]))) <= 4


def test_sincos_reduction_range():
    x = np.linspace(-3000.0, 3000.0, 200001).astype(np.float32)
    s = oracle.detmath(0, x)
    assert np.max(np.abs(s - np.sin(x.astype(np.float64)))) < 1e-6 * 3000 / 100


def test_log_uniform_inputs_and_special_values():
    u = np.arange(0, 2**32, 4093, dtype=np.uint64)
    x = ((u.astype(np.float32) + np.float32(1.0)) / np.float32(4294967296.0)).astype(np.float32)
    got = oracle.detmath(2, x)
    ref = np.log(x.astype(np.float64))
    assert np.max(ulp_err(got, ref)) <= 2.5
    x2 = np.geomspace(1e-37, 1e37, 100001).astype(np.float32)
    assert np.max(ulp_err(oracle.detmath(2, x2), np.log(x2.astype(np.float64)))) <= 2.5
    sp = oracle.detmath(2, np.array([1.0, 0.0, -1.0, np.inf], np.float32))
    assert sp[0] == 0.0 and np.isneginf(sp[1]) and np.isnan(sp[2]) and np.isposinf(sp[3])


def test_exp2_and_pow():
    x = np.linspace(-100, 100, 200001).astype(np.float32)
    got = oracle.detmath(3, x)
    assert np.max(ulp_err(got, np.exp2(x.astype(np.float64)))) <= 2
    cr = np.geomspace(1.0, 2e4, 50001).astype(np.float32)  # relative speeds [m/s]
    e = np.full_like(cr, 0.38)  # 2 - 2*omega for argon
    got = oracle.detmath(4, cr, e)
    assert np.max(ulp_err(got, cr.astype(np.float64) ** 0.38)) <= 8


def test_roundtrip_identities():
    x = np.linspace(0.01, 50, 10001).astype(np.float32)
    # exp2(log(x)/ln2) == x within a few ulp
    lg = oracle.detmath(2, x).astype(np.float64) / np.log(2)
    back = oracle.detmath(3, lg.astype(np.float32))
    assert np.max(ulp_err(back, x.astype(np.float64))) <= 8
