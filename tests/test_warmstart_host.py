"""Host-side logic of the ensemble and warm-start tooling (CPU)."""
from __future__ import annotations

import math

import numpy as np

from cudadsmc_amd import ensemble as ens
from cudadsmc_amd import neural
from cudadsmc_amd import warmstart as ws


def test_case_draws_cover_the_ensemble_ranges():
    cases = ens.draw_cases(400, seed=3)
    U = np.array([c.U for c in cases])
    n = np.array([c.n for c in cases])
    r = np.array([c.r for c in cases])
    N = np.array([c.N for c in cases])
    assert U.min() >= 600 and U.max() <= 3000
    assert n.min() >= 3e19 and n.max() <= 1e21
    assert r.min() >= 0.015 and r.max() <= 0.035
    assert N.min() >= 1e4 and N.max() <= 5e5
    # log-uniform: the median sits near the geometric mean
    assert abs(math.log10(np.median(N)) - 0.5 * (4 + math.log10(5e5))) < 0.15
    assert [c.seed for c in cases[:3]] == [0, 1, 2]
    again = ens.draw_cases(400, seed=3)
    assert [c.U for c in again] == list(U)


def test_settings_of_a_case():
    c = ens.Case(U=1500.0, n=1e20, T_inf=200.0, T_w=300.0, r=0.025, N=100000, seed=7)
    s = ens.settings(c)
    assert tuple(s.grid_dim) == ens.GRID
    assert abs(s.cell_size[0] - 0.4 / 64) < 1e-9
    assert s.target_particles == 100000 and s.seed == 7 and s.mode == 1
    assert abs(c.cd_scale() - 0.5 * 1e20 * 6.63e-26 * 1500 ** 2 * math.pi * 0.025 ** 2) < 1e-12


def test_targets_round_trip():
    rng = np.random.default_rng(0)
    n = rng.uniform(1e19, 1e21, 100)
    u, v, w = rng.normal(0, 800, (3, 100))
    T = rng.uniform(100, 3000, 100)
    t = neural.fields_to_targets(n, u, v, w, T, 2e20, 1800.0, 220.0)
    back = neural.targets_to_fields(t, 2e20, 1800.0, 220.0)
    for a, b in zip((n, u, v, w, T), back):
        np.testing.assert_allclose(b, a, rtol=2e-6, atol=1e-6 * np.max(np.abs(a)))
    # cells without samples map to 0 (masked in training)
    z = neural.fields_to_targets(np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), 1e20, 1000.0, 200.0)
    assert np.all(z == 0.0)


def test_params_normalisation_is_order_one():
    cases = ens.draw_cases(200, seed=1)
    p, s = ws.case_inputs(cases)
    assert p.shape == (200, 5) and s.shape == (200, 4)
    assert np.all(np.abs(p) <= 1.6)


def test_field_error_and_block_means():
    X, Y, Z = ens.GRID
    rng = np.random.default_rng(2)
    ref = {k: rng.uniform(1.0, 2.0, X * Y * Z) for k in ("n", "u", "v", "w", "T")}
    frac = np.ones(X * Y * Z)
    assert ws.field_error(ref, ref, frac, 1.0, 2) == 0.0
    f = dict(ref)
    f["T"] = ref["T"] * 1.01
    assert abs(ws.field_error(f, ref, frac, 1.0, 2) - 0.01) < 1e-9
    # zero-mean noise within each 2^3 block cancels in the block average
    g = dict(ref)
    noise = np.zeros((Z, Y, X))
    noise[:, :, 0::2] = 0.1
    noise[:, :, 1::2] = -0.1
    g["n"] = ref["n"] + noise.ravel()
    bm = ws.block_mean(np.asarray(g["n"]).reshape(Z, Y, X), 2)
    np.testing.assert_allclose(bm, ws.block_mean(np.asarray(ref["n"]).reshape(Z, Y, X), 2), atol=1e-12)
    # blocks that are not fully fluid are excluded
    frac2 = frac.copy()
    frac2[:8] = 0.5
    h = dict(ref)
    hn = np.asarray(ref["n"]).copy()
    hn[:8] *= 10
    h["n"] = hn
    assert ws.field_error(h, ref, frac2, 1.0, 2) < ws.field_error(h, ref, frac, 1.0, 2)


def test_drag_windows_batch_means():
    c = ens.Case(U=1000.0, n=1e20, T_inf=200.0, T_w=300.0, r=0.02, N=1000, seed=0)

    class Fake:
        def __init__(self):
            self.fx = 0.0

        def drag(self):
            return np.array([self.fx, 0, 0, 0])

    d = ens.DragWindows(c, 1e-6)
    s = Fake()
    # windows with forces 1, 2, 3 (x cd_scale): the cumulative average is fed
    total = 0.0
    for k, f in enumerate((1.0, 2.0, 3.0), start=1):
        total += f * c.cd_scale() * 10
        s.fx = total / (10 * k)
        d.update(s, 10 * k)
    m, e = d.cd()
    np.testing.assert_allclose(d.means, [1.0, 2.0, 3.0], rtol=1e-12)
    assert abs(m - 2.0) < 1e-12 and abs(e - 1.0 / math.sqrt(3)) < 1e-12


def _half_window(rng, lam, u, th, steps):
    """Per-cell moment sums of `steps` sampled steps for particles with bulk
    velocity u (C, 3) and thermal spread th (3kT/m): count, sum v, sum |v|^2."""
    C_ = len(lam)
    n = rng.poisson(lam * steps).astype(np.float64)
    s1 = n[:, None] * u + rng.normal(size=(C_, 3)) * np.sqrt(n * th / 3.0)[:, None]
    # sum |v|^2 = n |u|^2 + 2 u . (sum c) + sum |c|^2, thermal part with chi-square noise
    sc = s1 - n[:, None] * u
    s2 = n * (u * u).sum(1) + 2 * (u * sc).sum(1) + th * (n + rng.normal(size=C_) * np.sqrt(2.0 * n / 3.0))
    return np.column_stack([n, s1, s2])


def test_block_moments_and_moment_fields():
    X, Y, Z = ens.GRID
    C_ = X * Y * Z
    st = np.zeros((C_, 5))
    st[:, 0] = 2.0
    st[:, 1] = 2.0 * 300.0  # u = 300
    st[:, 4] = 2.0 * (300.0 ** 2 + 3 * 1e4)  # 3kT/m = 3e4
    bm = ws.block_moments(st, 4)
    assert bm.shape == (X * Y * Z // 64, 5)
    np.testing.assert_allclose(bm[:, 0], 128.0)
    f = ws.moment_fields(bm, 16)
    np.testing.assert_allclose(f[:, 0], 8.0)
    np.testing.assert_allclose(f[:, 1], 300.0)
    np.testing.assert_allclose(f[:, 2], 3e4, rtol=1e-12)
    # a block where cells move differently: the spread of cell means counts as thermal
    st2 = st.copy()
    st2[0::2, 1] = 2.0 * 200.0
    st2[0::2, 4] = 2.0 * (200.0 ** 2 + 3e4)
    f2 = ws.moment_fields(ws.block_moments(st2, 4), 16)
    np.testing.assert_allclose(f2[:, 1], 250.0)
    np.testing.assert_allclose(f2[:, 2], 3e4 + 50.0 ** 2, rtol=1e-12)


def _run_detector(drift, seed, halves=12):
    X, Y, Z = ens.GRID
    C_ = X * Y * Z
    rng = np.random.default_rng(seed)
    lam = np.full(C_, 8.0)
    u0 = np.zeros((C_, 3))
    u0[:, 0] = 1500.0
    th = np.full(C_, 3 * 1.2e5)
    det = ws.SteadyDetector(np.ones(C_ // 64, bool), 4, 20)
    cum = np.zeros((C_, 5))
    fired = []
    for h in range(halves):
        u = u0 * (1.0 + drift * h)
        cum = cum + _half_window(rng, lam, u, th, 20)
        if det.feed(cum):
            fired.append(h)
    return fired, det


def test_steady_detector_fires_on_a_stationary_field():
    for seed in range(3):
        fired, det = _run_detector(0.0, seed)
        # the first test is at the end of the second chunk (half windows 0..3)
        assert fired and fired[0] == 3
        d_c, d_h = det.last
        # pure noise: chunk-to-chunk change ~ sqrt(1/2) of the half-to-half change
        np.testing.assert_allclose(d_c / d_h, np.sqrt(0.5), rtol=0.15)


def test_steady_detector_waits_while_the_field_drifts():
    # 2% speed drift per half window is far above the noise of a chunk
    fired, det = _run_detector(0.02, 0)
    assert fired == []
    d_c, d_h = det.last
    assert d_c[1] > 1.5 * d_h[1]


def test_steady_detector_waits_for_a_drift_in_one_field():
    # density drifting alone holds the start back (the test pools the fields)
    X, Y, Z = ens.GRID
    C_ = X * Y * Z
    rng = np.random.default_rng(5)
    u = np.zeros((C_, 3))
    u[:, 0] = 1500.0
    th = np.full(C_, 3 * 1.2e5)
    det = ws.SteadyDetector(np.ones(C_ // 64, bool), 4, 20)
    cum = np.zeros((C_, 5))
    fired = []
    for h in range(10):
        cum = cum + _half_window(rng, np.full(C_, 8.0 * (1 + 0.03 * h)), u, th, 20)
        fired.append(det.feed(cum))
    assert not any(fired)
    assert det.last[0][1] ** 2 <= det.last[1][1] ** 2  # the speed alone would have passed
