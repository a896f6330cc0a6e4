"""Pins the oracle's Mersenne Twister (CPU only).

* The generic block-form generator instantiated with the MT19937 parameters
  must reproduce numpy's MT19937 (Knuth seeding) and the C++ standard's
  10000th output of std::mt19937 (4123659995 for the default seed 5489).
* Instantiated with the reference's DCMT rows and seed chain it must equal
  genrand_mt of the reference's src/genmtrand.c (golden vectors written by
  tests/golden/make_golden.py from the compiled reference source).
* The engine's own parameter table rows must pass the primitivity check and
  the stream seeding must equal the reference's lcg chain in REFERENCE mode.
"""
from __future__ import annotations

import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle

MT19937 = dict(n=624, m=397, a=0x9908B0DF, lower=0x7FFFFFFF, u=11, s=7, b=0x9D2C5680, t=15, c=0xEFC60000, l=18)
ROOT = Path(__file__).resolve().parent.parent


def mt19937(seed: int, n: int) -> np.ndarray:
    p = MT19937
    return oracle.mt_generic(p["n"], p["m"], p["a"], p["lower"], p["u"], p["s"], p["b"], p["t"], p["c"], p["l"], seed, n)


def test_mt19937_standard_10000th_output():
    out = mt19937(5489, 10000)
    assert int(out[-1]) == 4123659995


def test_mt19937_matches_numpy():
    bg = np.random.MT19937()
    bg._legacy_seeding(20240601)
    ref = bg.random_raw(2000).astype(np.uint32)
    assert np.array_equal(mt19937(20240601, 2000), ref)


def test_dcmt_rows_match_reference_genrand(golden):
    d = np.load(golden / "mt_dcmt_rows.npz")
    rows, seeds, outs = d["rows"], d["seeds"], d["outputs"]
    got = oracle.mt_streams(rows[:, 0], rows[:, 1], rows[:, 2], seeds, outs.shape[1])
    assert np.array_equal(got.T, outs)


def test_planted_twist_mistake_is_caught(golden):
    # a generator with a wrong middle word offset must not reproduce the golden vectors
    d = np.load(golden / "mt_dcmt_rows.npz")
    rows, seeds, outs = d["rows"], d["seeds"], d["outputs"]
    wrong = oracle.mt_generic(19, 8, int(rows[0, 0]), 1, 12, 7, int(rows[0, 1]), 15, int(rows[0, 2]), 18,
                              int(seeds[0]), 64)
    assert not np.array_equal(wrong, outs[0, :64])


def test_reference_seed_chain(golden):
    # InitGPUTwisters (mtwister.cu:41-64): seed_i = lcg^(i+1)(999)
    d = np.load(golden / "mt_dcmt_rows.npz")
    s, ref = 999, []
    for _ in range(len(d["seeds"])):
        s = (s * 279470273) % 4294967291
        ref.append(s)
    assert np.array_equal(np.asarray(ref, np.uint32), d["seeds"])


def test_uniform_float_range():
    # GetRandomFloatFast maps u to ((float)u + 1)/2^32: in (0, 1]
    u = np.array([0, 1, 2**24, 2**32 - 129, 2**32 - 1], np.uint64)
    f = (u.astype(np.float32) + np.float32(1.0)) / np.float32(4294967296.0)
    assert f.min() > 0 and f.max() <= 1.0
    assert f[-1] == 1.0


def _table_rows() -> list[int]:
    text = (ROOT / "cudadsmc_amd" / "csrc" / "host" / "mt_params_table.inc").read_text()
    return [int(tok.rstrip("u,"), 16) for tok in text.split() if tok.startswith("0x")]


def test_builtin_table_is_primitive():
    rows = _table_rows()
    assert len(rows) == 1024 and len(set(rows)) == 1024
    assert all(r & 0x80000000 for r in rows)
    gen = ROOT / "cudadsmc_amd" / "csrc" / "host" / "mt_params_gen.cpp"
    with tempfile.TemporaryDirectory() as td:
        exe = Path(td) / "gen"
        subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", str(gen), "-o", str(exe)], check=True)
        sample = rows[:24] + rows[-8:]
        # three known-primitive DCMT rows of the reference data and two random twists
        known = [0xE98E0000, 0xF03E0001, 0x827AFFFF]
        bad = [0x9A3B1C2D, 0xDEADBEEF]
        args = [f"0x{x:08X}" for x in sample + known + bad]
        out = subprocess.run([str(exe), "--check", *args], check=True, capture_output=True, text=True).stdout.split()
        verdict = dict(zip(out[0::2], out[1::2]))
    assert all(verdict[f"0x{x:08X}"] == "1" for x in sample + known)
    assert all(verdict[f"0x{x:08X}"] == "0" for x in bad)


def test_stream_statistics():
    # uniformity and inter-stream correlation of the engine's streams (oracle)
    rows = _table_rows()
    ns = 256
    a = np.array([rows[i % len(rows)] for i in range(ns)], np.uint32)
    b = np.full(ns, 0x9D2C5680, np.uint32)
    c = np.full(ns, 0xEFC60000, np.uint32)
    seeds = np.arange(1, ns + 1, dtype=np.uint32)
    x = oracle.mt_streams(a, b, c, seeds, 4000).astype(np.float64) / 2**32
    assert abs(x.mean() - 0.5) < 3 * np.sqrt(1 / 12 / x.size) * 1.5
    hist, _ = np.histogram(x, bins=64, range=(0, 1))
    exp = x.size / 64
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 63 + 5 * np.sqrt(2 * 63)
    cc = np.corrcoef(x.T)
    off = cc[~np.eye(ns, dtype=bool)]
    assert np.abs(off).max() < 6 / np.sqrt(4000)


# ------------------------------------------------------------- Hybrid Taus
def test_taus_lcg_component_closed_form():
    """x_n = a^n x_0 + c (a^n - 1) / (a - 1) mod 2^32 (geometric series)."""
    a, c, x0 = 1664525, 1013904223, 2718281828
    got = oracle.taus_draws([200, 300, 400, x0], 1000, component=4)
    M = 2 ** 32
    for n in (1, 2, 3, 17, 999, 1000):
        an = pow(a, n, M * (a - 1))
        want = (an * x0 + c * ((an - 1) // (a - 1))) % M
        assert int(got[n - 1]) == want


@pytest.mark.parametrize("component,k", [(3, 28), (2, 29), (1, 31)])
def test_taus_component_has_maximal_period(component, k):
    """taus88 (L'Ecuyer 1996): each component is a maximal-period Tausworthe
    generator, period 2^k - 1 on its k significant bits; a wrong shift or mask
    makes the characteristic polynomial non-primitive and the period shorter."""
    for v in (128, 0xDEADBEEF):
        assert oracle.taus_period(component, v, 2 ** k) == 2 ** k - 1


def test_taus_combination_and_float():
    seed = [12345, 67890, 13579, 24680]
    comb = oracle.taus_draws(seed, 500)
    parts = [oracle.taus_draws(seed, 500, component=c) for c in (1, 2, 3, 4)]
    np.testing.assert_array_equal(comb, parts[0] ^ parts[1] ^ parts[2] ^ parts[3])
    assert len(set(comb.tolist())) == 500


def test_mt_window_export_matches_the_recurrence():
    """The oracle generates in Matsumoto-Nishimura's block form; the engine
    keeps the last 19 words x_d .. x_{d+18} of the recurrence
    x_{k+19} = x_{k+9} ^ twist(x_k, x_{k+1}). The exported window must be that
    for every d, including mid-block."""
    a, seed = 0xB40E0000 | 0x80000000, 4242
    x = [seed]
    for i in range(1, 19):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + i) & 0xFFFFFFFF)
    for k in range(80):
        y = (x[k] & 0xFFFFFFFE) | (x[k + 1] & 1)
        x.append(x[k + 9] ^ (y >> 1) ^ (a if y & 1 else 0))
    for d in list(range(0, 45)) + [57, 60]:
        w = oracle.mt_window_after(a, 0x9D2C5680, 0xEFC60000, seed, d)
        assert w[:19].tolist() == x[d:d + 19], d
        assert w[19] == 0


@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox4x32_10_known_answers(ctr, key, want):
    """Known-answer vectors of Philox4x32-10 (Salmon et al., SC'11; the
    published test vectors: zeros, all ones, and the digits of pi)."""
    assert oracle.philox(ctr, key).tolist() == list(want)
