"""Pins of the oracle's random number generators.

KISS trunk (src/math/rnd/imp.cpp:9-34): the reference's golden file
rnd.ref.rnd.t0 and the reference's own rnd_ini/rnd_get compiled from source.

Branch generator (src/math/rnd/dev.h:27-72): no golden file exists, so it is
pinned by (a) the RZ fma against the hardware rounding mode, (b) an
independent exact-rational restatement in Python, (c) distribution laws.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")


# ---------------------------------------------------------------- KISS trunk
def test_kiss_matches_golden_text():
    """rnd.t0: rnd_ini(42,42,42,42) then 100 x printf("%10.6e", rnd_get)."""
    want = open(os.path.join(GOLD, "rnd.ref.rnd.t0.txt")).read().split("\n")
    want = [w for w in want if w.strip()]
    got = ["%10.6e" % x for x in O.kiss_fill(42, 42, 42, 42, 100)]
    assert got == want


@pytest.mark.skipif(not O.ref_rnd_available(), reason="reference rnd not built (oracle/_ref)")
@pytest.mark.parametrize("seed", [(42, 42, 42, 42), (0, 0, 0, 0), (390 + 14, 14 + 615, 12309, 23094),
                                  (1908, 1409, 290, 12968), (-7, 123456789, 5, 77)])
def test_kiss_matches_reference_build(seed):
    got = O.kiss_fill(*seed, 5000)
    ref = O.ref_rnd_fill(*seed, 5000)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_kiss_integer_stream_known_answer():
    """With seed (0,0,0,0) only the LCG term moves: x_{k+1} = 69069 x_k + 12345,
    y stays 0 (xorshift of 0), z = c = 0 (MWC of 0)."""
    ints = O.kiss_ints(0, 0, 0, 0, 50)
    x = 0
    for k in range(50):
        x = (69069 * x + 12345) & 0xFFFFFFFF
        assert int(ints[k]) == x


# ---------------------------------------------------------------- RZ fma
def test_fma_rz_matches_hardware_mode():
    rng = np.random.default_rng(7)
    n = 200000
    a = rng.uniform(-8, 8, n).astype(np.float32)
    b = rng.uniform(0, 1, n).astype(np.float32)
    c = rng.choice(np.float32([-8, 1, -1, 0.5]), n).astype(np.float32)
    # the logistic map's own operands: (8, x^2, -8) and (t, x^2, 1)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    x2 = (x * x).astype(np.float32)
    a = np.concatenate([a, np.full(n, 8, np.float32), (8 * x2 - 8).astype(np.float32)])
    b = np.concatenate([b, x2, x2])
    c = np.concatenate([c, np.full(n, -8, np.float32), np.ones(n, np.float32)])
    e = O.fma_rz(a, b, c)
    h = O.fma_rz(a, b, c, fenv=True)
    assert np.array_equal(e.view(np.uint32), h.view(np.uint32))
    # RZ never exceeds the exact value in magnitude, and differs from RN by
    # at most one unit in the last place
    exact = a.astype(np.float64) * b + c
    assert np.all(np.abs(e.astype(np.float64)) <= np.abs(exact))
    rn = exact.astype(np.float32)
    assert np.all(np.abs(rn.view(np.int32).astype(np.int64) - e.view(np.int32)) <= 1)


# ------------------------------------------- exact-rational branch generator
def _round_f32(x, mode):
    """Round the rational x to float32: mode 'rn' (ties to even) or 'rz'."""
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    m = abs(x)
    e = m.numerator.bit_length() - m.denominator.bit_length()
    if Fraction(2) ** e > m:
        e -= 1
    assert e >= -126
    ulp = Fraction(2) ** (e - 23)
    q, r = divmod(m, ulp)
    q = int(q)
    if mode == "rn":
        if r > ulp / 2 or (r == ulp / 2 and q % 2 == 1):
            q += 1
    return s * q * ulp


def _f(x):
    return Fraction(float(np.float32(x)))


GOLD_F, SILVER_F = _f(0.6180339887498948482), _f(0.4142135623730950488)
BRONZE_F, TIN_F = _f(0.00008877875787352212838853023), _f(0.00004602357186447026756768986)
SQRT2_F = _f(1.41421356237309514547)


def _rational_mean0var1(seed, u, v):
    rn = lambda q: _round_f32(q, "rn")
    rz = lambda q: _round_f32(q, "rz")
    fu_lo, fu, fv_lo, fv = (rn(Fraction(int(t))) for t in (u & 1023, u, v & 1023, v))
    s = rn(fu_lo * GOLD_F + rn(fu * BRONZE_F))
    s = rn(fv_lo * SILVER_F + s)
    s = rn(fv * TIN_F + s)
    p = s - (s.numerator // s.denominator)        # s - floor(s), exact
    x = rn(_f(seed) - p)
    for _ in range(9):
        x2 = rn(x * x)
        x = rz(rz(8 * x2 - 8) * x2 + 1)
    return float(rn(x * SQRT2_F))


def test_branch_rng_matches_exact_rational_restatement():
    rng = np.random.default_rng(11)
    seeds = O.kiss_fill(0, 0, 0, 0, 8)
    for seed in seeds:
        u = rng.integers(0, 9_000_000, 40)
        v = rng.integers(0, 9_000_000, 40)
        got = O.rng_ii(seed, u, v)
        for k in range(40):
            want = np.float32(_rational_mean0var1(seed, int(u[k]), int(v[k])))
            assert got[k].view(np.uint32) == want.view(np.uint32), (seed, u[k], v[k])


def test_branch_rng_special_case_zero_indices():
    """u = v = 0 gives p = 0: the value is sqrt2 * T4^9(seed) with the
    reference's rounding, computed here from the rational restatement."""
    for seed in O.kiss_fill(42, 42, 42, 42, 10):
        got = O.rng_ii(seed, [0], [0])[0]
        assert got == np.float32(_rational_mean0var1(seed, 0, 0))


def test_branch_rng_ii_equals_uu_for_nonnegative():
    rng = np.random.default_rng(3)
    u = rng.integers(0, 2**24, 5000)
    v = rng.integers(0, 2**24, 5000)
    s = np.float32(0.3183)
    assert np.array_equal(O.rng_ii(s, u, v).view(np.uint32), O.rng_uu(s, u, v).view(np.uint32))


def test_branch_rng_distribution():
    """Arcsine law on [-sqrt2, sqrt2]: mean 0, variance 1, CDF
    F(x) = 1/2 + asin(x/sqrt2)/pi."""
    rng = np.random.default_rng(5)
    seeds = O.kiss_fill(0, 0, 0, 0, 4)
    vals = np.concatenate([O.rng_ii(s, rng.integers(1, 10**6, 50000), rng.integers(0, 10**6, 50000))
                           for s in seeds]).astype(np.float64)
    assert np.all(np.abs(vals) <= np.float32(np.sqrt(2)) * (1 + 1e-6))
    assert abs(vals.mean()) < 0.01
    assert abs(vals.var() - 1.0) < 0.01
    xs = np.linspace(-1.3, 1.3, 11)
    cdf = np.array([(vals <= x).mean() for x in xs])
    assert np.max(np.abs(cdf - (0.5 + np.arcsin(xs / np.sqrt(2)) / np.pi))) < 0.01


def test_seed_table_is_kiss_streams():
    """Per-step seeds: index 26 is the bulk trunk (0,0,0,0); fragment trunks
    (390+s, s+615, 12309, 23094) with s from the single-rank is_plus rule."""
    tab = O.seeds_table(20)
    assert np.array_equal(tab[:, 26], O.kiss_fill(0, 0, 0, 0, 20))
    # fragment 24 = (+1,0,0): v = (2,1,1) -> s = 2 + 3*(1 + 3) = 14
    assert np.array_equal(tab[:, 24], O.kiss_fill(390 + 14, 14 + 615, 12309, 23094, 20))
    # anti-fragment pairs share a trunk, except (-1,0,+1)/(+1,0,-1) whose
    # is_plus signs coincide (src/fluforces/imp/ini.h:11-17)
    for f in range(26):
        a = O.frag_anti(f)
        same = np.array_equal(tab[:, f], tab[:, a])
        assert same == (f not in (7, 15)), f
