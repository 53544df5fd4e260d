"""The oracle's integrator, body force, wrap, diagnostics and DPD sigma
against closed forms and a thermostat invariant.

References: kick-drift src/scheme/move/dev/vv.h:1-10 (v += f/m dt, then
r += v dt); constant body force f += m a src/scheme/force/dev/main.h:1-4;
diagnostics src/io/diag/part/imp/main.h:37-63 (kBT = 0.5 sum|v|^2 / (1.5 N));
sigma = sqrt(2 kBT gamma / dt) src/pair/imp/main.h:54-62.
"""
import numpy as np
import pytest

import oracle as O


def _one(r, v):
    pp = np.zeros((1, 6), np.float32)
    pp[0, :3] = r
    pp[0, 3:] = v
    return pp


def test_kick_then_drift_from_rest():
    """From rest under a constant force the n-th position is the symplectic
    Euler sum (f/m) dt^2 n (n + 1) / 2; drift-then-kick would give
    n (n - 1) / 2, and the first step alone separates them (dt^2 vs 0)."""
    L, dt, m = (64, 64, 64), np.float32(1e-3), np.float32(2.0)
    f = np.array([[3.0, -1.5, 0.25]], np.float32)
    pp = _one((0.5, -0.25, 0.125), (0, 0, 0))
    r0 = pp[0, :3].astype(np.float64)
    a = f[0].astype(np.float64) / float(m)
    for n in range(1, 41):
        pp = O.integrate(L, pp, f, dt, mass=m, wrap=False)
        v_want = a * n * float(dt)
        r_want = r0 + a * float(dt) ** 2 * n * (n + 1) / 2
        assert np.allclose(pp[0, 3:], v_want, rtol=2e-6, atol=0)
        # float rounding of r (ulp 6e-8 near 0.5) accumulates one step at a time;
        # drift-then-kick would be off by n (f/m) dt^2 >= 1.2e-7 n
        assert np.allclose(pp[0, :3], r_want, rtol=0, atol=1e-7 + 3e-8 * n)
        if n == 1:
            # kick first: the position moved by (f/m) dt^2 already
            assert np.all(np.abs(pp[0, :3] - r0) > 0.5 * np.abs(a) * float(dt) ** 2)


def test_body_force_is_mass_times_acceleration():
    """f_total = f + m a: a force-free particle with body acceleration a moves
    exactly like one pushed by f = m a with no body force, and doubling the
    mass leaves its velocity change unchanged (it is a, not a/m)."""
    L, dt = (32, 32, 32), np.float32(5e-4)
    acc = np.array([0.01, 0.0, -0.02], np.float32)
    for m in (1.0, 2.0, 0.5):
        pp = _one((1.0, 2.0, -3.0), (0.1, -0.2, 0.3))
        zero = np.zeros((1, 3), np.float32)
        got = O.integrate(L, pp, zero, dt, mass=np.float32(m), accel=acc, wrap=False)
        want = O.integrate(L, pp, (np.float32(m) * acc)[None, :], dt, mass=np.float32(m), wrap=False)
        assert np.array_equal(got, want)
        dv = got[0, 3:].astype(np.float64) - pp[0, 3:].astype(np.float64)
        # within two float ulps of v (0.3): a / m instead of a would be off by >= 2.5e-6
        assert np.allclose(dv, acc.astype(np.float64) * float(dt), rtol=0, atol=6e-8)


def test_wrap_at_the_box_faces():
    """[-L/2, L/2): a drift that lands exactly on +L/2 maps to -L/2, one that
    lands exactly on -L/2 stays, crossings in either direction come back by L."""
    L, dt = (16, 16, 16), np.float32(1e-3)
    zero = np.zeros((1, 3), np.float32)
    # 7.5 + 500 * 0.001f rounds to exactly 8.0 in float
    pp = _one((7.5, -7.5, 0.0), (500.0, -500.0, 0.0))
    out = O.integrate(L, pp, zero, dt)
    assert out[0, 0] == np.float32(-8.0)
    assert out[0, 1] == np.float32(-8.0)
    # crossing both faces
    pp = _one((7.999, -7.999, 3.0), (2.0, -2.0, 0.0))
    out = O.integrate(L, pp, zero, dt)
    assert out[0, 0] == np.float32(np.float32(7.999) + np.float32(2.0) * dt - 16)
    assert out[0, 1] == np.float32(np.float32(-7.999) - np.float32(2.0) * dt + 16)
    assert -8 <= out[0, 0] < 8 and -8 <= out[0, 1] < 8
    # no wrap requested: the coordinate leaves the box
    out = O.integrate(L, _one((7.999, 0, 0), (2.0, 0, 0)), zero, dt, wrap=False)
    assert out[0, 0] > 8


def test_move_without_wrap_closed_form():
    """Membrane vertices: kick-drift with no body force and no wrap."""
    rng = np.random.default_rng(3)
    n, dt, m = 200, np.float32(1e-3), np.float32(1.5)
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-20, 20, (n, 3))
    pp[:, 3:] = rng.normal(0, 1, (n, 3))
    ff = rng.normal(0, 50, (n, 3)).astype(np.float32)
    out = O.move(pp, ff, dt, m)
    v1 = pp[:, 3:].astype(np.float64) + ff.astype(np.float64) / float(m) * float(dt)
    r1 = pp[:, :3].astype(np.float64) + v1 * float(dt)
    assert np.allclose(out[:, 3:], v1, rtol=0, atol=1e-6)
    assert np.allclose(out[:, :3], r1, rtol=0, atol=4e-6)
    assert np.any(np.abs(out[:, :3]) > 19)   # nothing wrapped into a box


def test_diag_hand_built():
    """Momentum, kBT = sum|v|^2 / (3 N) and max|v|^2 on velocities whose
    sums are exact."""
    pp = np.zeros((4, 6), np.float32)
    pp[:, 3:] = [(1, 0, 0), (-1, 0, 0), (0, 2, 0), (0, 0, -3)]
    d = O.diag(pp)
    assert list(d[:3]) == [0.0, 2.0, -3.0]
    assert d[3] == (1 + 1 + 4 + 9) / (3 * 4)
    assert d[4] == 9.0
    # a Maxwell set at kBT: the estimator is unbiased
    rng = np.random.default_rng(5)
    n = 200000
    pp = np.zeros((n, 6), np.float32)
    pp[:, 3:] = rng.normal(0, np.sqrt(0.0444), (n, 3))
    d = O.diag(pp)
    assert abs(d[3] / 0.0444 - 1) < 0.01
    assert d[4] == np.max(np.sum(pp[:, 3:].astype(np.float64) ** 2, axis=1))
    assert np.allclose(d[:3], pp[:, 3:].astype(np.float64).sum(0), rtol=0, atol=1e-9)
    assert O.diag(np.zeros((0, 6), np.float32))[3] == 0.0


@pytest.mark.parametrize("kBT,g,dt", [(1.0, 4.5, 1e-3), (0.0444, 8.0, 5e-4), (0.37, 6.25, 1e-3), (1.0, 0.0, 1e-3)])
def test_dpd_sigma_against_float64(kBT, g, dt):
    """sigma^2 dt = 2 kBT gamma (fluctuation-dissipation): float64 closed form,
    within the rounding of the three float operations and the root."""
    s = O.dpd_sigma(kBT, g, dt)
    want = np.sqrt(2.0 * np.float64(np.float32(kBT)) * np.float64(np.float32(g)) / np.float64(np.float32(dt)))
    assert abs(s - want) <= 4 * np.spacing(np.float32(want)) + 0.0


def test_thermostat_relaxes_to_kbt():
    """DPD thermostat invariant on the oracle's full step (cell list, forces
    with the branch RNG and sigma, kick-drift, wrap): a 16^3 box at density 3
    started at kBT = 0.5 from uniform random positions (the conservative
    potential heats it to 1.37 first) settles at the set kBT = 1 and keeps
    momentum zero (measured: 1.00-1.02 over steps 600-3000).  A sigma off by
    sqrt(2) (a dropped factor 2) or a wrong temperature estimator moves the
    plateau by a factor 2."""
    L = (16, 16, 16)
    n = 3 * 16 ** 3
    rng = np.random.default_rng(11)
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-8, 8, (n, 3))
    v = rng.normal(0, np.sqrt(0.5), (n, 3))
    pp[:, 3:] = v - v.mean(0)
    dt = 1e-3
    params = (25.0, 4.5, O.dpd_sigma(1.0, 4.5, dt), 0.25)
    seeds = O.seeds_table(900)
    temps = []
    for k in range(900):
        pp, _ = O.flu_step(L, pp, params, seeds[k], dt, nthreads=8)
        if k >= 600:
            temps.append(O.diag(pp)[3])
    t = float(np.mean(temps))
    assert 0.97 < t < 1.04, t
    # (total momentum is not a check here: the two sides of a wrapped pair in
    # the fragment pairs (-1,0,+1)/(+1,0,-1) draw independent noise, as in the
    # reference, so a periodic box random-walks its momentum)
