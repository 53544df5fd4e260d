"""The oracle's exact-sum forces (the production kernel's definition).

Each particle's force is the exact sum of its float pair forces, each
component rounded to a multiple of 2^-32, rounded to float once
(orc_dpd_forces_sum64).  Pinned by:
- the pair force being antisymmetric bit for bit, so evaluating a pair once
  (half shell) or from both sides (gather) gives the same terms;
- the half-shell traversal with reactions (orc_dpd_forces_half_sum64) and the
  full-shell gather giving identical bits, which holds only if every sum is
  exact (any rounding would depend on the order);
- the result being the float nearest to the double sum of the same float
  pair forces (orc_dpd_forces_f64), within the quantisation of pair forces
  below 2^-9;
- the float-order gather of the reference (orc_dpd_forces) staying within
  float summation error of that double sum.
"""
import numpy as np
import pytest

import oracle as O

CASES = [
    ((16, 16, 16), 3, 75.0, 4.5, 1.0, 1e-3, 0.25),
    ((7, 5, 3), 5, 25.0, 4.5, 0.5, 1e-3, 1.0),
    ((3, 3, 3), 4, 25.0, 4.5, 1.0, 1e-3, 0.5),
    ((12, 10, 9), 10, 2.5, 8.0, 0.0444, 5e-4, 0.25),
]


def random_box(L, rho, seed):
    rng = np.random.default_rng(seed)
    n = int(rho * np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 1.0, (n, 3)).astype(np.float32)
    return pp


def test_quantisation_keeps_forces_above_2m9():
    rng = np.random.default_rng(1)
    f = (rng.uniform(-1, 1, 100000) * 10.0 ** rng.uniform(-2.7, 3, 100000)).astype(np.float32)
    f = f[np.abs(f) >= 2.0 ** -9]
    q = np.rint(f * np.float32(2.0 ** 32)) * np.float32(2.0 ** -32)
    assert np.array_equal(q.view(np.uint32), f.view(np.uint32))


def test_pair_force_antisymmetric_bits():
    rng = np.random.default_rng(2)
    par = (25.0, 4.5, O.dpd_sigma(1.0, 4.5, 1e-3), 0.25)
    for k in range(2000):
        a = rng.normal(0, 1, 6).astype(np.float32)
        d = rng.normal(0, 1, 3)
        d *= rng.uniform(0, 1) / np.linalg.norm(d)   # inside the cutoff, near it too
        b = a.copy()
        b[:3] = (a[:3] - d).astype(np.float32)
        b[3:] = rng.normal(0, 1, 3).astype(np.float32)
        rnd = np.float32(rng.normal())
        fab, hit = O.pair_force(par, a, b, rnd)
        fba, hit2 = O.pair_force(par, b, a, rnd)
        assert hit == hit2
        # bit for bit up to the sign of zero terms, which exact sums ignore
        fab, mba = fab + np.float32(0), -fba + np.float32(0)
        assert np.array_equal(fab.view(np.uint32), mba.view(np.uint32)), k


@pytest.mark.parametrize("case", CASES)
def test_half_shell_equals_gather_exactly(case):
    L, rho, a, g, kBT, dt, spow = case
    pp = random_box(L, rho, 3)
    starts, counts, srt, _ = O.clist_build(L, pp)
    seeds = O.seeds_table(2)[1]
    par = (a, g, O.dpd_sigma(kBT, g, dt), spow)
    f_g, np_g = O.dpd_forces_sum64(L, srt, starts, counts, par, seeds, nthreads=4)
    f_h, np_h = O.dpd_forces_half_sum64(L, srt, starts, counts, par, seeds, nthreads=3)
    assert np_g == np_h
    assert np.array_equal(f_g.view(np.uint32), f_h.view(np.uint32))
    # the float half shell of the reference's order differs (rounding)
    f_r, _ = O.dpd_forces_half(L, srt, starts, counts, par, seeds, nthreads=3)
    assert not np.array_equal(f_r.view(np.uint32), f_h.view(np.uint32))


@pytest.mark.parametrize("case", CASES)
def test_sum64_is_the_rounded_double_sum(case):
    L, rho, a, g, kBT, dt, spow = case
    pp = random_box(L, rho, 4)
    starts, counts, srt, _ = O.clist_build(L, pp)
    seeds = O.seeds_table(1)[0]
    par = (a, g, O.dpd_sigma(kBT, g, dt), spow)
    f_q, np_q = O.dpd_forces_sum64(L, srt, starts, counts, par, seeds, nthreads=4)
    f_d, np_d = O.dpd_forces_f64(L, srt, starts, counts, par, seeds, nthreads=4)
    assert np_q == np_d
    # nearest float to the double sum, up to 2^-33 per quantised pair term
    ulp = np.spacing(np.abs(f_q)).astype(np.float64)
    maxpairs = 200   # above any particle's pair count at these densities
    err = np.abs(f_q.astype(np.float64) - f_d)
    assert np.all(err <= 0.5 * ulp + maxpairs * 2.0 ** -33), err.max()
    # the float-order gather is within float summation error of the same sum
    f_f, _ = O.dpd_forces(L, srt, starts, counts, par, seeds, nthreads=4)
    tot = np.abs(f_d).max()
    assert np.abs(f_f.astype(np.float64) - f_d).max() <= 64 * np.spacing(np.float32(tot))


def test_planted_sign_error_is_caught():
    """A reaction added with the wrong sign (a plausible half-shell mistake)
    breaks the equality with the gather: restated here in float64 numpy on a
    small box from the oracle's own pair forces."""
    L, rho = (4, 4, 4), 3
    par = (25.0, 4.5, O.dpd_sigma(1.0, 4.5, 1e-3), 0.25)
    pp = random_box(L, rho, 5)
    starts, counts, srt, _ = O.clist_build(L, pp)
    seeds = O.seeds_table(1)[0]
    f_g, _ = O.dpd_forces_f64(L, srt, starts, counts, par, seeds, periodic=False)
    n = len(srt)
    good = np.zeros((n, 3))
    bad = np.zeros((n, 3))
    h = np.array(L) / 2.0
    cell = np.floor(srt[:, :3].astype(np.float64) + h).astype(int)
    for i in range(n):
        for j in range(i):
            if np.any(np.abs(cell[i] - cell[j]) > 1):
                continue
            f, hit = O.pair_force(par, srt[i], srt[j], float(O.rng_ii(seeds[26], [i], [j])[0]))
            if not hit:
                continue
            good[i] += f
            good[j] -= f
            bad[i] += f
            bad[j] += f
    assert np.allclose(good, f_g, rtol=0, atol=1e-9)
    assert not np.allclose(bad, f_g, rtol=0, atol=1e-3)
