"""The wall oracle (oracle/udx_oracle_wall.c) against the reference's SDF
golden files, closed forms and brute force.

Fixtures: tests/golden/ref/sdf/{cyl,wall}.dat are the reference's
src/data/sdf/{cyl1/cyl.dat, wall1/wall.dat}; sdf.ref.sdf.0 and
wall.ref.sdf.1 are its src/test_data files for src/test/sdf/{0,1}.
"""
import os

import numpy as np
import pytest

import oracle as O
from oracle import wall as W

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")


def cyl():
    return W.read_sdf_dat(os.path.join(GOLD, "sdf", "cyl.dat"))


def test_sdf_eval_golden():
    """sdf.0: sdf_eval at (0, 0, z) of cyl.dat in a 16^3 box, printed %g."""
    ext, N, D = cyl()
    s = W.Sdf((16, 16, 16), ext, N, D)
    want = [l.split() for l in open(os.path.join(GOLD, "sdf.ref.sdf.0.txt")).read().split("\n") if l.strip()]
    for row in want:
        z = float(row[2])
        got = "%g" % s.eval([0.0, 0.0, z])[0]
        assert got == row[3], (z, got, row[3])


def test_sdf_resample_golden():
    """sdf.1: the scaled field B-spline-sampled at the 16^3 cell centres
    (the wall.h5 dump), printed %.10e."""
    ext, N, D = cyl()
    got = W.sample_box((16, 16, 16), ext, N, D)
    ref = np.loadtxt(os.path.join(GOLD, "wall.ref.sdf.1.txt"), skiprows=1)
    assert ref.size == got.size
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_texture_grid_size():
    """src/wall/sdf/imp/main.h:1-18 for the benchmark box and odd sizes."""
    for L, T in [((96, 96, 96), (256, 256, 256)), ((16, 16, 16), (256, 256, 256)), ((32, 16, 8), (256, 176, 128))]:
        ext = np.array(L, np.float32)
        s = W.Sdf(L, ext, (4, 4, 4), np.zeros(64, np.float32))
        assert s.T == T


def test_plane_channel_linear_precision():
    """A cubic B-spline reproduces linear functions and trilinear
    interpolation of a linear field is exact, so away from the kinks of
    |z| - h the texture gives |z| - h."""
    L = (24, 20, 32)
    Dn = W.plane_channel_sdf(L, 64, 10.0)
    s = W.Sdf(L, np.array(L, np.float32), (64, 64, 64), Dn)
    rng = np.random.default_rng(0)
    r = np.zeros((400, 3), np.float32)
    r[:, 0] = rng.uniform(-12, 12, 400)
    r[:, 1] = rng.uniform(-10, 10, 400)
    r[:, 2] = rng.choice([-1, 1], 400) * rng.uniform(2, 14, 400)
    got = s.eval(r)
    assert np.max(np.abs(got - (np.abs(r[:, 2]) - 10.0))) < 2e-4
    g = s.grad(r)
    assert np.allclose(g[:, 2], np.sign(r[:, 2]), atol=2e-3)
    # tangential components: FP32 rounding of the resampled texels divided by
    # the texel size
    assert np.allclose(g[:, :2], 0, atol=5e-4)


def channel(L=(16, 12, 24), h=8.0, rho=4, seed=1):
    Dn = W.plane_channel_sdf(L, 48, h)
    s = W.Sdf(L, np.array(L, np.float32), (48, 48, 48), Dn)
    rng = np.random.default_rng(seed)
    n = int(rho * np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    return s, pp


def test_freeze_labels_and_order():
    s, pp = channel()
    keep, wall = s.freeze(pp)
    sd = s.eval(pp[:, :3])
    assert np.array_equal(keep, pp[sd < 0])
    assert np.array_equal(wall, pp[(sd >= 0) & (sd <= 2)])


def test_wall_images_brute_force():
    L = (16, 12, 24)
    s, pp = channel(L)
    _, wall = s.freeze(pp)
    out = W.wall_images(L, wall)
    assert np.array_equal(out[:len(wall)], wall)
    lw = 0.5 * (np.array(L, np.float64) + 12)
    want = []
    for j in range(26):
        d = -np.array(O.frag_i2d3(j))
        for p in wall:
            r = p[:3] + (d * np.array(L)).astype(np.float32)
            if np.all((r.astype(np.float64) >= -lw) & (r.astype(np.float64) < lw)):
                want.append(np.concatenate([r, p[3:]]))
    assert np.array_equal(out[len(wall):], np.array(want, np.float32).reshape(-1, 6))


def wall_setup(L=(16, 12, 24), h=8.0, rho=4, seed=1):
    s, pp = channel(L, h, rho, seed)
    keep, wall = s.freeze(pp)
    allw = W.wall_images(L, wall)
    ws, wc, wsorted, _ = O.clist_build(W.margin_box(L), allw, sub=1)
    starts, counts, srt, _ = O.clist_build(L, keep)
    return s, srt, wsorted, ws


def test_wall_forces_brute_force():
    """Solvent-wall DPD == all pairs within the cutoff (wall velocity 0),
    with the noise mean0var1ii(seed, solvent index, wall index)."""
    L = (16, 12, 24)
    s, srt, wpp, ws = wall_setup(L)
    par = (25.0, 4.5, 3.0, 0.25)
    seed = np.float32(0.37)
    ff, npair = W.wall_forces(L, srt, s, wpp, ws, par, seed)
    far = s.far(srt[:, :3])
    want = np.zeros((len(srt), 3))
    count = 0
    rw = wpp[:, :3].astype(np.float64)
    for a in range(len(srt)):
        if far[a]:
            continue
        d = srt[a, :3].astype(np.float64) - rw
        r2 = (d * d).sum(1)
        for b in np.nonzero(r2 < 1)[0]:
            r = np.sqrt(r2[b])
            e = d[b] / r
            ev = e.dot(srt[a, 3:].astype(np.float64))
            xi = O.rng_ii(seed, [a], [b])[0]
            w = (1 - r) ** 0.25
            want[a] += (-par[1] * w * w * ev + par[2] * xi * w + par[0] * (1 - r)) * e
            count += 1
    assert npair == count and count > 0
    scale = np.sqrt((want ** 2).sum(1).mean())
    # float64 brute force vs the FP32 formula: w_R = (1-r)^(1/4) has an
    # unbounded slope at the cutoff, so a pair at r ~ 1 - 1e-6 turns the FP32
    # rounding of r into ~1e-4 of the noise term
    assert np.max(np.linalg.norm(ff - want, axis=1)) <= 1e-4 * scale
    # no particle far from the walls has a wall particle within the cutoff
    for a in np.nonzero(far)[0][:200]:
        assert np.min(((srt[a, :3] - wpp[:, :3]) ** 2).sum(1)) >= 1


def test_bounce_returns_particles_to_the_fluid():
    L = (16, 12, 24)
    s, _ = channel(L)
    rng = np.random.default_rng(4)
    n = 500
    pp = np.zeros((n, 6), np.float32)
    pp[:, 0] = rng.uniform(-8, 8, n)
    pp[:, 1] = rng.uniform(-6, 6, n)
    side = rng.choice([-1, 1], n)
    pp[:, 2] = side * rng.uniform(8.0, 8.05, n)       # just inside a wall
    pp[:, 5] = side * rng.uniform(0.5, 2.0, n)        # moving into it
    dt = 0.05
    out = s.bounce(pp, dt)
    assert np.all(s.eval(out[:, :3]) < 0)
    assert np.all(out[:, 5] * side < 0)               # normal velocity reversed
    assert np.array_equal(out[:, 3:], -pp[:, 3:])     # resting wall: v -> -v


def test_wvel_closed_forms():
    """Wall velocities (src/wall/wvel/dev.h:1-52) at hand-worked points."""
    r = np.array([[1.5, -2.0, 3.0], [3.0, 4.0, 0.0], [3.0, 4.0, 1.0], [3.0, 4.0, 2.5], [0.5, 0.0, -1.25]], np.float32)
    v = W.wvel_at(W.wvel("constant", u=(0.3, -0.2, 0.1)), r)
    assert np.array_equal(v, np.tile(np.float32([0.3, -0.2, 0.1]), (len(r), 1)))
    # shear: v[vdir] = gdot r[gdir]; half keeps r[gdir] <= 0
    v = W.wvel_at(W.wvel("shear", gdot=0.5, vdir=0, gdir=2), r)
    assert np.array_equal(v[:, 0], np.float32(0.5) * r[:, 2]) and not v[:, 1:].any()
    v = W.wvel_at(W.wvel("shear", gdot=0.5, vdir=1, gdir=0, half=1), r)
    assert np.array_equal(v[:, 1], np.float32(0.5) * np.minimum(r[:, 0], 0)) and not v[:, [0, 2]].any()
    assert not v[:, 1].any()                          # every x here is positive
    v = W.wvel_at(W.wvel("shear", gdot=-2.0, vdir=2, gdir=1, half=1), r)
    assert np.array_equal(v[:, 2], np.float32(-2.0) * np.minimum(r[:, 1], 0))
    # Hele-Shaw, u = 0.5, h = 4: radial u (1 - 4 z^2 / h^2) / |r_xy|^2, zero
    # from |z| = h / 2 on
    v = W.wvel_at(W.wvel("hele shaw", hu=0.5, hh=4.0), r)
    want = np.zeros((len(r), 3))
    for k, (x, y, z) in enumerate(r.astype(np.float64)):
        hfac = max(0.0, 0.5 * (1 - 4 * z * z / 16.0))
        want[k, :2] = hfac * np.array([x, y]) / (x * x + y * y)
    assert np.allclose(v, want, rtol=4e-7, atol=0)
    assert np.array_equal(v[1], np.float32([0.06, 0.08, 0])) and not v[3].any()
    assert not W.wvel_at(W.wvel("none"), r).any() and not W.wvel_at(None, r).any()


def test_wvel_shear_sin_rate():
    for gdot, w, t in [(0.5, 3.0, 0.4), (2.0, 0.1, 12.5), (-1.0, 7.0, 0.0)]:
        rate = W.wvel_shear_sin(gdot, w, t).gdot
        assert abs(rate - gdot * np.sin(np.float64(np.float32(w) * np.float32(t)))) <= 1e-6 * abs(gdot)


def moving_wall_particles(n, seed):
    rng = np.random.default_rng(seed)
    pp = np.zeros((n, 6), np.float32)
    pp[:, 0] = rng.uniform(-8, 8, n)
    pp[:, 1] = rng.uniform(-6, 6, n)
    side = rng.choice([-1, 1], n)
    pp[:, 2] = side * rng.uniform(8.0, 8.05, n)       # just inside a wall of the |z| < 8 channel
    pp[:, 3:5] = rng.normal(0, 1, (n, 2))
    pp[:, 5] = side * rng.uniform(0.5, 2.0, n)        # moving into it
    return pp, side


def test_bounce_off_moving_walls():
    """v' = 2 u_wall(r_w) - v at the crossing point r_w on the plane |z| = 8."""
    s, _ = channel((16, 12, 24))
    pp, side = moving_wall_particles(500, 6)
    dt = 0.05
    u = np.float32([0.3, -0.2, 0.1])
    out = s.bounce(pp, dt, W.wvel("constant", u=u))
    assert np.all(s.eval(out[:, :3]) < 0)
    assert np.allclose(out[:, 3:] + pp[:, 3:], 2 * u, rtol=0, atol=1e-6)
    # Couette walls v_x = gdot z: the lower wall moves at -8 gdot, the upper at +8 gdot
    gdot = np.float32(0.2)
    out = s.bounce(pp, dt, W.wvel("shear", gdot=gdot, vdir=0, gdir=2))
    assert np.all(s.eval(out[:, :3]) < 0)
    assert np.allclose(out[:, 3] + pp[:, 3], 2 * gdot * 8 * side, rtol=0, atol=2 * gdot * 0.01)
    assert np.array_equal(out[:, 4:], -pp[:, 4:])
    # half: only the lower (z < 0) wall moves
    out = s.bounce(pp, dt, W.wvel("shear", gdot=gdot, vdir=0, gdir=2, half=1))
    up = side > 0
    assert np.array_equal(out[up, 3:], -pp[up, 3:])
    assert np.allclose(out[~up, 3] + pp[~up, 3], -2 * gdot * 8, rtol=0, atol=2 * gdot * 0.01)


def test_wall_forces_galilean():
    """A constant wall velocity u enters only through v_solvent - u: the
    forces with walls at u equal those of solvent velocities v - u beside
    resting walls, bit for bit."""
    L = (16, 12, 24)
    s, srt, wpp, ws = wall_setup(L)
    par = (25.0, 4.5, 3.0, 0.25)
    seed = np.float32(0.37)
    u = np.float32([0.4, -0.3, 0.2])
    moved = srt.copy()
    moved[:, 3:] = srt[:, 3:] - u
    f_u, n_u = W.wall_forces(L, srt, s, wpp, ws, par, seed, wv=W.wvel("constant", u=u))
    f_0, n_0 = W.wall_forces(L, moved, s, wpp, ws, par, seed)
    assert n_u == n_0 > 0
    assert np.array_equal(f_u, f_0)
    f_r, _ = W.wall_forces(L, srt, s, wpp, ws, par, seed)
    assert not np.array_equal(f_u, f_r)
