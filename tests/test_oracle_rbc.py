"""The RBC membrane oracle (oracle/udx_oracle_rbc.c) against the reference's
golden files, the energy it is the gradient of, and mesh invariants.

Fixtures: tests/golden/ref/cells/{498,642,1986}.off are the reference's
src/data/cells/rbc meshes (its sc/642.off, the stress-free shape, is the same
file as 642.off); av.ref.rbc.u.area_volume.dev.0 and
forces.ref.rbc.u.{simple,stree.free} are its src/test_data files for
src/test/rbc/u/{area_volume/dev,simple,stress.free}, run with
conf/rbc.test.cfg (tests/golden/ref/conf/rbc.test.cfg), the initial condition
"two" (identity and a shift by (1, 2, 3)) in a 16^3 box, positions and forces
printed %g after scaling by 0.1 (simple) or 0.025 (stress free).
"""
import os

import numpy as np
import pytest

from oracle import rbc as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")


def mesh(nv):
    return R.Mesh(*R.read_off(os.path.join(GOLD, "cells", "%d.off" % nv)))


def two_cells(m, L=(16, 16, 16)):
    T = np.eye(4)
    T[:3, 3] = [1, 2, 3]
    return R.place(m.rr, [np.eye(4), T], L)


# totArea / totVolume of the 642 template as the test scripts pass them
# (off.area / off.volume, printed with 6 digits)
A642, V642 = 80.0594, 43.3586


def test_area_volume_golden():
    """av.0: 1986.off with the identity initial condition in a 20^3 box."""
    m = mesh(1986)
    pp = R.place(m.rr, [np.eye(4)], (20, 20, 20))
    A, V = R.area_volume(m, pp, 1)
    want = [float(x) for x in open(os.path.join(GOLD, "av.ref.rbc.u.area_volume.dev.0.txt")).read().split()]
    assert "%g" % A == "%g" % want[0] and abs(V - want[1]) < 0.05
    assert abs(A - want[0]) < 1e-3 and abs(V / want[1] - 1) < 1e-5


def test_area_volume_closed_form():
    """Area and volume of an icosahedron of circumradius 1, translated away
    from the origin (the volume formula is origin independent for a closed
    oriented surface)."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v[0])
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
                  [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int32)
    m = R.Mesh(v.astype(np.float32), f)
    pp = np.zeros((12, 6), np.float32)
    pp[:, :3] = (v + [3.0, -2.0, 1.5]).astype(np.float32)
    A, V = R.area_volume(m, pp, 1)
    a = 1 / np.sin(2 * np.pi / 5)        # edge of the unit-circumradius icosahedron
    assert abs(A - 5 * 3 ** 0.5 * a * a) < 1e-5
    assert abs(V - 5 * (3 + 5 ** 0.5) / 12 * a ** 3) < 1e-5


def test_adjacency_rings():
    """Each vertex's ring lists its neighbours in face order: (v, n_k,
    n_k+1) is a face, adj1 is the vertex across the edge (n_k, n_k+1), the
    ring starts at the smallest neighbour and closes."""
    for nv in (498, 642):
        m = mesh(nv)
        faces = {tuple(np.roll(t, k)) for t in m.tt for k in range(3)}
        deg = np.bincount(m.tt.ravel(), minlength=m.nv)
        assert m.md == deg.max()
        for v in range(m.nv):
            ring = m.adj0[v * m.md: v * m.md + deg[v]]
            assert np.all(m.adj0[v * m.md + deg[v]: (v + 1) * m.md] == -1)
            assert ring[0] == ring.min() and len(set(ring)) == deg[v]
            for k in range(deg[v]):
                a, b = ring[k], ring[(k + 1) % deg[v]]
                assert (v, a, b) in faces
                assert (b, a, m.adj1[v * m.md + k]) in faces
            mm = m.map(v * m.md)
            assert mm[1] == ring[0] and mm[2] == ring[1] and mm[3] == ring[2 % deg[v]]


def minsin_per_vertex(m, pp):
    """smallest |sin theta| over the dihedrals each vertex takes part in"""
    P = pp[:, :3].astype(np.float64)
    out = np.full(len(P), 9.0)

    def sin_t(a, b, c, d):
        k, n = np.cross(a - b, a - c), np.cross(b - d, b - c)
        cs = k.dot(n) / np.sqrt(k.dot(k) * n.dot(n))
        return np.sqrt(max(1 - cs * cs, 0.0))
    for v in range(len(P)):
        for j in range(m.md):
            mm = m.map(v * m.md + j)
            if mm is None:
                continue
            i0, i1, i2, i3, i4 = mm[:5]
            s1, s2 = sin_t(P[i0], P[i2], P[i1], P[i4]), sin_t(P[i1], P[i0], P[i2], P[i3])
            for w in (i0, i1, i2, i4):
                out[w] = min(out[w], s1)
            for w in (i0, i1, i2, i3):
                out[w] = min(out[w], s2)
    return out


@pytest.mark.parametrize("name,scale,stress_free", [("simple", 0.1, False), ("stree.free", 0.025, True)])
def test_forces_golden(name, scale, stress_free):
    """Membrane + Kantor forces of two 642-vertex cells.  Where a vertex takes
    part in a nearly flat dihedral (|sin theta| < 0.05; this mesh has some
    down to 1e-3) the reference's own result is ill-conditioned: its
    1/sin(theta) changes sign with the rounding of the positions, and its
    two identical cells differ there by up to 0.08 in the printed units.
    Elsewhere the oracle must match the printed forces to their 6 digits."""
    m = mesh(642)
    pp = two_cells(m)
    ref = np.loadtxt(os.path.join(GOLD, "forces.ref.rbc.u.%s.txt" % name))
    assert np.abs(ref[:, :3] - scale * pp[:, :3]).max() < 1e-5
    p = R.par(**R.RBC_TEST, totArea=A642, totVolume=V642)
    av = R.area_volume(m, pp, 2)
    if stress_free:
        ff = R.forces(p, m, pp, 2, av, stress_free=m.shape())
    else:
        l0, a0 = m.stressful(A642)
        ff = R.forces(p, m, pp, 2, av, l0=l0, a0=a0)
    err = np.abs(ref[:, 3:] - scale * ff).max(1)
    well = minsin_per_vertex(m, pp) > 0.05
    assert well.sum() > 400
    fmax = np.abs(ref[:, 3:]).max()
    assert err[well].max() <= 2e-4 * max(fmax, 1.0)
    # the ill-conditioned rest: no larger than the reference's own spread
    selfdiff = np.abs(ref[:642, 3:] - ref[642:, 3:]).max()
    assert err.max() <= 1.5 * selfdiff + 1e-3


def perturbed(nv, sigma, seed):
    m = mesh(nv)
    rng = np.random.default_rng(seed)
    pp = np.zeros((nv, 6), np.float32)
    pp[:, :3] = (m.rr + rng.normal(0, sigma, (nv, 3))).astype(np.float32)
    return m, pp


TERMS = {
    # term bit: parameters that switch on only this term
    1: dict(ka=2450.0), 2: dict(kv=3750.0), 4: dict(kd=100.0), 8: dict(ks=14.8), 16: dict(kb=17.78, phi=6.97),
}


@pytest.mark.parametrize("term", sorted(TERMS))
@pytest.mark.parametrize("stress_free", [False, True])
def test_forces_are_minus_energy_gradient(term, stress_free):
    """Each membrane term is -grad of its energy: global area ka (A - A0)^2 /
    (2 A0), volume kv (V - V0)^2 / (2 V0), local area kd (a - a0)^2 / (2 a0),
    WLC + POW springs in closed form, bending kb (1 - cos(theta - phi)) per
    dihedral.  Central differences, h = 1e-4, on a noisy 498-vertex cell
    whose reference shape is the noise-free template."""
    m, pp = perturbed(498, 0.03, 5 + term)
    base = dict(gammaC=0.0, gammaT=0.0, kBT=0.0, kb=0.0, phi=0.0, ks=0.0, x0=0.5, mpow=2.0, ka=0.0, kd=0.0, kv=0.0)
    base.update(TERMS[term])
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    if term == 1 or term == 2:
        A0, V0 = A0 * 0.97, V0 * 1.04          # away from equilibrium
    p = R.par(**base, totArea=np.float32(A0), totVolume=np.float32(V0))
    sf = m.shape() if stress_free else None
    l0, a0 = m.stressful(np.float32(A0))
    av = R.area_volume(m, pp, 1)
    ff = R.forces(p, m, pp, 1, av, stress_free=sf, l0=l0, a0=a0)
    r0 = pp[:, :3].astype(np.float64)
    h = 1e-4
    rng = np.random.default_rng(1)
    scale = np.abs(ff).max()
    assert scale > 0
    for v in rng.choice(m.nv, 12, replace=False):
        for c in range(3):
            rp, rm = r0.copy(), r0.copy()
            rp[v, c] += h
            rm[v, c] -= h
            g = (R.energy(p, m, rp, 1, term, sf, l0, a0) - R.energy(p, m, rm, 1, term, sf, l0, a0)) / (2 * h)
            assert abs(ff[v, c] + g) <= 2e-4 * scale, (v, c, ff[v, c], -g)


def test_internal_forces_balance():
    """Area, springs and bending are internal: total force and torque vanish;
    the viscous pair forces cancel in pairs (Newton's third law)."""
    m, pp = perturbed(642, 0.05, 3)
    pp[:, 3:] = np.random.default_rng(4).normal(0, 1, (m.nv, 3)).astype(np.float32)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    p = R.par(**R.RBC_TEST, totArea=np.float32(A0 * 0.98), totVolume=np.float32(V0 * 1.03))
    av = R.area_volume(m, pp, 1)
    ff = R.forces(p, m, pp, 1, av, stress_free=m.shape())
    fs = np.abs(ff).sum()
    assert np.abs(ff.sum(0)).max() <= 1e-9 * fs
    torque = np.cross(pp[:, :3].astype(np.float64), ff).sum(0)
    # the viscous term gammaC (du.dr) dr is central but gammaT du is not
    # aligned with dr, so only the force balance holds for it; check the
    # torque without viscosity
    p.gammaT = 0.0
    p.gammaC = 0.0
    ff2 = R.forces(p, m, pp, 1, av, stress_free=m.shape())
    torque = np.cross(pp[:, :3].astype(np.float64), ff2).sum(0)
    assert np.abs(torque).max() <= 1e-8 * np.abs(ff2).sum() * 10


def test_stretch():
    f = np.arange(12, dtype=np.float32).reshape(4, 3)
    ff = np.ones((8, 3), np.float32)
    out = R.stretch(2, 4, f, ff)
    assert np.array_equal(out, 1 + np.concatenate([f, f]))
