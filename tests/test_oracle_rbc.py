"""The RBC membrane oracle (oracle/udx_oracle_rbc.c) against the reference's
golden files, the energy it is the gradient of, and mesh invariants.

Fixtures: tests/golden/ref/cells/{498,642,1986}.off are the reference's
src/data/cells/rbc meshes (its sc/642.off, the stress-free shape, is the same
file as 642.off); av.ref.rbc.u.area_volume.dev.0 and
forces.ref.rbc.u.{simple,stree.free} are its src/test_data files for
src/test/rbc/u/{area_volume/dev,simple,stress.free}, run with
conf/rbc.test.cfg (tests/golden/ref/conf/rbc.test.cfg), the initial condition
"two" (identity and a shift by (1, 2, 3)) in a 16^3 box, positions and forces
printed %g after scaling by 0.1 (simple) or 0.025 (stress free);
eng.ref.mesh.eng_kantor is its src/test/mesh/eng_kantor output on
src/data/cells/rbc/sc/2562.off (tests/golden/ref/cells/sc2562.off).
"""
import os

import numpy as np
import pytest

from oracle import rbc as R
from membranes import C4, COEF, GOLD, adj_setup, icosahedron, mesh


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
    v, f = icosahedron()
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
            # every slot: i1, i2, i3 are the slot's neighbour and its two
            # cyclic successors (also across the end of a padded ring), i4
            # the vertex across (i0, i1); padded slots are invalid; the
            # second cell's slots carry its vertex offset
            # (src/rbc/adj/imp/map.h:2-27 hst0)
            for c in (0, 1):
                for j in range(m.md):
                    mm = m.map((c * m.nv + v) * m.md + j)
                    if j >= deg[v]:
                        assert mm is None
                        continue
                    want = [c * m.nv + v] + [c * m.nv + x for x in
                                             (ring[j], ring[(j + 1) % deg[v]], ring[(j + 2) % deg[v]],
                                              m.adj1[v * m.md + j])] + [c]
                    assert list(mm) == want, (nv, v, c, j)
        # both ring ends occur: full rings (j + 1 wraps by the modulus) and
        # padded ones (wraps at the first -1)
        assert (deg == m.md).any() and (deg < m.md).any()


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


# ------------------------------------------------------------------ FSI
def fsi_setup(L=(12, 10, 14), rho=4, seed=3, out=(1.0, 1.2, 0.4)):
    """solvent in the box and a 498-vertex cell centred near the upper
    corner so that some vertices lie outside the box on every axis"""
    import oracle as O
    rng = np.random.default_rng(seed)
    n = int(rho * np.prod(L))
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    starts, counts, srt, _ = O.clist_build(L, spp)
    m = mesh(498)
    c = np.array(L, np.float32) / 2 - np.array(out, np.float32)
    opp = np.zeros((m.nv, 6), np.float32)
    opp[:, :3] = (m.rr + c).astype(np.float32)
    opp[:, 3:] = rng.normal(0, 0.3, (m.nv, 3)).astype(np.float32)
    return srt, starts, opp


def fsi_brute(L, srt, opp, par, seeds):
    """all (instance, solvent) pairs: vertices and their periodic images
    (vertices outside the box, shifted back by a period on every subset of
    their outside axes), the solvent cell within one cell of the instance's
    cell, r^2 < 1; noise index = vertex or position in the halo list"""
    import oracle as O
    Lh = np.array([x // 2 for x in L])
    inst = [(opp[v], v, v, 0) for v in range(len(opp))]
    aid = 0
    for i in range(26):
        di = np.array(O.frag_i2d3(i))
        for v in range(len(opp)):
            d = -1 + (opp[v, :3] >= -Lh).astype(int) + (opp[v, :3] >= Lh).astype(int)
            s = -di
            if not np.any(d) or np.any((s != 0) & (s != d)) or not np.any(s):
                continue
            q = opp[v].copy()
            q[:3] = (q[:3] + (np.array(L) * di).astype(np.float32)).astype(np.float32)
            inst.append((q, v, aid, 1))
            aid += 1
    fo = np.zeros((len(opp), 3))
    fs = np.zeros((len(srt), 3))
    cs = np.floor(srt[:, :3].astype(np.float64) + np.array(L) / 2).astype(int)
    npair = 0
    for q, v, lid, halo in inst:
        cq = Lh + np.floor(q[:3]).astype(int)
        near = np.all(np.abs(cs - cq) <= 1, axis=1)
        for b in np.nonzero(near)[0]:
            xi = O.rng_uu(seeds[halo], np.array([lid], np.uint32), np.array([b], np.uint32))[0]
            f, ok = O.pair_force(par, q, srt[b], xi)
            npair += ok
            fo[v] += f
            fs[b] -= f
    return fo, fs, npair, aid


def test_fsi_forces_brute_force():
    import oracle as O
    L = (12, 10, 14)
    srt, starts, opp = fsi_setup(L)
    par = (2.6666667, 8.0, O.dpd_sigma(0.0444, 8.0, 5e-4), 0.25)
    seeds = np.array([0.31, 0.77], np.float32)
    fo, fs, n = R.fsi_forces(L, srt, starts, opp, par, seeds)
    bo, bs, nb, nimg = fsi_brute(L, srt, opp, par, seeds)
    assert nimg > 50 and n == nb and n > 1000
    scale = max(np.abs(bo).max(), np.abs(bs).max())
    assert np.abs(fo - bo).max() <= 1e-5 * scale
    assert np.abs(fs - bs).max() <= 1e-5 * scale
    # momentum: every pair acts on both sides
    assert np.abs(fo.astype(np.float64).sum(0) + fs.astype(np.float64).sum(0)).max() <= 1e-4 * scale


def test_fsi_inside_vertices_make_no_images():
    import oracle as O
    L = (12, 10, 14)
    srt, starts, opp = fsi_setup(L, out=(4.0, 3.5, 2.0))      # the whole cell inside
    assert np.all(np.abs(opp[:, :3]) < np.array(L) / 2)
    par = (2.6666667, 8.0, 1.0, 0.25)
    fo1, fs1, n1 = R.fsi_forces(L, srt, starts, opp, par, np.array([0.3, 0.1], np.float32))
    fo2, fs2, n2 = R.fsi_forces(L, srt, starts, opp, par, np.array([0.3, 0.9], np.float32))
    assert n1 == n2 and np.array_equal(fo1, fo2) and np.array_equal(fs1, fs2)   # the halo seed is unused


def test_fsi_seeds_trunk():
    import oracle as O
    s = R.fsi_seeds(3)
    assert np.array_equal(s.ravel(), O.kiss_fill(1908, 1409, 290, 12968, 6))



def test_kantor_energy_golden():
    """eng.ref.mesh.eng_kantor: the reference's Kantor energy utility
    (src/u/mesh/eng_kantor/main.cpp, src/mesh/eng_kantor/imp/main.h:30-39)
    on its 2562-vertex stress-free RBC mesh with kb = 1 and angle 0 prints,
    per vertex, the cylindrical coordinates and the mean of the energies
    2 kb (1 - cos theta) of the vertex's edges (src/mesh/scatter/imp/main.h).
    Sum over vertices of mean times degree = twice the total edge energy.  The
    oracle's bending energy kb (1 - cos theta) per edge, the potential whose
    negative gradient the Kantor force is (checked by central differences),
    is half of the utility's per-edge value: the totals differ by exactly 2
    (to the 6 printed digits)."""
    rr, tt = R.read_off(os.path.join(GOLD, "cells", "sc2562.off"))
    m = R.Mesh(rr, tt)
    ref = np.loadtxt(os.path.join(GOLD, "eng.ref.mesh.eng_kantor.txt"))
    assert ref.shape == (m.nv, 4)
    assert np.abs(np.hypot(rr[:, 0], rr[:, 1]) - ref[:, 0]).max() < 1e-4      # same vertex order
    assert np.abs(rr[:, 2] - ref[:, 2]).max() < 1e-5
    edges = {(min(u, v), max(u, v)) for a, b, c in tt for u, v in ((a, b), (b, c), (c, a))}
    assert len(edges) == 3 * m.nv - 6                                        # closed genus-0 mesh
    deg = np.zeros(m.nv)
    for u, v in edges:
        deg[u] += 1
        deg[v] += 1
    utility = (ref[:, 3] * deg).sum() / 2
    p = R.par(kb=1.0, phi=0.0, ka=0.0, kd=0.0, kv=0.0, ks=0.0, x0=0.5, mpow=2.0, totArea=1.0, totVolume=1.0)
    E = R.energy(p, m, rr.astype(np.float64), 1, 16)
    assert abs(utility / (2 * E) - 1) < 1e-6


# ------------------------------------------------------------------ adjoint
@pytest.mark.parametrize("stress_free", [False, True])
def test_adjoint_terms_sum_to_the_force(stress_free):
    m, pp, kw, sf, l0, a0, g = adj_setup(stress_free=stress_free)
    p = R.par(**kw)
    gx, terms, gp = R.adjoint(p, m, pp, 2, g, sf, l0, a0)
    F = R.forces(p, m, pp, 2, R.area_volume(m, pp, 2), stress_free=sf, l0=l0, a0=a0)
    tot = sum(kw[c] * terms[k] for k, c in enumerate(COEF))
    # the force rounds the spring length and powf to float, the adjoint does
    # not: 2e-6 of the largest (spring-dominated) force at most
    assert np.abs(tot - F).max() <= 1e-5 * np.abs(F).max()
    assert np.allclose(gp, [(g * terms[k]).sum() for k in range(5)], rtol=1e-12, atol=0)


@pytest.mark.parametrize("term", range(5))
def test_adjoint_is_the_force_jacobian(term):
    """gx = (dF/dx) g for each term alone: central differences of the force
    along g (h = 1e-3 on float positions: 2e-3 of the largest component),
    and g . gx = -d^2 E / d eps^2 along g from the FP64 energy (1e-5).  The
    bending term is checked on an icosahedron: the RBC meshes have nearly
    flat dihedrals where the reference's 1/sin(theta) (clamped at 1e-3 and
    signed) is not differentiable."""
    m, pp, kw, sf, l0, a0, g = adj_setup(seed=3 + term, ico=term == 4)
    base = dict(kw, ka=0.0, kd=0.0, kv=0.0, ks=0.0, kb=0.0)
    base[COEF[term]] = kw[COEF[term]]
    p = R.par(**base)
    gx, _, _ = R.adjoint(p, m, pp, 2, g, sf, l0, a0)
    h = 1e-3

    def force(x):
        q = pp.copy()
        q[:, :3] = x.astype(np.float32)
        return R.forces(p, m, q, 2, R.area_volume(m, q, 2), stress_free=sf, l0=l0, a0=a0)
    x0 = pp[:, :3].astype(np.float64)
    fd = (force(x0 + h * g) - force(x0 - h * g)) / (2 * h)
    assert np.abs(fd - gx).max() <= 2e-3 * np.abs(gx).max()
    bit = (1, 4, 2, 8, 16)[term]          # energy term bits of orc_rbc_energy
    e = lambda x: R.energy(p, m, x, 2, bit, sf, l0, a0)
    hh = 1e-4
    d2 = (e(x0 + hh * g) - 2 * e(x0) + e(x0 - hh * g)) / (hh * hh)
    assert abs((g * gx).sum() + d2) <= 1e-5 * abs(d2) + 1e-8 * np.abs(g * gx).sum()


def test_c4_fp64_gradient_is_the_central_difference():
    """The FP64 CPU gradient of the membrane calibration loss (4 cells, 4
    relaxation steps) through the oracle adjoint against central differences
    with h = 1e-3: along a random direction of the initial positions (spontaneous
    angle 0, where the bending energy is smooth: with the reference's 6.97
    degrees its 1/sin(theta) flips sign at flat dihedrals and differences
    across one are meaningless) within 1e-5, and for each coefficient
    (h = 1e-3 of it, the reference's angle) within 1e-5."""
    c = C4(nc=4, seed=2, phi=0.0)
    _, gx, _ = c.grad64()
    h = 1e-3
    fd = (c.loss64(x0=c.x0 + h * c.u)[0] - c.loss64(x0=c.x0 - h * c.u)[0]) / (2 * h)
    assert abs(fd - (gx * c.u).sum()) <= 1e-5 * abs(fd)
    c = C4(nc=4, seed=2)
    _, _, gk = c.grad64()
    k = c.coef()
    for j in range(5):
        kp, km = k.copy(), k.copy()
        kp[j] += h * k[j]
        km[j] -= h * k[j]
        fdk = (c.loss64(coef=kp)[0] - c.loss64(coef=km)[0]) / (2 * h * k[j])
        assert abs(fdk - gk[j]) <= 1e-5 * abs(gk[j]), COEF[j]


def test_redistribute_whole_cells():
    """Cells move as a whole by their bounding-box centre: kept cells first,
    moved cells after them in fragment order, each shifted by -L d, every
    centre back inside the box; a cell whose centre sits on +L/2 moves
    (the interval is [-L/2, L/2)), and its shifted centre may round to just
    below -L/2, as in the reference (float positions, float shift)."""
    L = (16, 12, 20)
    m = mesh(498)
    rng = np.random.default_rng(4)
    centres = [[0.0, 0.0, 0.0], [8.5, 0.0, 0.0], [-3.0, -7.0, 11.0], [2.0, 3.0, -1.0], [0.0, 6.0, 0.0]]
    T = []
    for c in centres:
        A = np.eye(4)
        A[:3, 3] = c
        T.append(A)
    pp = R.place(m.rr, T, (0, 0, 0))
    # the fifth cell: bounding-box centre exactly on +Ly/2
    x = pp[4 * m.nv:, 1]
    x += np.float32(6.0) - np.float32(0.5) * (x.min() + x.max())
    out, fid = R.redistribute(L, pp, len(centres), m.nv)
    d = np.array([[0, 0, 0], [1, 0, 0], [0, -1, 1], [0, 0, 0], [0, 1, 0]])
    want_fid = (d[:, 0] + 2) % 3 + 3 * ((d[:, 1] + 2) % 3) + 9 * ((d[:, 2] + 2) % 3)
    assert np.array_equal(fid, want_fid)
    order = [0, 3] + sorted([1, 2, 4], key=lambda c: want_fid[c])
    src = pp.reshape(5, m.nv, 6)
    got = out.reshape(5, m.nv, 6)
    for k, c in enumerate(order):
        assert np.array_equal(got[k, :, 3:], src[c, :, 3:])
        assert np.array_equal(got[k, :, :3], src[c, :, :3] + (-d[c] * np.array(L)).astype(np.float32))
        lo, hi = got[k, :, :3].min(0), got[k, :, :3].max(0)
        ctr = np.float32(0.5) * (lo + hi)
        assert np.all(ctr >= -np.array(L) / 2 - 1e-5) and np.all(ctr < np.array(L) / 2)
