"""The Juelicher bending oracle (oracle/udx_oracle_jul.c) against what the
mathematics fixes: the edge list covers every mesh edge once with its two
triangles; the dihedral angle of a regular icosahedron is its closed form on
every edge; the dihedral-gradient formula of the reference is minus the
gradient of its angle (central differences); and the force terms against
central differences of float64 numpy restatements of
E = sum_v 2 A_v (L_v / (4 A_v) - H0)^2:
  f = -grad_x sum_e c_e theta_e l_e(x) + grad_x sum_e c_e l_e theta_e(x)
      + sum_v (L_v^2 / (8 A_v^2) - 2 H0^2) d A_v / d x_v,
  c_e = (L_j / (4 A_j) - H0) + (L_k / (4 A_k) - H0) (frozen at x0),
which is -grad E except for the two terms the reference writes differently:
the dihedral term has the opposite sign (its gradient formula is minus the
gradient of its angle) and each triangle's area gradient is weighted by the
coefficient of the vertex it moves only.  The area-difference force is
-(C/4) grad sum_e theta_e(x0) l_e(x) + (C/2) grad sum_e l_e(x0) theta_e(x).
The reference has no test or golden file of this model.
"""
import numpy as np
import pytest

from membranes import icosahedron, mesh
from oracle import rbc as R

H0 = -0.5
KAD = np.float32(2 * np.float32(1.0) / np.float32(np.pi))


def theta64(x, dih):
    a, b, c, d = (x[dih[:, q]] for q in range(4))
    o3 = np.einsum("ij,ij->i", a - d, np.cross(b - d, c - d))
    k = np.cross(b - a, c - a)
    l = np.cross(c - d, b - d)
    return np.arctan2(o3 * np.linalg.norm(b - c, axis=1), np.einsum("ij,ij->i", k, l))


def quantities(x, tt, dih, nv):
    A_t = 0.5 * np.linalg.norm(np.cross(x[tt[:, 1]] - x[tt[:, 0]], x[tt[:, 2]] - x[tt[:, 0]]), axis=1)
    A = np.zeros(nv)
    for q in range(3):
        np.add.at(A, tt[:, q], A_t / 3)
    th = theta64(x, dih)
    le = np.linalg.norm(x[dih[:, 2]] - x[dih[:, 1]], axis=1)
    L = np.zeros(nv)
    np.add.at(L, dih[:, 1], le * th)
    np.add.at(L, dih[:, 2], le * th)
    return A, L, th, le


def energy(x, tt, dih, nv):
    A, L, _, _ = quantities(x, tt, dih, nv)
    return (2 * A * (L / (4 * A) - H0) ** 2).sum()


def grad_fd(fun, x, verts, h=1e-5):
    g = np.zeros((len(verts), 3))
    for n, v in enumerate(verts):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, c] += h
            xm[v, c] -= h
            g[n, c] = (fun(xp) - fun(xm)) / (2 * h)
    return g


def test_edges_cover_the_mesh_once():
    for m in (mesh(498), mesh(642), R.Mesh(*icosahedron())):
        dih = R.jul_edges(m)
        tt = [tuple(t) for t in m.tt]
        assert len(dih) * 2 == 3 * len(tt)
        keys = {tuple(sorted((j, k))) for _, j, k, _ in dih}
        assert len(keys) == len(dih)
        cyc = lambda t: {t, t[1:] + t[:1], t[2:] + t[:2]}
        tset = set()
        for t in tt:
            tset |= cyc(t)
        for i, j, k, l in dih:
            assert (j, k, i) in tset and (k, j, l) in tset


def test_icosahedron_dihedral_closed_form():
    v, f = icosahedron()
    m = R.Mesh(v.astype(np.float32), f)
    dih = R.jul_edges(m)
    want = np.pi - np.arccos(-np.sqrt(5) / 3)      # pi minus the interior dihedral angle
    th = np.array([R.jul_theta(*(v[d] for d in e)) for e in dih])
    assert np.allclose(np.abs(th), want, rtol=0, atol=1e-12)
    assert np.all(np.sign(th) == np.sign(th[0]))   # a convex surface bends one way
    assert np.allclose(th, theta64(v, dih), rtol=0, atol=1e-12)


def noisy(seed, sigma=0.02):
    m = mesh(498)
    rng = np.random.default_rng(seed)
    x = m.rr.astype(np.float64) + rng.normal(0, sigma, m.rr.shape)
    pp = np.zeros((m.nv, 6), np.float32)
    pp[:, :3] = x.astype(np.float32)
    return m, pp, pp[:, :3].astype(np.float64)


def ddih_formula(a, b, c, d):
    """the reference's dihedral gradient (juelicher/dev/main.h:104-128)"""
    def area(a, b, c):
        return 0.5 * np.linalg.norm(np.cross(b - a, c - a))

    def normal(a, b, c):
        n = np.cross(b - a, c - a)
        return n / np.linalg.norm(n)

    def cot(a, b, c):
        return np.dot(a - b, c - b) / (2 * area(a, b, c))

    n, k, e = normal(a, b, c), normal(c, b, d), np.linalg.norm(b - c)
    cn, bn, bk, ck = cot(b, c, a), cot(a, b, c), cot(c, b, d), cot(d, c, b)
    return np.array([n * e / (2 * area(a, b, c)), -cn / e * n - ck / e * k, -bn / e * n - bk / e * k,
                     k * e / (2 * area(c, b, d))])


def test_dihedral_gradient_formula_is_minus_the_angle_gradient():
    rng = np.random.default_rng(3)
    for _ in range(5):
        X = np.array([[0.4, -0.8, 0.1], [0, 0, 0], [1.0, 0.1, -0.05], [0.6, 0.7, 0.3]]) + rng.normal(0, 0.1, (4, 3))
        dih = np.array([[0, 1, 2, 3]])
        g = grad_fd(lambda Y: theta64(Y, dih)[0], X, range(4), h=1e-6)
        assert np.allclose(ddih_formula(*X), -g, rtol=0, atol=1e-7)


@pytest.mark.parametrize("seed", [1, 2])
def test_bending_force_terms(seed):
    m, pp, x = noisy(seed)
    dih = R.jul_edges(m)
    f, fad = R.jul_forces(m, pp, 1)
    A, L, th, le = quantities(x, m.tt, dih, m.nv)
    c = (L[dih[:, 1]] / (4 * A[dih[:, 1]]) - H0) + (L[dih[:, 2]] / (4 * A[dih[:, 2]]) - H0)
    p1 = lambda y: (c * th * np.linalg.norm(y[dih[:, 2]] - y[dih[:, 1]], axis=1)).sum()
    p2 = lambda y: (c * le * theta64(y, dih)).sum()
    verts = np.random.default_rng(seed).choice(m.nv, 12, replace=False)
    own = np.zeros((len(verts), 3))      # d A_v / d x_v
    for n, v in enumerate(verts):
        own[n] = grad_fd(lambda y: quantities(y, m.tt, dih, m.nv)[0][v], x, [v])[0]
    want = -grad_fd(p1, x, verts) + grad_fd(p2, x, verts) + (L[verts] ** 2 / (8 * A[verts] ** 2) - 2 * H0 ** 2)[:, None] * own
    scale = np.abs(f).max()
    assert np.abs(f[verts] - want).max() <= 2e-5 * scale
    # and it is not the energy gradient: the two differing terms matter
    g = grad_fd(lambda y: energy(y, m.tt, dih, m.nv), x, verts)
    assert np.abs(f[verts] + g).max() > 1e-2 * scale


def test_area_difference_force():
    m, pp, x = noisy(3)
    dih = R.jul_edges(m)
    _, fad = R.jul_forces(m, pp, 1)
    A, L, th0, le0 = quantities(x, m.tt, dih, m.nv)
    C = (L.sum() / 4 - H0 * A.sum()) * (4 * KAD * np.pi / A.sum())
    phi1 = lambda y: (th0 * np.linalg.norm(y[dih[:, 2]] - y[dih[:, 1]], axis=1)).sum()
    phi2 = lambda y: (le0 * theta64(y, dih)).sum()
    verts = np.random.default_rng(7).choice(m.nv, 12, replace=False)
    want = -(C / 4) * grad_fd(phi1, x, verts) + (C / 2) * grad_fd(phi2, x, verts)
    scale = np.abs(fad).max()
    assert np.abs(fad[verts] - want).max() <= 2e-5 * scale


def test_cells_are_independent():
    m, pp, _ = noisy(4)
    two = np.concatenate([pp, pp + np.float32([5, 0, 0, 0, 0, 0])])
    f1, g1 = R.jul_forces(m, pp, 1)
    f2, g2 = R.jul_forces(m, two, 2)
    assert np.array_equal(f2[:m.nv], f1) and np.array_equal(g2[:m.nv], g1)
    assert np.allclose(f2[m.nv:], f1, atol=1e-5 * np.abs(f1).max())
