"""The membrane bounce-back oracle (oracle/udx_oracle_mbb.c) against what
the mathematics fixes: the crossing time and barycentric point of a particle
with a moving triangle from float64 bisection of the signed volume, the
triangle inertia from numerical integration, the periodic mesh images from
the bounding boxes, and on a whole step: every particle keeps its side of
every membrane, every bounce is the particle's earliest crossing over all
triangle instances, and momentum is conserved between solvent and vertices.
"""
import numpy as np
import pytest

import bench
from membranes import bounce_case
from oracle import mbb as B


def crossing64(dt, s, vs, r, v):
    """first root in [0, dt] of det(s2 - s1, s3 - s1, r - s1)(t) by sampling
    and bisection in float64, and the barycentric (u, v) there; None when
    the signed volume has the same sign at both ends"""
    s = [np.asarray(x, np.float64) for x in s]
    vs = [np.asarray(x, np.float64) for x in vs]
    r, v = np.asarray(r, np.float64), np.asarray(v, np.float64)

    def f(t):
        t = np.asarray(t, np.float64)[..., None]
        a = s[1] + t * vs[1] - s[0] - t * vs[0]
        b = s[2] + t * vs[2] - s[0] - t * vs[0]
        return (np.cross(a, b) * (r + t * v - s[0] - t * vs[0])).sum(-1)

    if f(0.0) * f(dt) > 0:
        return None
    ts = np.linspace(0, dt, 2001)
    fs = f(ts)
    k = int(np.nonzero(np.sign(fs[1:]) != np.sign(fs[:-1]))[0][0])
    lo, hi = ts[k], ts[k + 1]
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.sign(f(mid)) == np.sign(f(lo)):
            lo = mid
        else:
            hi = mid
    h = 0.5 * (lo + hi)
    a1 = s[1] + h * vs[1] - s[0] - h * vs[0]
    a2 = s[2] + h * vs[2] - s[0] - h * vs[0]
    g = r + h * v - s[0] - h * vs[0]
    (u, w), *_ = np.linalg.lstsq(np.stack([a1, a2], 1), g, rcond=None)
    return h, u, w


@pytest.mark.parametrize("motion", ["static", "translating", "deforming"])
def test_intersect_against_bisection(motion):
    rng = np.random.default_rng({"static": 1, "translating": 2, "deforming": 3}[motion])
    dt = 0.05
    hits = wrong = 0
    for _ in range(400):
        s = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        if motion == "static":
            vs = np.zeros((3, 3), np.float32)
        elif motion == "translating":
            vs = np.tile(rng.normal(0, 2, 3), (3, 1)).astype(np.float32)
        else:
            vs = rng.normal(0, 4, (3, 3)).astype(np.float32)
        # a particle that crosses the plane of the triangle near a random
        # point of it (or just outside) within the step
        bc = rng.dirichlet([1, 1, 1]) * rng.uniform(0.6, 1.3)
        tc = rng.uniform(0.05, 0.95) * dt
        x = sum(bc[j] * (s[j] + tc * vs[j]) for j in range(3))
        v = rng.normal(0, 3, 3)
        r = (x - tc * v).astype(np.float32)
        v = v.astype(np.float32)
        st, h, u, w, side = B.intersect(dt, *s, *vs, r, v)
        ref = crossing64(dt, s, vs, r, v)
        if ref is None:
            assert st == B.BB_NOCROSS
            continue
        h64, u64, w64 = ref
        inside = u64 >= 0 and w64 >= 0 and u64 + w64 <= 1
        margin = min(u64, w64, 1 - u64 - w64)
        if abs(margin) < 1e-4:
            continue                      # on an edge: either answer
        if not inside:
            assert st in (B.BB_WTRIANGLE, B.BB_HFAIL)
            wrong += 1
            continue
        assert st == B.BB_SUCCESS, (st, ref)
        # the reference forms r - s1 and v - vs1 in float: an ulp of the
        # coordinates (|x| <= 2) over the normal speed relative to the triangle
        sh = [s[j].astype(np.float64) + h64 * vs[j] for j in range(3)]
        nrm = np.cross(sh[1] - sh[0], sh[2] - sh[0])
        vn = abs(np.dot(v - vs.astype(np.float64).T @ np.array([1 - u64 - w64, u64, w64]), nrm / np.linalg.norm(nrm)))
        assert abs(h - h64) <= 1e-6 * dt + 4 * 2.0 ** -23 / vn
        assert abs(u - u64) <= 2e-5 and abs(w - w64) <= 2e-5
        hits += 1
    assert hits > 100 and wrong > 20


def test_intersect_static_plane_closed_form():
    """triangle in z = 0 at rest, particle from z = 0.1 to z = -0.3 over
    dt = 0.4: crossing at t = 0.1, above the point (0.25, 0.5)"""
    s = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    z = np.zeros((3, 3), np.float32)
    st, h, u, w, side = B.intersect(0.4, *s, *z, [0.25, 0.5, 0.1], [0.0, 0.0, -1.0])
    assert st == B.BB_SUCCESS and side == 1.0
    assert abs(h - 0.1) < 1e-7 and abs(u - 0.25) < 1e-7 and abs(w - 0.5) < 1e-7
    st, *_ = B.intersect(0.05, *s, *z, [0.25, 0.5, 0.1], [0.0, 0.0, -1.0])
    assert st == B.BB_NOCROSS
    st, *_ = B.intersect(0.4, *s, *z, [0.8, 0.8, 0.1], [0.0, 0.0, -1.0])
    assert st == B.BB_WTRIANGLE


def test_inertia_closed_form():
    """1.5 m times the inertia tensor of the uniform triangle about its
    centroid (the reference's normalisation), from a quadrature of 2e5 points"""
    rng = np.random.default_rng(4)
    for _ in range(5):
        a, b, c = rng.uniform(-2, 2, (3, 3)).astype(np.float32)
        mass = np.float32(0.7)
        got = B.inertia(mass, a, b, c)
        w = rng.dirichlet([1, 1, 1], 200000)
        p = w @ np.stack([a, b, c]).astype(np.float64)
        p -= p.mean(0)
        J = np.eye(3) * (p * p).sum(1).mean() - p.T @ p / len(p)
        want = 1.5 * mass * J
        got_m = np.array([[got[0], got[1], got[2]], [got[1], got[3], got[4]], [got[2], got[4], got[5]]])
        assert np.allclose(got_m, want, rtol=0, atol=2e-2 * np.abs(want).max())
        # exact against the closed-form moments of the uniform triangle
        x = np.stack([a, b, c]).astype(np.float64)
        xc = x - x.mean(0)
        C = (xc.T @ xc) / 12.0        # covariance of the uniform triangle
        want = 1.5 * mass * (np.eye(3) * np.trace(C) - C)
        assert np.allclose(got_m, want, rtol=1e-5, atol=1e-6)
        inv = B.inverse(got)
        inv_m = np.array([[inv[0], inv[1], inv[2]], [inv[1], inv[3], inv[4]], [inv[2], inv[4], inv[5]]])
        assert np.allclose(inv_m @ got_m, np.eye(3), atol=1e-4)


def test_m2v_momentum():
    rng = np.random.default_rng(5)
    a, b, c = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
    mass = 0.5
    P = np.float32([0.3, -0.2, 0.1])
    v = B.m2v(mass, np.concatenate([P, np.zeros(3, np.float32)]), a, b, c)
    assert np.array_equal(v[0], v[1]) and np.array_equal(v[1], v[2])
    assert np.allclose(v[0], P / 3 / mass, rtol=1e-6)
    Lm = np.float32([0.02, 0.05, -0.03])
    v = B.m2v(mass, np.concatenate([np.zeros(3, np.float32), Lm]), a, b, c)
    com = (a.astype(np.float64) + b + c) / 3
    assert np.allclose(v.sum(0), 0, atol=1e-6)
    for x, vx in zip((a, b, c), v):
        assert abs(np.dot(x - com, vx)) <= 1e-5 * np.linalg.norm(vx) * np.linalg.norm(x - com) + 1e-7
    assert B.nonzero([1e-5] * 6) and not B.nonzero([1e-5] * 5 + [1e-7])


def test_instances_from_bounding_boxes():
    L = (16, 12, 20)
    nv = 2
    rc = np.zeros((4 * nv, 6), np.float32)
    rc[0:2, :3] = [[-1, 0, 0], [1, 0, 0]]              # inside: no image
    rc[2:4, :3] = [[6.5, 0, 0], [7.2, 0, 0]]           # past x = 7 (16/2 - 1): one image at x - 16
    rc[4:6, :3] = [[-7.5, -5.5, 9.2], [-6.0, -4.0, 9.5]]   # corner -x -y +z: 7 images
    rc[6:8, :3] = [[0, 4.99, 0], [0, 5.0, 0]]          # y reaches 5 = 12/2 - 1: one image
    im, sh = B.instances(L, 4, nv, rc)
    assert list(im[:4]) == [0, 1, 2, 3]
    img = sorted((int(m), tuple(float(x) for x in s)) for m, s in zip(im[4:], sh[4:]))
    want = sorted([(1, (-16.0, 0.0, 0.0)), (3, (0.0, -12.0, 0.0))] +
                  [(2, (16.0 * dx, 12.0 * dy, -20.0 * dz)) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)
                   if dx or dy or dz])
    assert img == want


def inside_any(pts, case, rr_key):
    """inside any instance (original or a periodic image) of any cell"""
    L = np.array(case["L"], np.float64)
    nv, tt = case["nv"], case["tt"]
    rr = case[rr_key][:, :3].astype(np.float64)
    out = np.zeros(len(pts), bool)
    for k in range(case["nm"]):
        x = rr[k * nv:(k + 1) * nv]
        for sx in (-1, 0, 1):
            for sy in (-1, 0, 1):
                for sz in (-1, 0, 1):
                    y = x + np.array([sx, sy, sz]) * L
                    lo, hi = y.min(0) - 0.5, y.max(0) + 0.5
                    sel = np.all((pts >= lo) & (pts <= hi), 1)
                    if sel.any():
                        out[sel] |= bench.inside_mesh(pts[sel], y, tt)
    return out


def test_bounce_keeps_every_particle_on_its_side():
    c = bounce_case(seed=1)
    pp, rc, mm, col, ncol = B.bounce(c["dt"], 1.0, 0.5, c["L"], c["nm"], c["tt"], c["rp"], c["rc"], c["starts"],
                                     c["pp0"], c["pp"])
    assert ncol > 20 and ncol == int((col >= 0).sum())
    L = np.array(c["L"], np.float64)
    start = inside_any(c["pp0"][:, :3].astype(np.float64), c, "rp")
    moved = inside_any(c["pp"][:, :3].astype(np.float64), c, "rc")
    end = inside_any(pp[:, :3].astype(np.float64), c, "rc")
    assert int((start != moved).sum()) >= ncol * 0.8       # the unbounced motion would cross
    assert np.array_equal(start, end)
    # particles not bounced are untouched
    same = col < 0
    assert np.array_equal(pp[same], c["pp"][same])
    # images acted: some bounced particle sits near the -x face
    assert np.any(pp[col >= 0, 0] < -L[0] / 2 + 1.5)


def test_bounce_is_the_earliest_crossing():
    c = bounce_case(seed=2)
    pp, rc, mm, col, ncol = B.bounce(c["dt"], 1.0, 0.5, c["L"], c["nm"], c["tt"], c["rp"], c["rc"], c["starts"],
                                     c["pp0"], c["pp"])
    im, sh = B.instances(c["L"], c["nm"], c["nv"], c["rc"])
    dt, nv, tt = np.float32(c["dt"]), c["nv"], c["tt"]
    rcs, rps = c["rc"][:, :3], c["rp"]
    idx = np.nonzero(col >= 0)[0][:40]
    assert len(idx) > 20
    for i in idx:
        r0 = c["pp0"][i, :3]
        v = ((c["pp"][i, :3] - r0).astype(np.float64) / np.float64(dt)).astype(np.float32)
        best = (np.inf, -1)
        for k in range(len(im)):
            base = im[k] * nv
            P = rcs[base:base + nv] + sh[k]
            Q = rps[base:base + nv] + sh[k]
            V = (P - Q) / dt
            S = P - V * dt
            near = np.all(np.abs(S - r0) < 3, 1)
            for t in np.nonzero(near[tt[:, 0]] | near[tt[:, 1]] | near[tt[:, 2]])[0]:
                j = tt[t]
                st, h, *_ = B.intersect(dt, S[j[0]], S[j[1]], S[j[2]], V[j[0]], V[j[1]], V[j[2]], r0, v)
                if st == B.BB_SUCCESS and h < best[0]:
                    best = (h, im[k] * len(tt) + t)
        assert best[1] == col[i]


def test_bounce_conserves_momentum():
    c = bounce_case(seed=3)
    mass_f, mass_m = 1.0, 0.5
    pp, rc, mm, col, ncol = B.bounce(c["dt"], mass_f, mass_m, c["L"], c["nm"], c["tt"], c["rp"], c["rc"],
                                     c["starts"], c["pp0"], c["pp"])
    dp_flu = mass_f * (pp[:, 3:].astype(np.float64) - c["pp"][:, 3:]).sum(0)
    P_tri = mm[:, :3].astype(np.float64).sum(0)
    assert np.allclose(dp_flu + P_tri, 0, atol=1e-4 * np.abs(dp_flu).max())
    active = np.array([B.nonzero(m) for m in mm])
    assert active.sum() > 20
    dp_vert = mass_m * (rc[:, 3:].astype(np.float64) - c["rc"][:, 3:]).sum(0)
    assert np.allclose(dp_vert, mm[active, :3].astype(np.float64).sum(0), atol=1e-4 * np.abs(dp_vert).max())
