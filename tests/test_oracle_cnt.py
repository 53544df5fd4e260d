"""The membrane contact oracle (oracle/udx_oracle_cnt.c) against closed forms
and brute force: the capped repulsive Lennard-Jones term, and on cells that
touch each other and straddle the periodic faces, every vertex pair within
the cutoff (bulk, noise indices (higher, lower)) and every pair of a periodic
image of an outside vertex with a vertex (halo, noise indices (image, vertex)),
summed in float64; total momentum is conserved.  The reference has no test
of this module.
"""
import numpy as np

import oracle as O
from membranes import mesh, rotation
from oracle import rbc as R

PAR = (20.0, 8.0, 3.0, 0.25, 0.3, 1.0)   # a, g, s, spow, ljs, lje


def test_lennard_jones_closed_form():
    for s, e in ((0.5, 1.0), (0.3, 2.5)):
        for r in (0.25, 0.31, 0.4, 0.5, 0.55, 0.7):
            f, ok = R.cnt_pair((0, 0, 0, 0.25, s, e), [r, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0], 0.0)
            t6 = (s / r) ** 6
            want = min(max(24 * e / r * t6 * (2 * t6 - 1), 0.0), 1e4)
            # float rounding of the powers and of 2 t6 - 1
            tol = 1e-6 * 24 * e / r * t6 * (2 * t6 + 1)
            assert ok and abs(f[0] - want) <= tol and f[1] == 0 and f[2] == 0


def contact_case(seed, L=(16, 16, 16)):
    m = mesh(498)
    rng = np.random.default_rng(seed)
    centres = [(L[0] / 2 - 0.5, 0.0, 0.0), (-L[0] / 2 + 4.6, 0.3, -0.2), (1.0, -3.0, 2.0), (1.2, -3.0, 6.8)]
    pp = np.zeros((len(centres) * m.nv, 6), np.float32)
    for k, c in enumerate(centres):
        x = m.rr.astype(np.float64) @ rotation(rng).T * 1.1 + np.array(c)
        pp[k * m.nv:(k + 1) * m.nv, :3] = x.astype(np.float32)
    pp[:, 3:] = rng.normal(0, 0.3, (len(pp), 3)).astype(np.float32)
    return pp


def halo_list(L, pp):
    """images of the vertices outside the domain: receive fragment i = 0..25
    (direction d_i), vertices ascending, each shifted by +L d_i when it lies
    outside toward -d_i on every axis where d_i is nonzero"""
    L = np.array(L)
    out = []
    for i in range(26):
        t = np.array([i % 3, (i // 3) % 3, i // 9])
        d = np.where(t == 0, 1, np.where(t == 1, -1, 0))
        for v, p in enumerate(pp):
            code = -1 + (p[:3] >= -(L // 2)).astype(int) + (p[:3] >= L // 2).astype(int)
            if not code.any():
                continue
            if np.all((d == 0) | (code == -d)):
                q = p.copy()
                q[:3] = p[:3] + (L * d).astype(np.float32)
                out.append((v, q))
    return out


def test_images_match_the_halo_list():
    L = (16, 16, 16)
    pp = contact_case(1, L)
    img, iv = R.cnt_images(L, pp)
    want = halo_list(L, pp)
    assert len(want) == len(img) > 0
    assert np.array_equal(iv, [v for v, _ in want])
    assert np.array_equal(img, np.array([q for _, q in want]))


def test_forces_brute_force():
    L = (16, 16, 16)
    pp = contact_case(2, L)
    sb, sh = np.float32(0.123), np.float32(0.456)
    got, npair = R.cnt_forces(L, pp, PAR, sb, sh)
    want = np.zeros((len(pp), 3))
    count = 0
    x = pp[:, :3].astype(np.float64)
    for a in range(len(pp)):
        d2 = ((x[:a] - x[a]) ** 2).sum(1)
        for b in np.nonzero(d2 < 1)[0]:
            f, ok = R.cnt_pair(PAR, pp[a], pp[b], O.rng_ii(sb, [a], [b])[0])
            if ok:
                want[a] += f
                want[b] -= f
                count += 1
    for h, (v, q) in enumerate(halo_list(L, pp)):
        d2 = ((x - q[:3].astype(np.float64)) ** 2).sum(1)
        for b in np.nonzero(d2 < 1)[0]:
            f, ok = R.cnt_pair(PAR, q, pp[b], O.rng_ii(sh, [h], [b])[0])
            if ok:
                want[v] += f
                want[b] -= f
                count += 1
    assert npair == count and count > 1000
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-5 * scale
    # momentum: every pair acts on two vertices with opposite forces
    assert np.abs(got.astype(np.float64).sum(0)).max() <= 1e-4 * scale
