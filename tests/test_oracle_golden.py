"""The oracle against the reference's golden files and brute force.

Fixtures (tests/golden/ref/) are copies of the reference's
src/test_data/*.ref.*.txt and of its inputs src/data/pair/{2,rnd}.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")


def _gold(name):
    return os.path.join(GOLD, name)


# ------------------------------------------------------------ gen_random
@pytest.mark.skipif(not O.ref_gen_random_available(), reason="reference gen_random not built")
@pytest.mark.parametrize("args", [(4, 6, 8, 128), (2, 6, 8, 64), (20, 6, 8, 64), (4, 4, 4, 640)])
def test_gen_random_matches_reference_build(args):
    want = O.ref_gen_random_text(*args)
    got = "".join("%f %f %f %f %f %f\n" % tuple(r) for r in O.gen_random(*args, roundtrip=False))
    assert got == want


# ------------------------------------------------------------ cell list
@pytest.mark.parametrize("name,gen", [("t0", (4, 6, 8, 128)), ("t1", (2, 6, 8, 64)), ("t2", (20, 6, 8, 64))])
def test_clist_golden(name, gen):
    """src/test/clist/{0,1,2}/main: box 4x6x8; flag then (start, count) per cell."""
    pp = O.gen_random(*gen)
    starts, counts, out, order = O.clist_build((4, 6, 8), pp)
    ref = np.loadtxt(_gold("clist.ref.clist.%s.txt" % name), dtype=np.int64)
    assert ref[0] == 0
    assert np.array_equal(ref[1::2], starts[:-1])
    assert np.array_equal(ref[2::2], counts)
    # every particle lies in its cell (the harness' validity flag)
    L = np.array([4, 6, 8])
    cells = np.repeat(np.arange(counts.size), counts)
    cz, cy, cx = cells // 24, (cells // 4) % 6, cells % 4
    r = out[:, :3].astype(np.float64) + L / 2
    assert np.all(np.floor(r[:, 0]) == cx) and np.all(np.floor(r[:, 1]) == cy) and np.all(np.floor(r[:, 2]) == cz)


@pytest.mark.parametrize("sub", [1, 2, 4, 8])
def test_clist_stable_order_and_brute_force(sub):
    """order = (cell, x slab of width 1/sub, input index); cell and slab
    starts against brute-force counts."""
    rng = np.random.default_rng(0)
    L = (5, 7, 3)
    pp = np.zeros((500, 6), np.float32)
    pp[:, :3] = rng.uniform(-4, 4, (500, 3))   # some outside the box
    pp[:7, 0] = [-2.5, -2.25, 2.25, 0.0, 0.125, -0.375, 2.4999998]   # on slab faces
    starts, counts, out, order, bstarts = O.clist_build(L, pp, sub=sub, slabs=True)
    h = np.array(L) / 2.0
    r = pp[:, :3].astype(np.float64)
    inside = np.all((r >= -h) & (r < h), axis=1)
    c = np.floor(r + h).astype(np.int64)
    cid = c[:, 0] + L[0] * (c[:, 1] + L[1] * c[:, 2])
    slab = np.floor((r[:, 0] + h[0]) * sub).astype(np.int64) - c[:, 0] * sub
    assert np.all((slab[inside] >= 0) & (slab[inside] < sub))
    want = sorted(np.nonzero(inside)[0], key=lambda i: (cid[i], slab[i], i))
    assert list(order) == want
    assert np.array_equal(out, pp[want])
    bid = (cid * sub + slab)[inside]
    assert np.array_equal(np.diff(bstarts), np.bincount(bid, minlength=np.prod(L) * sub))
    assert np.array_equal(starts, bstarts[::sub])
    assert np.array_equal(counts, np.bincount(cid[inside], minlength=np.prod(L)))
    assert starts[-1] == inside.sum()


# ------------------------------------------------------------ exclusive scan
@pytest.mark.parametrize("name,n", [("t0", 16), ("t1", 64)])
def test_scan_golden(name, n):
    """scan.t0/t1: exclusive scan of `seq n` (src/algo/scan/cpu/imp.h:4-12)."""
    x = np.arange(1, n + 1)
    ref = np.loadtxt(_gold("scan.ref.scan.%s.txt" % name), dtype=np.int64)
    assert np.array_equal(np.concatenate([[0], np.cumsum(x)[:-1]]), ref)


# ------------------------------------------------------------ pair force
_KIND = {"SOLVENT": 0, "O": 0, "0": 0, "SOLID": 1, "S": 1, "1": 1, "WALL": 2, "W": 2, "2": 2}
_COLOR = {"BLUE": 0, "B": 0, "0": 0, "RED": 1, "R": 1, "1": 1}


def _pair_cases(inp):
    lines = [l for l in open(inp).read().split("\n") if l.strip()]
    for k in range(len(lines) // 2):
        la, lb = lines[2 * k].split(), lines[2 * k + 1].split()
        a = np.array([float(x) for x in la[:6]], np.float32)
        b = np.array([float(x) for x in lb[:6]], np.float32)
        yield k, a, b, _KIND[la[6]], _KIND[lb[6]], _COLOR[la[7]], _COLOR[lb[7]]


def _param_index(ka, kb, ca, cb):
    """Parameter slot chosen by the pair harness (src/u/pair/main.cu:30-66)."""
    k0, k1 = min(ka, kb), max(ka, kb)
    if k0 == 0 and k1 in (1, 2):
        return (ca + 1) * (ca + 2) // 2 - 1       # mirrored view, colour of a
    if k0 == 0:
        c0, c1 = min(ca, cb), max(ca, cb)
        return c1 * (c1 + 1) // 2 + c0            # colour pair id
    if k0 == 1 and k1 == 1:
        return None                               # DPD + LJ: not on this path
    return 0                                      # grey view


@pytest.mark.parametrize("inp,ref,A,G", [
    ("pair_input_2.txt", "pair.ref.pair.0.txt", [2.6666667] * 3, [8.0] * 3),
    ("pair_input_rnd.txt", "pair.ref.pair.1.txt", [10.0, 30.0, 20.0], [100.0, 300.0, 200.0]),
])
def test_pair_force_golden(inp, ref, A, G):
    """pair.0 / pair.1: one pair_force per input pair with xi = 0 (no RND).
    Tolerance 2e-5 of |f|: the fixture prints 6 significant digits ("%g") and
    was computed with fast-math powf/rsqrtf (src/make/main.mk:27); one pair
    of pair.1 cancels its conservative and dissipative parts to |f| = 1.7
    and sits at 1.4e-5."""
    refl = [l for l in open(_gold(ref)).read().split("\n") if l.strip()]
    checked = 0
    for k, a, b, ka, kb, ca, cb in _pair_cases(_gold(inp)):
        j = _param_index(ka, kb, ca, cb)
        if j is None:
            continue
        f, _ = O.pair_force((np.float32(A[j]), np.float32(G[j]), 0.0, 0.25), a, b, 0.0)
        fr = np.array([float(x) for x in refl[k].split()])
        assert np.linalg.norm(f - fr) <= 2e-5 * max(np.linalg.norm(fr), 1e-6), (k, f, fr)
        checked += 1
    assert checked >= 2


def test_pair_force_closed_form():
    """f = [(-g w^2 (e.dv) + s xi w) + a (1-r)] e with w = (1-r)^(1/4)."""
    rng = np.random.default_rng(1)
    for _ in range(500):
        a = rng.uniform(-1, 1, 6).astype(np.float32)
        b = a + rng.uniform(-0.6, 0.6, 6).astype(np.float32)
        xi = np.float32(rng.uniform(-1.4, 1.4))
        p = (25.0, 4.5, 3.0, 0.25)
        f, is_pair = O.pair_force(p, a, b, xi)
        d = a[:3].astype(np.float64) - b[:3]
        r = np.linalg.norm(d)
        if r >= 1:
            assert not is_pair and not f.any()
            continue
        e = d / r
        dv = a[3:].astype(np.float64) - b[3:]
        w = (1 - r) ** 0.25
        f0 = (-p[1] * w * w * e.dot(dv) + p[2] * xi * w) + p[0] * (1 - r)
        want = f0 * e
        scale = p[0] + p[1] * np.abs(dv).sum() + p[2] * 1.5
        assert np.linalg.norm(f - want) <= 1e-5 * scale


# ------------------------------------------------------------ bulk forces
@pytest.mark.parametrize("name,n,scale", [("t0", 64, 1e4), ("t1", 640, 1e2)])
def test_bulk_forces_golden(name, n, scale):
    """bulk.t0/t1 (src/test/bulk/{0,1}/main): 4^3 box, gen_random, kBT = 0,
    a = 2.6666667, g = 8, spow = 0.25, bulk kernel only (no periodic images).
    The fixture is `uscale K ff.txt | sort` of r v f lines."""
    pp = O.gen_random(4, 4, 4, n)
    starts, counts, out, _ = O.clist_build((4, 4, 4), pp)
    par = (np.float32(2.6666667), 8.0, O.dpd_sigma(0.0, 8.0, 5e-4), 0.25)
    ff, _ = O.dpd_forces((4, 4, 4), out, starts, counts, par, np.zeros(27, np.float32), periodic=False)
    ref = np.loadtxt(_gold("ff.ref.bulk.%s.txt" % name)) / scale
    key = lambda row: tuple(np.round(np.asarray(row[:6], np.float64), 4))
    table = {key(r): r[6:] for r in ref}
    assert len(table) == n
    for p, f in zip(out, ff):
        fr = table[key(p)]
        assert np.linalg.norm(f - fr) <= 1e-5 * max(np.linalg.norm(fr), 1.0)


def _brute_forces(L, pp, par):
    """O(N^2) minimum-image DPD forces at xi = 0, float64."""
    L = np.asarray(L, np.float64)
    r = pp[:, :3].astype(np.float64)
    v = pp[:, 3:].astype(np.float64)
    f = np.zeros_like(r)
    for i in range(len(pp)):
        d = r[i] - r
        d -= L * np.round(d / L)
        r2 = (d * d).sum(1)
        m = (r2 < 1) & (np.arange(len(pp)) != i) & (r2 >= 1e-6)
        rr = np.sqrt(r2[m])
        e = d[m] / rr[:, None]
        ev = (e * (v[i] - v[m])).sum(1)
        w = (1 - rr) ** par[3]
        f0 = -par[1] * w * w * ev + par[0] * (1 - rr)
        f[i] = (f0[:, None] * e).sum(0)
    return f


@pytest.mark.parametrize("L", [(3, 3, 3), (4, 5, 6), (7, 3, 4)])
def test_periodic_forces_brute_force(L):
    """Full-shell periodic gather == all-pairs minimum image (xi = 0)."""
    rng = np.random.default_rng(sum(L))
    n = 3 * int(np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)
    pp[:, 3:] = rng.normal(0, 1, (n, 3))
    starts, counts, out, _ = O.clist_build(L, pp)
    par = (25.0, 4.5, 0.0, 0.25)
    ff, npair = O.dpd_forces(L, out, starts, counts, par, np.zeros(27, np.float32), periodic=True)
    want = _brute_forces(L, out, par)
    scale = np.sqrt((want ** 2).sum(1).mean())
    assert np.max(np.linalg.norm(ff - want, axis=1)) <= 1e-5 * scale
    # pair count: every unordered pair found from both sides
    r = out[:, :3].astype(np.float64)
    d = r[:, None] - r[None]
    d -= np.array(L) * np.round(d / np.array(L))
    r2 = (d * d).sum(-1)
    assert npair == int(((r2 < 1).sum() - n))


def test_newton_third_law_kbt0():
    """Without noise the per-pair forces are antisymmetric, so sum f = 0."""
    L = (6, 6, 6)
    rng = np.random.default_rng(9)
    n = 3 * 216
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-3, 3, (n, 3))
    pp[:, 3:] = rng.normal(0, 1, (n, 3))
    starts, counts, out, _ = O.clist_build(L, pp)
    ff, _ = O.dpd_forces(L, out, starts, counts, (25.0, 4.5, 0.0, 0.25), np.zeros(27, np.float32))
    assert np.abs(ff.astype(np.float64).sum(0)).max() < 1e-3 * np.abs(ff).max()


def _fragment_lists(L, starts, counts):
    """Packed fragment lists built the way the halo packer walks them
    (src/exch/flu/dev/map.h:1-98, src/exch/flu/dev/pack.h:2-40): the cells of
    fragment box (origin, extent) in x-fastest order, each cell's particles in
    array order.  Returns {fid: {particle: position}}."""
    pos = {}
    for fid in range(26):
        d = O.frag_i2d3(fid)
        org = [L[c] - 1 if d[c] == 1 else 0 for c in range(3)]
        ext = [L[c] if d[c] == 0 else 1 for c in range(3)]
        lst = []
        for z in range(ext[2]):
            for y in range(ext[1]):
                for x in range(ext[0]):
                    cid = (org[0] + x) + L[0] * ((org[1] + y) + L[1] * (org[2] + z))
                    lst.extend(range(starts[cid], starts[cid] + counts[cid]))
        pos[fid] = {p: k for k, p in enumerate(lst)}
    return pos


@pytest.mark.parametrize("L", [(4, 3, 5), (3, 3, 3)])
def test_wrapped_pair_noise_matches_fragment_lists(L):
    """Random force only (a = g = 0): each pair's xi comes from the bulk rule
    mean0var1ii(seed, max, min) or, across the periodic boundary, from the
    halo rule mean0var1uu(seed_fid, ...) with positions in the packed
    fragment lists (src/fluforces/halo/dev/main.h:7-12)."""
    rng = np.random.default_rng(42)
    n = 3 * int(np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)
    starts, counts, out, _ = O.clist_build(L, pp)
    seeds = O.seeds_table(3)[2]
    par = (0.0, 0.0, 7.0, 0.25)
    ff, _ = O.dpd_forces(L, out, starts, counts, par, seeds)
    frag = _fragment_lists(L, starts, counts)
    Lf = np.array(L, np.float64)
    r = out[:, :3].astype(np.float64)
    want = np.zeros((n, 3))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            m = np.round((r[i] - r[j]) / Lf).astype(int)   # image of j at r_j + m L
            d = r[i] - (r[j] + m * Lf)
            r2 = d.dot(d)
            if r2 >= 1 or r2 < 1e-6:
                continue
            if not m.any():
                xi = O.rng_ii(seeds[26], [max(i, j)], [min(i, j)])[0]
            else:
                fid = O.frag_d2i(*m)
                anti = O.frag_anti(fid)
                aid, bid = frag[fid][i], frag[anti][j]
                a1, a2 = (aid, bid) if fid < anti else (bid, aid)
                xi = O.rng_uu(seeds[fid], [a1], [a2])[0]
            rr = np.sqrt(r2)
            want[i] += par[2] * float(xi) * (1 - rr) ** 0.25 * d / rr
    scale = np.sqrt((want ** 2).sum(1).mean())
    assert np.max(np.linalg.norm(ff - want, axis=1)) <= 1e-5 * scale


@pytest.mark.parametrize("L,rho", [((9, 8, 7), 5), ((3, 4, 3), 3)])
def test_half_shell_traversal_matches_the_gather(L, rho):
    """The CPU baseline's force pass (the reference's half shell with
    reactions + one-sided wrapped pairs) against the oracle's full-shell
    gather: the same pairs, forces within FP32 summation order."""
    rng = np.random.default_rng(4)
    n = rho * int(np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    starts, counts, srt, _ = O.clist_build(L, pp)
    par = (25.0, 4.5, O.dpd_sigma(1.0, 4.5, 1e-3), 0.25)
    seeds = O.seeds_table(2)[1]
    f1, n1 = O.dpd_forces(L, srt, starts, counts, par, seeds)
    f2, n2 = O.dpd_forces_half(L, srt, starts, counts, par, seeds, nthreads=3)
    assert n1 == n2
    mag = np.linalg.norm(f1.astype(np.float64), axis=1)
    err = np.linalg.norm(f2.astype(np.float64) - f1, axis=1) / np.maximum(mag, 1e-2 * np.sqrt((mag ** 2).mean()))
    assert err.max() < 1e-5


@pytest.mark.parametrize("L", [(4, 5, 6), (3, 3, 4)])
def test_stress_brute_force(L):
    """Per-particle stress (src/pair/dev.h:87-100: s_ab = 1/2 f_a dr_b summed
    over the particle's pairs, both particles of a pair getting the same)
    against all-pairs minimum image in float64 (xi = 0); the forces equal
    the force-only gather bit for bit; tr s = 1/2 sum f0 r."""
    rng = np.random.default_rng(31 + sum(L))
    n = 4 * int(np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)
    pp[:, 3:] = rng.normal(0, 1, (n, 3))
    starts, counts, out, _ = O.clist_build(L, pp)
    par = (25.0, 4.5, 0.0, 0.25)
    z = np.zeros(27, np.float32)
    ff, ss, npair = O.dpd_forces_stress(L, out, starts, counts, par, z)
    f2, n2 = O.dpd_forces(L, out, starts, counts, par, z)
    assert npair == n2 and np.array_equal(ff, f2)
    r = out[:, :3].astype(np.float64)
    v = out[:, 3:].astype(np.float64)
    want = np.zeros((n, 6))
    tr = np.zeros(n)
    for i in range(n):
        d = r[i] - r
        d -= np.array(L) * np.round(d / np.array(L))
        r2 = (d * d).sum(1)
        m = (r2 < 1) & (np.arange(n) != i) & (r2 >= 1e-6)
        rr = np.sqrt(r2[m])
        e = d[m] / rr[:, None]
        ev = (e * (v[i] - v[m])).sum(1)
        w = (1 - rr) ** par[3]
        f0 = -par[1] * w * w * ev + par[0] * (1 - rr)
        f = f0[:, None] * e
        dm = d[m]
        want[i] = 0.5 * np.array([(f[:, 0] * dm[:, 0]).sum(), (f[:, 0] * dm[:, 1]).sum(), (f[:, 0] * dm[:, 2]).sum(),
                                  (f[:, 1] * dm[:, 1]).sum(), (f[:, 1] * dm[:, 2]).sum(), (f[:, 2] * dm[:, 2]).sum()])
        tr[i] = 0.5 * (f0 * rr).sum()
    scale = np.abs(want).max()
    assert np.abs(ss - want).max() <= 1e-5 * scale
    assert np.abs(ss[:, 0] + ss[:, 3] + ss[:, 5] - tr).max() <= 1e-5 * scale
