"""Solvent, membranes and FSI in one step (udx_flu_set_rbc) against the
oracle's suspension step (oracle/rbc.py suspension_step) on the same seeds.

Two 498-vertex cells scaled to a diameter of 7.8 in a 16^3 box at density 10:
one whose bounding-box centre starts outside the box (the first step's
redistribution shifts it by a box length and moves it after the other cell)
and one that crosses the y boundary (FSI through periodic images).  Exact
mode: the solvent DPD and FSI kernels are bit-exact with the oracle, the
membrane force is within the float rounding of its double sum (device and
host pow/sqrt may differ by an ulp), so after one step the solvent state is
identical bit for bit and the vertices within 1e-6; after three steps both
agree within 1e-5 of the box scale (positions) and of the largest velocity."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import rbc as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from ctc_amd import udx as U  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return U.Context()


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def case(L=(16, 16, 16), rho=10, seed=11):
    rr, tt = R.read_off(os.path.join(GOLD, "cells", "498.off"))
    m = R.Mesh((rr * 1.5).astype(np.float32), tt)
    rng = np.random.default_rng(seed)
    T = []
    for centre in ([0.5 * L[0] + 0.3, -2.0, 1.0], [-3.0, 0.5 * L[1] - 1.5, -2.5]):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        A = np.eye(4)
        A[:3, :3] = q * np.sign(np.linalg.det(q))
        A[:3, 3] = centre
        T.append(A)
    opp = R.place(m.rr, T, (0, 0, 0))
    opp[:, 3:] = rng.normal(0, 0.05, (len(opp), 3)).astype(np.float32)
    n = int(rho * np.prod(L))
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, 0.21, (n, 3)).astype(np.float32)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, kb=40.0, totArea=np.float32(A0), totVolume=np.float32(V0))
    l0, a0 = m.stressful(kw["totArea"])
    return L, m, spp, opp, kw, l0, a0


KBT, DT = 0.0444, 5e-4


def run_gpu(ctx, L, m, spp, opp, kw, nsteps, seeds, dpd, fsi_par, substeps, exact=True):
    nc = len(opp) // m.nv
    flu = U.Flu(ctx, L, len(spp), U.PairDpd(*dpd), DT, 1.0, exact=exact)
    flu.set_pp(dev(spp))
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    r.set_stressful(kw["totArea"])
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * m.nv)
    vert = dev(opp)
    flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(*fsi_par), 0.5, substeps, nc, vert)
    flu.run(nsteps, seeds)
    sol = torch.zeros((len(spp), 6), dtype=torch.float32, device="cuda")
    flu.get_pp(sol)
    out = host(sol), host(vert), flu.pairs(), fsi.pairs()
    flu.set_rbc(None)
    flu.close()
    r.close()
    return out


def run_oracle(L, m, spp, opp, kw, l0, a0, nsteps, seeds, dpd, fsi_par, substeps):
    nc = len(opp) // m.nv
    fseeds = R.fsi_seeds(nsteps)
    p = R.par(**kw)
    npair = nfsi = 0
    for k in range(nsteps):
        spp, opp, a, b = R.suspension_step(L, spp, opp, m, nc, dpd, seeds[k], fsi_par, fseeds[k], p, DT,
                                           substeps=substeps, l0=l0, a0=a0, nthreads=8)
        npair += a
        nfsi += b
    return spp, opp, npair, nfsi


@pytest.mark.parametrize("nsteps,substeps", [(1, 1), (3, 2)])
def test_suspension_step(ctx, nsteps, substeps):
    L, m, spp, opp, kw, l0, a0 = case()
    say("suspension: %d steps, %d membrane sub-steps" % (nsteps, substeps))
    dpd = (2.5, 8.0, O.dpd_sigma(KBT, 8.0, DT), 0.25)
    fsi_par = (2.6666667, 8.0, O.dpd_sigma(KBT, 8.0, DT), 0.25)
    seeds = O.seeds_table(nsteps)
    gs, gv, gpairs, gfsi = run_gpu(ctx, L, m, spp, opp, kw, nsteps, seeds, dpd, fsi_par, substeps)
    rs, rv, rpairs, rfsi = run_oracle(L, m, spp, opp, kw, l0, a0, nsteps, seeds, dpd, fsi_par, substeps)
    assert gpairs == rpairs and gfsi == rfsi
    # the first cell's centre started outside: it was shifted by -L x and now comes last
    assert np.abs(gv[m.nv:, 0] - (opp[:m.nv, 0] - L[0])).max() < 0.1
    vmax = np.abs(rv[:, 3:]).max()
    if nsteps == 1:
        assert np.array_equal(gs.view(np.uint32), rs.view(np.uint32))
        assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-6
        assert np.abs(gv[:, 3:] - rv[:, 3:]).max() <= 1e-6 * vmax
    else:
        assert np.abs(gs[:, :3] - rs[:, :3]).max() <= 1e-5 * L[0]
        assert np.abs(gs[:, 3:] - rs[:, 3:]).max() <= 1e-5 * np.abs(rs[:, 3:]).max()
        assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-5 * L[0]
        assert np.abs(gv[:, 3:] - rv[:, 3:]).max() <= 1e-5 * vmax


def test_suspension_rejects_bad_calls(ctx):
    L, m, spp, opp, kw, l0, a0 = case()
    flu = U.Flu(ctx, L, len(spp), U.PairDpd(2.5, 8.0, 1.0, 0.25), DT, 1.0)
    r = U.Rbc(ctx, m.tt, m.nv, 1)
    fsi = U.Fsi(ctx, L, m.nv)
    with pytest.raises(U.UdxError, match="capacity"):
        flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(2.6, 8.0, 1.0, 0.25), 0.5, 1, 2, dev(opp))
    flu.close()
    r.close()


@pytest.mark.parametrize("nc", [7, 1500])
def test_distribute_matches_oracle(ctx, nc):
    """udx_rbc_distribute against the oracle's redistribution: random cell
    centres in and around the box (every direction), more cells than one
    1024-thread chunk of the ordering kernel; bit-identical output."""
    L = (16, 12, 20)
    rr, tt = R.read_off(os.path.join(GOLD, "cells", "498.off"))
    m = R.Mesh(rr, tt)
    rng = np.random.default_rng(nc)
    T = []
    for _ in range(nc):
        A = np.eye(4)
        A[:3, 3] = rng.uniform(-0.75, 0.75, 3) * np.array(L)
        T.append(A)
    pp = R.place(m.rr, T, (0, 0, 0))
    pp[:, 3:] = rng.normal(0, 1, (len(pp), 3)).astype(np.float32)
    want, fid = R.redistribute(L, pp, nc, m.nv)
    assert (fid != 26).sum() > 0 and (fid == 26).sum() > 0
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    d = dev(pp)
    r.distribute(L, nc, d)
    got = host(d)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    r.close()


@pytest.mark.parametrize("lam,bounce,nsteps", [(0.0, False, 1), (2.0, False, 1), (0.0, True, 1), (2.0, False, 3)])
def test_suspension_step_between_walls(ctx, lam, bounce, nsteps):
    """solvent + membranes in a plane channel (|z| < 9 of a 16 x 16 x 24 box),
    cells close to a wall: one exact-mode step from the internal seed trunks
    (the wall trunk drawing for the solvent, then for the vertices) against
    the oracle: the solvent bit for bit, the vertices within 1e-6, the same
    solvent-wall, vertex-wall and FSI pair counts; with the membrane
    bounce-back also the same bounced particles (hotter solvent, longer step).
    Three steps: the wall trunk's draws alternate solvent, vertices, solvent,
    ... over the steps; positions within 1e-5 of the box, velocities within
    1e-5 of the largest."""
    from oracle import wall as W
    from membranes import mesh as tmesh, rotation
    L, h, kBT = (16, 16, 24), 9.0, 0.0444
    dt, vsig = (5e-3, 1.0) if bounce else (2e-3, np.sqrt(kBT))
    m = tmesh(498)
    rng = np.random.default_rng(5)
    # the first cell 0.4 from the lower wall, the second 0.45 from the upper
    # one and across the x boundary, the third in the middle
    centres = [(-3.0, -3.0, None, 0.4), (7.5, 3.0, None, -0.45), (0.0, 4.0, 0.0, 0.0)]
    opp = np.zeros((len(centres) * m.nv, 6), np.float32)
    for k, (cx, cy, cz, gap) in enumerate(centres):
        r = m.rr.astype(np.float64) @ rotation(rng).T
        if cz is None:
            cz = -h + gap - r[:, 2].min() if gap > 0 else h + gap - r[:, 2].max()
        opp[k * m.nv:(k + 1) * m.nv, :3] = (r + np.array([cx, cy, cz])).astype(np.float32)
    opp[:, 3:] = rng.normal(0, 0.05, (len(opp), 3)).astype(np.float32)
    nc = len(centres)
    ext, N = np.array(L, np.float32), (48, 48, 48)
    D = W.plane_channel_sdf(L, 48, h)
    q = U.Sdf(ctx, L, ext=ext, N=N, D=D)
    ref = W.Sdf(L, ext, N, D.ravel())
    n0 = 10 * L[0] * L[1] * L[2]
    pp = np.zeros((n0, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n0, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, vsig, (n0, 3)).astype(np.float32)
    ppd = dev(pp)
    w = U.Wall(ctx, q, ppd, exact=True)
    solvent = host(ppd)[:w.n_solvent].copy()
    keep, frozen = ref.freeze(pp)
    wst, _, wpp, _ = O.clist_build(W.margin_box(L), W.wall_images(L, frozen), sub=1)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, kb=40.0, totArea=np.float32(A0), totVolume=np.float32(V0))
    l0, a0 = m.stressful(kw["totArea"])
    dpd = (2.6666667, 8.0, O.dpd_sigma(kBT, 8.0, dt), 0.25)
    say("suspension between walls: %d solvent, %d cells, lambda %g, bounce %d" % (len(solvent), nc, lam, bounce))
    flu = U.Flu(ctx, L, len(solvent), U.PairDpd(*dpd), dt, 1.0, exact=True)
    sd = dev(solvent)
    flu.set_pp(sd)
    flu.set_wall(w)
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    r.set_stressful(kw["totArea"])
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * m.nv)
    vert = dev(opp)
    flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(*dpd), 0.5, 1, nc, vert)
    flu.set_rbc_repulsion(lam)
    if bounce:
        bb = U.Mbb(ctx, m.nv, m.tt, nc, len(solvent))
        flu.set_rbc_bounce(bb)
    w.pairs(reset=True)
    flu.run(nsteps)
    out = torch.zeros_like(sd)
    nout = flu.get_pp(out)
    gs, gv = host(out)[:nout], host(vert)
    gp, gw, gwm, gf = flu.pairs(), flu.wall_pairs(), w.pairs(), fsi.pairs()
    gn = bb.count() if bounce else 0
    flu.set_rbc(None)
    flu.close()
    r.close()
    wseeds = O.kiss_fill(42, 42, 42, 42, 2 * nsteps)
    table, fseeds = O.seeds_table(nsteps), R.fsi_seeds(nsteps)
    rs, rv, rp, rf, rw, rwm, rn = solvent, opp, 0, 0, 0, 0, 0
    for k in range(nsteps):
        seeds28 = np.concatenate([table[k], [wseeds[2 * k]]]).astype(np.float32)
        out = R.suspension_wall_step(L, rs, rv, m, nc, dpd, seeds28, dpd, fseeds[k], np.float32(wseeds[2 * k + 1]),
                                     R.par(**kw), dt, ref, wpp, wst, lam=lam, l0=l0, a0=a0, nthreads=16,
                                     bounce=bounce)
        rs, rv = out[0], out[1]
        rp, rf, rw, rwm = rp + out[2], rf + out[3], rw + out[4], rwm + out[5]
        rn += out[6] if bounce else 0
    say("pairs gpu", gp, gw, gwm, gf, gn, "oracle", rp, rw, rwm, rf, rn)
    assert (gp, gw, gwm, gf, gn) == (rp, rw, rwm, rf, rn) and rwm > 0 and (rn > 0 or not bounce)
    if nsteps == 1:
        assert np.array_equal(gs.view(np.uint32), rs.view(np.uint32))
        assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-6 * L[0]
    else:
        assert np.abs(gs[:, :3] - rs[:, :3]).max() <= 1e-5 * L[0]
        assert np.abs(gs[:, 3:] - rs[:, 3:]).max() <= 1e-5 * np.abs(rs[:, 3:]).max()
        assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-5 * L[0]
    assert np.abs(gv[:, 3:] - rv[:, 3:]).max() <= 1e-5 * max(np.abs(rv[:, 3:]).max(), 1.0)
    assert np.all(ref.eval(gv[:, :3]) < 0)
    if bounce:
        bb.close()
    w.close()
    q.close()


def _octahedron(scale=1.5):
    rr = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], np.float32) * scale
    tt = np.array([(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)], np.int32)
    return R.Mesh(rr, tt)


def test_switching_to_a_mesh_with_more_cells(ctx):
    """udx_flu_set_rbc with a 498-vertex mesh (2 cells), then with an
    octahedron (6 vertices, 100 cells: fewer vertices in all, more cells than
    the per-cell arrays held): the per-cell arrays are reallocated, and one
    exact step equals the same step of a solvent that only ever had the
    octahedra."""
    L = (16, 16, 16)
    rng = np.random.default_rng(21)
    n = 10 * 16 ** 3
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, 0.21, (n, 3)).astype(np.float32)
    big = R.Mesh(*R.read_off(os.path.join(GOLD, "cells", "498.off")))
    octa = _octahedron()
    nc = 100
    T = []
    for _ in range(nc):
        A = np.eye(4)
        A[:3, 3] = rng.uniform(-0.5, 0.5, 3) * np.array(L)
        T.append(A)
    opp = R.place(octa.rr, T, (0, 0, 0))
    opp[:, 3:] = rng.normal(0, 0.05, (len(opp), 3)).astype(np.float32)
    bpp = R.place(big.rr, [np.eye(4), np.eye(4)], (0, 0, 0))
    tmpl = np.zeros((octa.nv, 6), np.float32)
    tmpl[:, :3] = octa.rr
    A0, V0 = R.area_volume(octa, tmpl, 1)
    kw = dict(R.RBC_TEST, kb=40.0, totArea=np.float32(A0), totVolume=np.float32(V0))
    dpd = (2.5, 8.0, O.dpd_sigma(KBT, 8.0, DT), 0.25)
    fsi_par = U.PairDpd(2.6666667, 8.0, O.dpd_sigma(KBT, 8.0, DT), 0.25)
    seeds = O.seeds_table(1)
    outs = []
    for switch in (True, False):
        say("switch meshes: %s" % switch)
        flu = U.Flu(ctx, L, n, U.PairDpd(*dpd), DT, 1.0, exact=True)
        flu.set_pp(dev(spp))
        rb = U.Rbc(ctx, big.tt, big.nv, 2)
        ro = U.Rbc(ctx, octa.tt, octa.nv, nc)
        ro.set_stressful(kw["totArea"])
        fsi = U.Fsi(ctx, L, 2 * big.nv)
        bvert = dev(bpp)
        vert = dev(opp)
        if switch:
            flu.set_rbc(rb, fsi, U.rbc_params(**kw), fsi_par, 0.5, 1, 2, bvert)
        flu.set_rbc(ro, fsi, U.rbc_params(**kw), fsi_par, 0.5, 1, nc, vert)
        flu.run(1, seeds)
        sol = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
        flu.get_pp(sol)
        outs.append((host(sol), host(vert)))
        flu.set_rbc(None)
        flu.close()
        rb.close()
        ro.close()
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))


def test_zero_cells_leave_the_wall_trunk_alone(ctx):
    """With a wall and a membrane handle that holds no cells, the reference
    skips the vertices' wall force and its seed draw
    (src/sim/walls/imp/main.h:54-58): three steps from the internal trunks
    equal three steps of the same solvent with no membranes, bit for bit."""
    from oracle import wall as W
    L, h = (16, 16, 24), 9.0
    rng = np.random.default_rng(8)
    ext, N = np.array(L, np.float32), (48, 48, 48)
    q = U.Sdf(ctx, L, ext=ext, N=N, D=W.plane_channel_sdf(L, 48, h))
    n0 = 10 * L[0] * L[1] * L[2]
    pp = np.zeros((n0, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n0, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 0.21, (n0, 3)).astype(np.float32)
    ppd = dev(pp)
    w = U.Wall(ctx, q, ppd, exact=True)
    solvent = host(ppd)[:w.n_solvent].copy()
    m = _octahedron()
    dpd = (2.6666667, 8.0, O.dpd_sigma(KBT, 8.0, DT), 0.25)
    outs = []
    for with_rbc in (True, False):
        say("wall, zero cells: membranes attached %s" % with_rbc)
        flu = U.Flu(ctx, L, len(solvent), U.PairDpd(*dpd), DT, 1.0, exact=True)
        sd = dev(solvent)
        flu.set_pp(sd)
        flu.set_wall(w)
        r = U.Rbc(ctx, m.tt, m.nv, 4)
        fsi = U.Fsi(ctx, L, 4 * m.nv)
        if with_rbc:
            flu.set_rbc(r, fsi, U.rbc_params(**R.RBC_TEST), U.PairDpd(*dpd), 0.5, 1, 0, None)
        flu.run(3)
        out = torch.zeros_like(sd)
        nout = flu.get_pp(out)
        outs.append(host(out)[:nout])
        flu.set_rbc(None)
        flu.close()
        r.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    w.close()
    q.close()
