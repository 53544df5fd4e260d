"""GPU membrane bounce-back (udx_mbb_*) through the C ABI against the oracle.

1. The standalone bounce on the oracle's set-ups (cells straddling the
   periodic boundary, deforming over the step): the same particles bounce off
   the same triangles, and the bounced solvent, the triangle momenta and the
   vertex velocities equal the oracle's bit for bit.  The device's double
   acos/cos/pow may differ from the host's by an ulp in the cubic's roots;
   their float results have not differed in these cases.
2. The suspension step with bounce-back (exact mode): one step equals the
   oracle's solvent bit for bit with the same bounced count; vertices within
   1e-6 (membrane force pow/sqrt ulps).
3. Physics at a larger size: production steps of a 32^3 suspension with and
   without bounce-back; with it no solvent particle changes side of any
   membrane, without it some do.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from membranes import GOLD, bounce_case
from oracle import mbb as B
from oracle import rbc as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import bench  # noqa: E402
from ctc_amd import udx as U  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return U.Context()


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bounce_matches_oracle(ctx, seed):
    c = bounce_case(seed=seed)
    say("standalone bounce, seed", seed, "particles", len(c["pp"]))
    want_pp, want_rc, want_mm, want_col, want_n = B.bounce(c["dt"], 1.0, 0.5, c["L"], c["nm"], c["tt"], c["rp"],
                                                           c["rc"], c["starts"], c["pp0"], c["pp"])
    nm, nv = c["nm"], c["nv"]
    b = U.Mbb(ctx, nv, c["tt"], nm, len(c["pp"]))
    rp = np.zeros((nm * nv, 6), np.float32)
    rp[:, :3] = c["rp"]
    rc, pp = dev(c["rc"]), dev(c["pp"])
    mm = torch.zeros((nm * len(c["tt"]), 6), dtype=torch.float32, device="cuda")
    b.bounce(c["dt"], 1.0, 0.5, c["L"], nm, dev(rp), rc, dev(c["starts"], torch.int32), dev(c["pp0"]), pp, mm)
    col = host(b.collisions(len(c["pp"])))
    assert b.count() == want_n > 20
    assert np.array_equal(col, want_col)
    assert np.array_equal(bits(host(pp)), bits(want_pp))
    assert np.array_equal(bits(host(mm)), bits(want_mm))
    assert np.array_equal(bits(host(rc)), bits(want_rc))
    # a second call on the same handle (keys and counters reset) gives the same
    rc2, pp2 = dev(c["rc"]), dev(c["pp"])
    b.bounce(c["dt"], 1.0, 0.5, c["L"], nm, dev(rp), rc2, dev(c["starts"], torch.int32), dev(c["pp0"]), pp2)
    assert torch.equal(pp2, pp) and torch.equal(rc2, rc)
    # a pair list too short for the call: the in-place search gives the same
    b.set_pair_capacity(8)
    rc3, pp3 = dev(c["rc"]), dev(c["pp"])
    b.bounce(c["dt"], 1.0, 0.5, c["L"], nm, dev(rp), rc3, dev(c["starts"], torch.int32), dev(c["pp0"]), pp3)
    assert torch.equal(pp3, pp) and torch.equal(rc3, rc) and b.count() == want_n
    b.close()


def suspension_setup(L, lattice, seed, scale=1.0):
    rr, tt = U.read_off(os.path.join(GOLD, "cells", "498.off"))
    rr = (rr * np.float32(scale)).astype(np.float32)
    m = R.Mesh(rr, tt)
    opp = np.zeros((lattice[0] * lattice[1] * lattice[2] * m.nv, 6), np.float32)
    opp[:, :3] = bench.place_cells(rr, tt, L, lattice, seed)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, kb=40.0, totArea=np.float32(A0), totVolume=np.float32(V0))
    return m, opp, kw


def test_suspension_step_with_bounce(ctx):
    L, kBT, dt = (16, 16, 16), 0.0444, 5e-3
    m, opp, kw = suspension_setup(L, (2, 2, 2), 7)
    nc = len(opp) // m.nv
    rng = np.random.default_rng(11)
    n = 10 * L[0] * L[1] * L[2]
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, 1.0, (n, 3)).astype(np.float32)
    opp[:, 3:] = rng.normal(0, 0.05, (len(opp), 3)).astype(np.float32)
    dpd = (2.6666667, 8.0, O.dpd_sigma(kBT, 8.0, dt), 0.25)
    seeds = O.seeds_table(1)
    l0, a0 = m.stressful(kw["totArea"])
    say("suspension step with bounce: %d solvent, %d cells" % (n, nc))
    flu = U.Flu(ctx, L, n, U.PairDpd(*dpd), dt, 1.0, exact=True)
    flu.set_pp(dev(spp))
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    r.set_stressful(kw["totArea"])
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * m.nv)
    bb = U.Mbb(ctx, m.nv, m.tt, nc, n)
    vert = dev(opp)
    flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(*dpd), 0.5, 1, nc, vert)
    flu.set_rbc_bounce(bb)
    flu.run(1, seeds)
    sol = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    flu.get_pp(sol)
    gs, gv, gn = host(sol), host(vert), bb.count()
    flu.set_rbc(None)
    flu.close()
    r.close()
    rs, rv, rp, rf, rn = R.suspension_step(L, spp, opp, m, nc, dpd, seeds[0], dpd, R.fsi_seeds(1)[0], R.par(**kw),
                                           dt, l0=l0, a0=a0, nthreads=16, bounce=True)
    say("bounced: gpu %d oracle %d" % (gn, rn))
    assert gn == rn > 0
    assert np.array_equal(bits(gs), bits(rs))
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-6 * L[0]
    assert np.abs(gv[:, 3:] - rv[:, 3:]).max() <= 1e-5 * max(np.abs(rv[:, 3:]).max(), 1.0)
    bb.close()


def sides(pts, opp, nv, tt, L):
    """for each point: the index of the cell (or its periodic image) it is
    inside, -1 outside all"""
    out = np.full(len(pts), -1, np.int64)
    Lv = np.array(L, np.float64)
    order = np.argsort(pts[:, 0], kind="stable")
    xs = pts[order, 0]
    for k in range(len(opp) // nv):
        x = opp[k * nv:(k + 1) * nv, :3].astype(np.float64)
        for s in np.stack(np.meshgrid(*[[-1, 0, 1]] * 3, indexing="ij"), -1).reshape(-1, 3):
            y = x + s * Lv
            lo, hi = y.min(0) - 0.1, y.max(0) + 0.1
            i0, i1 = np.searchsorted(xs, lo[0]), np.searchsorted(xs, hi[0], side="right")
            if i1 <= i0:
                continue
            cand = order[i0:i1]
            q = pts[cand]
            sel = cand[np.all((q[:, 1:] >= lo[1:]) & (q[:, 1:] <= hi[1:]), 1)]
            if len(sel):
                ins = bench.inside_mesh(pts[sel], y, tt)
                out[sel[ins]] = k
    return out


@pytest.mark.parametrize("bounce", [True, False])
def test_no_solvent_crosses_membranes(ctx, bounce):
    """4 production steps of a 32^3 suspension (27 cells, dt 5e-4, solvent
    started hot, sigma_v = 2, so that a step has ~15 crossings): with
    bounce-back every particle stays on its side of every membrane; without
    it some cross"""
    L, kBT, dt, nsteps = (32, 32, 32), 0.0444, 5e-4, 4
    m, opp, kw = suspension_setup(L, (3, 3, 3), 5)
    nc = len(opp) // m.nv
    rng = np.random.default_rng(21)
    n = 10 * L[0] * L[1] * L[2]
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, 2.0, (n, 3)).astype(np.float32)
    say("crossing test, bounce", bounce, ": %d solvent, %d cells, %d steps" % (n, nc, nsteps))
    dpd = (2.6666667, 8.0, O.dpd_sigma(kBT, 8.0, dt), 0.25)
    flu = U.Flu(ctx, L, n, U.PairDpd(*dpd), dt, 1.0)
    flu.set_pp(dev(spp))
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    r.set_stressful(kw["totArea"])
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * m.nv)
    bb = U.Mbb(ctx, m.nv, m.tt, nc, n)
    vert = dev(opp)
    flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(*dpd), 0.5, 1, nc, vert)
    if bounce:
        flu.set_rbc_bounce(bb)
    sol = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    flu.get_pp(sol)
    a = host(sol).copy()
    changed = bounced = 0
    for k in range(nsteps):
        va = host(vert).copy()
        flu.run(1)
        flu.get_pp(sol)
        b, vb = host(sol).copy(), host(vert).copy()
        # the membranes move little in a step (the bounce searches 0.1 around
        # each triangle's start); whole-cell periodic shifts aside
        dx = vb[:, :3] - va[:, :3]
        dx -= np.round(dx / np.array(L)) * np.array(L)
        assert np.abs(dx).max() < 0.02
        # the step's output is in the cell order of its start positions
        _, _, _, order = O.clist_build(L, a)
        s0 = sides(a[order, :3].astype(np.float64), va, m.nv, m.tt, L)
        s1 = sides(b[:, :3].astype(np.float64), vb, m.nv, m.tt, L)
        changed += int((s0 != s1).sum())
        bounced += bb.count() if bounce else 0
        say("step %d: %d changed side, %d bounced, largest vertex move %.3g" % (k, changed, bounced,
                                                                                np.abs(dx).max()))
        a = b
    flu.set_rbc(None)
    flu.close()
    r.close()
    bb.close()
    if bounce:
        assert bounced > 30 and changed == 0
    else:
        assert changed > 30


def test_contact_matches_oracle(ctx):
    """udx_cnt_apply_seed against the contact oracle on cells that touch and
    straddle the periodic faces: forces added to a nonzero array bit for bit,
    the same pair count"""
    from test_oracle_cnt import PAR, contact_case
    L = (16, 16, 16)
    pp = contact_case(2, L)
    rng = np.random.default_rng(3)
    ff0 = rng.normal(0, 1, (len(pp), 3)).astype(np.float32)
    sb, sh = np.float32(0.123), np.float32(0.456)
    want, npair = R.cnt_forces(L, pp, PAR, sb, sh, ff=ff0)
    c = U.Cnt(ctx, L, len(pp))
    ffd = dev(ff0)
    c.apply(U.PairDpdLj(*PAR), dev(pp), ffd, seeds=(sb, sh))
    assert c.pairs() == npair > 1000
    assert np.array_equal(bits(host(ffd)), bits(want))
    # a second call (counters reset by the first) gives the same forces
    ffd2 = dev(ff0)
    c.apply(U.PairDpdLj(*PAR), dev(pp), ffd2, seeds=(sb, sh))
    assert torch.equal(ffd2, ffd)
    c.close()


def test_suspension_step_with_contact(ctx):
    """one exact-mode suspension step with membrane contact: vertices within
    1e-6 of the oracle (membrane force ulps), and the contact moves them by
    more than that"""
    from test_oracle_cnt import PAR
    L, kBT, dt = (16, 16, 16), 0.0444, 5e-3
    m, opp, kw = suspension_setup(L, (2, 2, 2), 7)
    nc = len(opp) // m.nv
    # two cells pressed together so that contact pairs exist
    opp[:m.nv, 0] += np.float32(2.2)
    rng = np.random.default_rng(11)
    n = 10 * L[0] * L[1] * L[2]
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, np.sqrt(kBT), (n, 3)).astype(np.float32)
    dpd = (2.6666667, 8.0, O.dpd_sigma(kBT, 8.0, dt), 0.25)
    seeds = O.seeds_table(1)
    l0, a0 = m.stressful(kw["totArea"])
    say("suspension step with contact: %d solvent, %d cells" % (n, nc))
    flu = U.Flu(ctx, L, n, U.PairDpd(*dpd), dt, 1.0, exact=True)
    flu.set_pp(dev(spp))
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    r.set_stressful(kw["totArea"])
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * m.nv)
    cnt = U.Cnt(ctx, L, nc * m.nv)
    vert = dev(opp)
    flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(*dpd), 0.5, 1, nc, vert)
    flu.set_rbc_contact(cnt, U.PairDpdLj(*PAR))
    flu.run(1, seeds)
    gv, gpairs = host(vert), cnt.pairs()
    flu.set_rbc(None)
    flu.close()
    r.close()
    cs = O.kiss_fill(7119, 187, 18278, 15674, 2)
    args = (L, spp, opp, m, nc, dpd, seeds[0], dpd, R.fsi_seeds(1)[0], R.par(**kw), dt)
    _, rv, _, _ = R.suspension_step(*args, l0=l0, a0=a0, nthreads=16, cnt_par=PAR, cnt_seeds=cs)
    _, rv0, _, _ = R.suspension_step(*args, l0=l0, a0=a0, nthreads=16)
    say("contact pairs %d; vertex shift by contact %.3g" % (gpairs, np.abs(rv[:, :3] - rv0[:, :3]).max()))
    assert gpairs > 0
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-6 * L[0]
    assert np.abs(gv[:, 3:] - rv[:, 3:]).max() <= 1e-5 * max(np.abs(rv[:, 3:]).max(), 1.0)
    assert np.abs(rv[:, 3:] - rv0[:, 3:]).max() > 1e-3
    cnt.close()
