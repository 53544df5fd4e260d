"""Parity at the benchmark configurations' full sizes, through the C ABI.

- C1 (periodic 16^3, density 3, a = 75, gamma = 4.5, kBT = 1, dt = 1e-3,
  spow = 0.25): the full 1,000 steps in exact mode equal the oracle bit for
  bit, with the same pair count (ordered pairs over the run, halved once).
- C2 (wall-bounded 96^3, density 10, plane channel |z| - 40 from a 128^3
  field): the frozen walls of 8,847,360 particles and one exact-mode step
  (cell list, solvent DPD, wall DPD, body force, kick-drift, bounce, wrap)
  equal the oracle bit for bit (state and forces); the production force
  kernel on the same state is within the contract metric
  |df| / max(|f|, 1e-3 F_rms) <= 1e-5 per particle with equal pair counts;
  the half-shell mode equals the oracle's exact sums bit for bit; then 20
  production-mode steps leave no solvent particle inside the walls.
- C3 (64^3 suspension, 567 cells of 498 vertices): one exact-mode step of
  solvent + membranes + FSI; the solvent equals the oracle bit for bit and the
  vertices agree within 1e-6 (the membrane force's pow/sqrt ulps).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import rbc as R
from oracle import wall as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import bench  # noqa: E402
from ctc_amd import udx as U  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")
THREADS = 16


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


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_c1_1000_steps_bit_exact(ctx):
    L, n = (16, 16, 16), 3 * 16 ** 3
    rng = np.random.default_rng(1)
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-8, 8, (n, 3)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    pp[:, 3:] -= pp[:, 3:].mean(0)
    dt, nsteps = 1e-3, 1000
    params = (75.0, 4.5, O.dpd_sigma(1.0, 4.5, dt), 0.25)
    seeds = O.seeds_table(nsteps)
    say("C1: %d particles, %d steps" % (n, nsteps))
    flu = U.Flu(ctx, L, n, U.PairDpd(*params), dt, 1.0, exact=True)
    d = dev(pp)
    flu.set_pp(d)
    flu.run(nsteps, seeds)
    out = torch.zeros_like(d)
    flu.get_pp(out)
    got, pairs = host(out), flu.pairs()
    flu.close()
    # the step of O.flu_step with the ordered pair count kept: pairs across
    # the periodic boundary are tested on each side with its own shifted
    # coordinates, so a pair at the cutoff can count on one side only and a
    # step's ordered count can be odd; the library halves the total once
    ref, ordered = pp.copy(), 0
    for k in range(nsteps):
        starts, counts, srt, _ = O.clist_build(L, ref)
        ff, npair = O.dpd_forces(L, srt, starts, counts, params, seeds[k], nthreads=THREADS)
        ref = O.integrate(L, srt, ff, dt)
        ordered += npair
    assert pairs == ordered // 2
    assert np.array_equal(bits(got), bits(ref))


def test_c2_full_size_step_bit_exact(ctx):
    L, h, rho, dt = (96, 96, 96), 40.0, 10, 5e-4
    N = (128, 128, 128)
    D = bench.plane_channel(L, 128, h)
    ext = np.array(L, np.float32)
    rng = np.random.default_rng(1234)
    n = rho * L[0] * L[1] * L[2]
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, np.sqrt(0.0444), (n, 3)).astype(np.float32)
    say("C2: %d particles" % n)
    q = U.Sdf(ctx, L, ext=ext, N=N, D=D)
    d = dev(pp)
    w = U.Wall(ctx, q, d, exact=True)
    solvent = host(d)[:w.n_solvent].copy()
    ref = W.Sdf(L, ext, N, D.ravel())
    keep, frozen = ref.freeze(pp)
    allw = W.wall_images(L, frozen)
    wst, _, wpp, _ = O.clist_build(W.margin_box(L), allw, sub=1)
    assert np.array_equal(bits(solvent), bits(keep))
    assert np.array_equal(bits(host(w.particles())), bits(wpp))
    assert np.array_equal(host(w.starts()), wst)
    say("C2: walls equal; one exact step")
    params = (2.5, 8.0, O.dpd_sigma(0.0444, 8.0, dt), 0.25)
    accel = (0.01, 0.0, 0.0)
    seeds = O.seeds_table(1)
    seeds28 = np.concatenate([seeds, O.kiss_fill(42, 42, 42, 42, 1)[:, None]], axis=1).astype(np.float32)
    sd = dev(solvent)
    gpu = {}
    for mode in (U.MODE_EXACT, U.MODE_FAST, U.MODE_HALF):
        flu = U.Flu(ctx, L, len(solvent), U.PairDpd(*params), dt, 1.0, accel, mode=mode)
        flu.set_pp(sd)
        flu.set_wall(w)
        flu.run(1, seeds28)
        out = torch.zeros_like(sd)
        m = flu.get_pp(out)
        ff = torch.zeros((len(solvent), 3), dtype=torch.float32, device="cuda")
        flu.get_ff(ff)
        gpu[mode] = (host(out)[:m], host(ff)[:m], flu.pairs(), flu.wall_pairs())
        flu.close()
    # the oracle step, keeping its forces (solvent DPD + wall DPD)
    starts, counts, srt, _ = O.clist_build(L, keep)
    ff, npair = O.dpd_forces(L, srt, starts, counts, params, seeds28[0, :27], nthreads=THREADS)
    ff, nw = W.wall_forces(L, srt, ref, wpp, wst, params, seeds28[0, 27], ff)
    want = W.integrate_wall(L, srt, ff, dt, ref, 1.0, accel)
    got, gff, gp, gw = gpu[U.MODE_EXACT]
    assert (gp, gw) == (npair // 2, nw) and nw > 0
    assert np.array_equal(bits(gff), bits(ff))
    assert np.array_equal(bits(got), bits(want))
    # the production kernel (the one the benchmark times) on the same state:
    # per particle |df| / max(|f|, 1e-3 F_rms) <= 1e-5, equal pair counts
    _, fff, fp, fw = gpu[U.MODE_FAST]
    assert (fp, fw) == (npair // 2, nw)
    mag = np.linalg.norm(ff.astype(np.float64), axis=1)
    frms = np.sqrt((mag ** 2).mean())
    err = np.linalg.norm(fff.astype(np.float64) - ff, axis=1) / np.maximum(mag, 1e-3 * frms)
    say("C2 production kernel: max rel err %.3g (F_rms %.4g, min |f| %.3g, %d particles)"
        % (err.max(), frms, mag.min(), len(mag)))
    assert err.max() <= 1e-5
    # the half-shell mode: the oracle's exact-sum solvent forces plus the same
    # wall forces bit for bit, equal pair counts, and the contract metric
    # against the double sum of the same pair forces
    _, fhh, hp, hw = gpu[U.MODE_HALF]
    assert (hp, hw) == (npair // 2, nw)
    fq0, nq = O.dpd_forces_sum64(L, srt, starts, counts, params, seeds28[0, :27], nthreads=THREADS)
    assert nq == npair
    fq, _ = W.wall_forces(L, srt, ref, wpp, wst, params, seeds28[0, 27], fq0)
    assert np.array_equal(bits(fhh), bits(fq))
    fd, _ = O.dpd_forces_f64(L, srt, starts, counts, params, seeds28[0, :27], nthreads=THREADS)
    fd += fq.astype(np.float64) - fq0   # the wall forces
    magd = np.linalg.norm(fd, axis=1)
    frmsd = np.sqrt((magd ** 2).mean())
    errh = np.linalg.norm(fhh.astype(np.float64) - fd, axis=1) / np.maximum(magd, 1e-3 * frmsd)
    errf = np.linalg.norm(fff.astype(np.float64) - fd, axis=1) / np.maximum(magd, 1e-3 * frmsd)
    say("C2 half-shell mode: max rel err %.3g vs the double sum (the gather: %.3g)" % (errh.max(), errf.max()))
    assert errh.max() <= 1e-5
    say("C2: 20 production steps")
    flu = U.Flu(ctx, L, len(solvent), U.PairDpd(*params), dt, 1.0, accel)
    flu.set_pp(sd)
    flu.set_wall(w)
    flu.run(20)
    out = torch.zeros_like(sd)
    m = flu.get_pp(out)
    fin = host(out)[:m]
    flu.close()
    s = ref.eval(fin[:, :3])
    assert np.all(s < 0), int((s >= 0).sum())
    w.close()
    q.close()


def test_c3_full_size_step(ctx):
    L, kBT, dt = (64, 64, 64), 0.0444, 5e-4
    rr, tt = U.read_off(os.path.join(GOLD, "cells", "498.off"))
    rr = (rr * np.float32(1.5)).astype(np.float32)
    m = R.Mesh(rr, tt)
    opp = np.zeros((567 * m.nv, 6), np.float32)
    opp[:, :3] = bench.place_cells(rr, tt, L, (9, 9, 7), 2024)
    nc = len(opp) // m.nv
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, kb=40.0, totArea=np.float32(A0), totVolume=np.float32(V0))
    l0, a0 = m.stressful(kw["totArea"])
    rng = np.random.default_rng(99)
    n = 10 * L[0] * L[1] * L[2]
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, np.sqrt(kBT), (n, 3)).astype(np.float32)
    dpd = (2.6666667, 8.0, O.dpd_sigma(kBT, 8.0, dt), 0.25)
    fsi_par = dpd
    seeds = O.seeds_table(1)
    say("C3: %d solvent, %d cells" % (n, nc))
    flu = U.Flu(ctx, L, n, U.PairDpd(*dpd), dt, 1.0, exact=True)
    flu.set_pp(dev(spp))
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    r.set_stressful(kw["totArea"])
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * m.nv)
    vert = dev(opp)
    flu.set_rbc(r, fsi, U.rbc_params(**kw), U.PairDpd(*fsi_par), 0.5, 1, nc, vert)
    flu.run(1, seeds)
    sol = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    flu.get_pp(sol)
    gs, gv, gp, gf = host(sol), host(vert), flu.pairs(), fsi.pairs()
    flu.set_rbc(None)
    flu.close()
    r.close()
    rs, rv, rp, rf = R.suspension_step(L, spp, opp, m, nc, dpd, seeds[0], fsi_par, R.fsi_seeds(1)[0], R.par(**kw),
                                       dt, l0=l0, a0=a0, nthreads=THREADS)
    assert (gp, gf) == (rp, rf)
    assert np.array_equal(bits(gs), bits(rs))
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-6 * L[0]
    assert np.abs(gv[:, 3:] - rv[:, 3:]).max() <= 1e-6 * max(np.abs(rv[:, 3:]).max(), 1.0)
