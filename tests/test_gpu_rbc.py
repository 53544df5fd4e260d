"""GPU parity of the RBC membrane path against the CPU oracle, through the C
ABI: ring adjacency identical, FP64 area/volume within 1e-12 (the block
reduction sums in another order), forces within the FP32 rounding of the
double sum plus the ulp differences of powf between the device and host
libraries."""
import os
import sys

import numpy as np
import pytest

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


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def mesh(nv):
    return R.Mesh(*R.read_off(os.path.join(GOLD, "cells", "%d.off" % nv)))


def rotations(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        A = np.eye(4)
        A[:3, :3] = q * np.sign(np.linalg.det(q))
        A[:3, 3] = rng.uniform(-6, 6, 3) + 16
        out.append(A)
    return out


def test_adjacency_identical(ctx):
    for nv in (498, 642, 1986):
        m = mesh(nv)
        g = U.Rbc(ctx, m.tt, m.nv, 4)
        assert g.md == m.md
        a0, a1 = g.adjacency()
        assert np.array_equal(a0, m.adj0) and np.array_equal(a1, m.adj1)
        g.close()


def test_area_volume_fp64(ctx):
    m = mesh(498)
    pp = R.place(m.rr, rotations(7, 1), (32, 32, 32))
    g = U.Rbc(ctx, m.tt, m.nv, 16)
    got = host(g.area_volume(7, dev(pp)))
    want = R.area_volume(m, pp, 7)
    assert np.allclose(got, want, rtol=1e-12, atol=0)
    g.close()


def check_forces(got, want, fmax):
    err = np.abs(got.astype(np.float64) - want)
    # float rounding of the double sum + powf ulps inside the spring terms
    tol = 2 ** -23 * np.abs(want) + 2e-6 * fmax
    assert np.all(err <= tol), "max err %g (fmax %g)" % (err.max(), fmax)


@pytest.mark.parametrize("stress_free", [False, True])
def test_forces_golden_setup(ctx, stress_free):
    """The reference's forces.ref.rbc.u.{simple,stree.free} setup: two
    642-vertex cells, conf/rbc.test.cfg, zero velocities."""
    say("rbc golden setup, stress free", stress_free)
    m = mesh(642)
    T = np.eye(4)
    T[:3, 3] = [1, 2, 3]
    pp = R.place(m.rr, [np.eye(4), T], (16, 16, 16))
    A0, V0 = 80.0594, 43.3586
    p = R.par(**R.RBC_TEST, totArea=A0, totVolume=V0)
    av = R.area_volume(m, pp, 2)
    g = U.Rbc(ctx, m.tt, m.nv, 4)
    if stress_free:
        want = R.forces(p, m, pp, 2, av, stress_free=m.shape())
        g.set_stressfree(m.rr)
    else:
        l0, a0 = m.stressful(A0)
        want = R.forces(p, m, pp, 2, av, l0=l0, a0=a0)
        g.set_stressful(A0)
    ff = torch.zeros((2 * m.nv, 3), dtype=torch.float32, device="cuda")
    g.force(U.rbc_params(**R.RBC_TEST, totArea=A0, totVolume=V0), 5e-4, 2, dev(pp), ff)
    got = host(ff)
    check_forces(got, want, np.abs(want).max())
    g.close()


@pytest.mark.parametrize("nv,nc,bending", [(498, 9, True), (642, 3, False), (1986, 2, True)])
def test_forces_random_cells(ctx, nv, nc, bending):
    """Noisy rotated cells with random velocities (viscous terms on), the
    stress-full model away from the target area and volume, forces added to
    a nonzero array."""
    say("rbc random cells", nv, nc, bending)
    m = mesh(nv)
    rng = np.random.default_rng(nv + nc)
    pp = R.place(m.rr, rotations(nc, nv), (32, 32, 32))
    pp[:, :3] += rng.normal(0, 0.02, (len(pp), 3)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 0.5, (len(pp), 3)).astype(np.float32)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, totArea=np.float32(A0 * 1.02), totVolume=np.float32(V0 * 0.97))
    if not bending:
        kw["kb"] = 0.0
    p = R.par(**kw)
    av = R.area_volume(m, pp, nc)
    l0, a0 = m.stressful(kw["totArea"])
    ff0 = rng.normal(0, 1, (nc * m.nv, 3)).astype(np.float32)
    want = R.forces(p, m, pp, nc, av, l0=l0, a0=a0, bending=bending)
    g = U.Rbc(ctx, m.tt, m.nv, nc)
    g.set_stressful(kw["totArea"])
    g.set_bending(bending)
    ffd = dev(ff0)
    g.force(U.rbc_params(**kw), 5e-4, nc, dev(pp), ffd)
    got = host(ffd) - ff0
    # the subtraction of ff0 adds one rounding of |ff0 + f|
    fmax = np.abs(want).max()
    err = np.abs(got.astype(np.float64) - want)
    tol = 2 ** -23 * (np.abs(want) + np.abs(ff0) + 1) + 2e-6 * fmax
    assert np.all(err <= tol), "max err %g (fmax %g)" % (err.max(), fmax)
    g.close()


def test_stretch(ctx):
    nv, nc = 498, 3
    rng = np.random.default_rng(2)
    f = rng.normal(0, 1, (nv, 3)).astype(np.float32)
    ff0 = rng.normal(0, 1, (nc * nv, 3)).astype(np.float32)
    ffd = dev(ff0)
    U.rbc_stretch_apply(ctx, nc, nv, dev(f), ffd)
    assert np.array_equal(host(ffd), R.stretch(nc, nv, f, ff0))


# ------------------------------------------------------------------ FSI
def fsi_case(L, rho, seed, out):
    import oracle as O
    rng = np.random.default_rng(seed)
    n = int(rho * np.prod(L))
    spp = np.zeros((n, 6), np.float32)
    spp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp[:, 3:] = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    starts, counts, srt, _ = O.clist_build(L, spp)
    m = mesh(498)
    opp = R.place(m.rr, rotations(3, seed), (0, 0, 0))
    # three cells: one near the upper corner (outside on every axis), one
    # near the lower corner, one inside
    c = [np.array(L, np.float32) / 2 - np.array(out, np.float32), -np.array(L, np.float32) / 2 + 0.7,
         np.zeros(3, np.float32)]
    for k in range(3):
        sl = slice(k * m.nv, (k + 1) * m.nv)
        opp[sl, :3] = (opp[sl, :3] - opp[sl, :3].mean(0) + c[k]).astype(np.float32)
    opp[:, 3:] = rng.normal(0, 0.3, (len(opp), 3)).astype(np.float32)
    return srt, starts, opp


@pytest.mark.parametrize("exact", [True, False])
def test_fsi_forces(ctx, exact):
    """Vertex and solvent forces added to nonzero arrays: exact mode bit for
    bit with the oracle (same order on both sides), production mode within
    the force tolerance; pair counts equal."""
    import oracle as O
    L = (16, 12, 14)
    srt, starts, opp = fsi_case(L, 5, 2, (1.0, 1.2, 0.4))
    par = (2.6666667, 8.0, O.dpd_sigma(0.0444, 8.0, 5e-4), 0.25)
    seeds = np.array([0.31, 0.77], np.float32)
    rng = np.random.default_rng(9)
    fo0 = rng.normal(0, 1, (len(opp), 3)).astype(np.float32)
    fs0 = rng.normal(0, 1, (len(srt), 3)).astype(np.float32)
    fo_ref, fs_ref, n_ref = R.fsi_forces(L, srt, starts, opp, par, seeds, fo0, fs0)
    say("fsi pairs", n_ref, "exact", exact)
    f = U.Fsi(ctx, L, len(opp), exact=exact)
    spp_d, fs_d, st_d = dev(srt), dev(fs0), dev(starts, torch.int32)
    f.bind_solvent(spp_d, fs_d, U.C.c_void_p(st_d.data_ptr()))
    fo_d = dev(fo0)
    f.apply(U.PairDpd(*par), dev(opp), fo_d, seeds=seeds)
    fo, fs = host(fo_d), host(fs_d)
    assert f.pairs() == n_ref and n_ref > 1000
    if exact:
        assert np.array_equal(fo.view(np.uint32), fo_ref.view(np.uint32))
        assert np.array_equal(fs.view(np.uint32), fs_ref.view(np.uint32))
    else:
        for got, want, base in ((fo, fo_ref, fo0), (fs, fs_ref, fs0)):
            df = want.astype(np.float64) - base
            frms = np.sqrt((df ** 2).sum(1).mean())
            tol = 1e-5 * np.maximum(np.abs(df), 0.1 * frms) + 2.4e-7 * np.abs(want)
            assert np.all(np.abs(got - want) <= tol)
    f.close()


def test_fsi_trunk(ctx):
    """Two apply calls draw (bulk, halo) seeds from KISS(1908, 1409, 290, 12968)."""
    import oracle as O
    L = (12, 12, 12)
    srt, starts, opp = fsi_case(L, 3, 4, (1.0, 1.0, 1.0))
    par = (2.6666667, 8.0, 1.0, 0.25)
    seeds = R.fsi_seeds(2)
    fo_ref = np.zeros((len(opp), 3), np.float32)
    fs_ref = np.zeros((len(srt), 3), np.float32)
    for k in range(2):
        fo_ref, fs_ref, _ = R.fsi_forces(L, srt, starts, opp, par, seeds[k], fo_ref, fs_ref)
    f = U.Fsi(ctx, L, len(opp), exact=True)
    spp_d, fs_d, st_d = dev(srt), torch.zeros((len(srt), 3), device="cuda"), dev(starts, torch.int32)
    f.bind_solvent(spp_d, fs_d, U.C.c_void_p(st_d.data_ptr()))
    fo_d = torch.zeros((len(opp), 3), device="cuda")
    opp_d = dev(opp)
    f.apply(U.PairDpd(*par), opp_d, fo_d)
    f.apply(U.PairDpd(*par), opp_d, fo_d)
    assert np.array_equal(host(fo_d), fo_ref) and np.array_equal(host(fs_d), fs_ref)
    f.close()


@pytest.mark.parametrize("nv,nc", [(498, 3), (642, 2)])
def test_juelicher_bending(ctx, nv, nc):
    """bending "juelicher": the membrane force without bending plus the
    oracle's Juelicher terms (f + fad), accumulated as the kernel does.  The
    device's double atan2 may differ from the host's by an ulp before the
    float rounding; the bound is 2 float ulps of the largest term."""
    say("juelicher", nv, nc)
    m = mesh(nv)
    rng = np.random.default_rng(nv + 7)
    pp = R.place(m.rr, rotations(nc, nv + 1), (32, 32, 32))
    pp[:, :3] += rng.normal(0, 0.02, (len(pp), 3)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 0.5, (len(pp), 3)).astype(np.float32)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, totArea=np.float32(A0 * 1.02), totVolume=np.float32(V0 * 0.97))
    g = U.Rbc(ctx, m.tt, m.nv, nc)
    g.set_stressful(kw["totArea"])
    g.set_bending_type("none")
    base = torch.zeros((nc * m.nv, 3), dtype=torch.float32, device="cuda")
    g.force(U.rbc_params(**kw), 5e-4, nc, dev(pp), base)
    g.set_bending_type("juelicher")
    got = torch.zeros_like(base)
    g.force(U.rbc_params(**kw), 5e-4, nc, dev(pp), got)
    f, fad = R.jul_forces(m, pp, nc)
    want = host(base) + (f + fad)
    got = host(got)
    scale = np.abs(f + fad).max()
    exact = np.mean(got == want)
    say("juelicher: %.4f of the components bit-identical" % exact)
    assert np.abs(got - want).max() <= 2 * 2 ** -23 * (scale + np.abs(want).max())
    # the adjoint refuses the Juelicher model
    with pytest.raises(U.UdxError, match="Kantor"):
        g.adjoint(U.rbc_params(**kw), nc, dev(pp), dev(np.zeros((nc * m.nv, 3))), torch.zeros_like(base))
    g.set_bending_type("kantor")
    g.close()
