"""GPU parity of the SDF walls against the CPU wall oracle, through the C ABI.

Bars: bit-exact for the texture, sdf_eval / cheap / gradient probes, the
freeze split, the wall particles and their margin-box cell list, bounce-back,
exact-mode wall forces and wall-bounded trajectories; the production wall
force within the force tolerance of the solvent kernel; the repulsion within
the ulp error of expf.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import wall as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from ctc_amd import udx as U  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")
CYL = os.path.join(GOLD, "sdf", "cyl.dat")
FTOL = 1e-5


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


def texture(q):
    T, _ = q.size()
    out = torch.empty(int(np.prod(T)), dtype=torch.float32, device="cuda")
    U._copy_dev(out, U.lib().udx_sdf_texture(q.h), out.numel() * 4)
    return host(out)


def channel_field(L, h, N=48):
    return np.array(L, np.float32), (N, N, N), W.plane_channel_sdf(L, N, h)


def gpu_sdf(ctx, L, src):
    if src == "cyl":
        return U.Sdf(ctx, L, path=CYL)
    ext, N, D = src
    return U.Sdf(ctx, L, ext=ext, N=N, D=D)


def cpu_sdf(L, src):
    if src == "cyl":
        ext, N, D = W.read_sdf_dat(CYL)
        return W.Sdf(L, ext, N, D)
    ext, N, D = src
    return W.Sdf(L, ext, N, D.ravel())


SDF_CASES = [((16, 16, 16), "cyl"), ((24, 20, 32), "chan"), ((32, 16, 8), "cyl")]


def case_src(L, kind):
    return "cyl" if kind == "cyl" else channel_field(L, 0.3 * L[2])


@pytest.mark.parametrize("L,kind", SDF_CASES)
def test_texture_bit_exact(ctx, L, kind):
    say("texture", L, kind)
    src = case_src(L, kind)
    q, ref = gpu_sdf(ctx, L, src), cpu_sdf(L, src)
    T, thr = q.size()
    assert T == ref.T
    assert np.float32(thr) == np.float32(ref.cheap_threshold)
    got = texture(q)
    assert np.array_equal(bits(got), bits(ref.tex))
    q.close()


def probe_points(L, n, seed):
    rng = np.random.default_rng(seed)
    lw = np.array(L, np.float64) / 2 + W.MARGIN
    return (rng.uniform(-1, 1, (n, 3)) * lw * 0.999).astype(np.float32)


@pytest.mark.parametrize("L,kind", SDF_CASES)
def test_probe_bit_exact(ctx, L, kind):
    say("probe", L, kind)
    src = case_src(L, kind)
    q, ref = gpu_sdf(ctx, L, src), cpu_sdf(L, src)
    r = probe_points(L, 2000, 7)
    rd = dev(r)
    assert np.array_equal(bits(host(q.probe(rd, 0))), bits(ref.eval(r)))
    assert np.array_equal(bits(host(q.probe(rd, 1))), bits(ref.cheap(r)))
    assert np.array_equal(bits(host(q.probe(rd, 2))), bits(ref.grad(r)))
    q.close()


def channel_pp(L, rho, seed, vel=0.3):
    rng = np.random.default_rng(seed)
    n = int(rho * np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, vel, (n, 3)).astype(np.float32)
    return pp


class Setup:
    """One channel: the GPU walls from udx_wall_gen and the oracle's freeze,
    images and margin-box cell list of the same particles."""

    def __init__(self, ctx, L, h, rho, seed, exact=True):
        self.L = L
        src = channel_field(L, h)
        self.q, self.ref = gpu_sdf(ctx, L, src), cpu_sdf(L, src)
        pp = channel_pp(L, rho, seed)
        ppd = dev(pp)
        self.w = U.Wall(ctx, self.q, ppd, exact=exact)
        self.solvent = host(ppd)[:self.w.n_solvent].copy()
        self.keep, frozen = self.ref.freeze(pp)
        allw = W.wall_images(L, frozen)
        self.nfrozen = len(frozen)
        self.wstarts, _, self.wpp, _ = O.clist_build(W.margin_box(L), allw, sub=1)

    def close(self):
        self.w.close()
        self.q.close()


WALL_CASES = [((16, 12, 24), 8.0, 4, 1), ((20, 16, 16), 5.5, 10, 2)]


@pytest.mark.parametrize("L,h,rho,seed", WALL_CASES)
def test_wall_gen_bit_exact(ctx, L, h, rho, seed):
    say("wall_gen", L, h, rho)
    s = Setup(ctx, L, h, rho, seed)
    assert np.array_equal(bits(s.solvent), bits(s.keep))
    nf, nt = s.w.counts()
    assert (nf, nt) == (s.nfrozen, len(s.wpp))
    assert np.array_equal(bits(host(s.w.particles())), bits(s.wpp))
    assert np.array_equal(host(s.w.starts()), s.wstarts)
    s.close()


@pytest.mark.parametrize("L,h,rho,seed", WALL_CASES)
@pytest.mark.parametrize("exact", [True, False])
def test_wall_forces(ctx, L, h, rho, seed, exact):
    """Wall force added to a nonzero force array: exact mode bit for bit,
    production mode within FTOL * max(|f|, 0.1 F_rms); pair counts equal."""
    say("wall force", L, h, rho, "exact", exact)
    s = Setup(ctx, L, h, rho, seed, exact=exact)
    starts, counts, srt, _ = O.clist_build(L, s.solvent)
    par = (25.0, 4.5, O.dpd_sigma(0.0444, 4.5, 5e-4), 0.25)
    seed_f = np.float32(0.6123)
    ff0 = np.random.default_rng(3).normal(0, 1, (len(srt), 3)).astype(np.float32)
    want, npair = W.wall_forces(L, srt, s.ref, s.wpp, s.wstarts, par, seed_f, ff0)
    ffd = dev(ff0)
    s.w.pairs(reset=True)
    s.w.force(U.PairDpd(*par), dev(srt), ffd, seed=seed_f)
    got = host(ffd)
    assert s.w.pairs() == npair and npair > 0
    if exact:
        assert np.array_equal(bits(got), bits(want))
    else:
        df = want - ff0
        frms = np.sqrt((df ** 2).sum(1).mean())
        # the kernel's error on the wall force plus one rounding of the sum
        tol = FTOL * np.maximum(np.abs(df), 0.1 * frms) + 2.4e-7 * np.abs(want)
        assert np.all(np.abs(got - want) <= tol)
    s.close()


def test_bounce_bit_exact(ctx):
    L = (16, 12, 24)
    src = channel_field(L, 8.0)
    q, ref = gpu_sdf(ctx, L, src), cpu_sdf(L, src)
    rng = np.random.default_rng(4)
    n = 4000
    pp = np.zeros((n, 6), np.float32)
    pp[:, 0] = rng.uniform(-8, 8, n)
    pp[:, 1] = rng.uniform(-6, 6, n)
    side = rng.choice([-1, 1], n)
    pp[:, 2] = side * rng.uniform(6.0, 8.3, n)        # fluid and just inside the walls
    pp[:, 3:5] = rng.normal(0, 1, (n, 2))
    pp[:, 5] = side * rng.uniform(-2.0, 2.0, n)
    dt = np.float32(0.05)
    ppd = dev(pp)
    q.bounce(float(dt), ppd)
    got = host(ppd)
    want = ref.bounce(pp, dt)
    assert np.count_nonzero(np.any(want != pp, axis=1)) > 100
    assert np.array_equal(bits(got), bits(want))
    q.close()


def test_repulse_tolerance(ctx):
    L = (16, 12, 24)
    src = channel_field(L, 8.0)
    q, ref = gpu_sdf(ctx, L, src), cpu_sdf(L, src)
    rng = np.random.default_rng(5)
    n = 4000
    pp = np.zeros((n, 6), np.float32)
    pp[:, :2] = rng.uniform(-6, 6, (n, 2))
    pp[:, 2] = rng.choice([-1, 1], n) * rng.uniform(4.0, 9.0, n)
    ffd = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    q.repulse(2.0, dev(pp), ffd)
    got = host(ffd)
    want = ref.repulse(pp, 2.0)
    assert np.count_nonzero(want) > 100
    # expf: a few ulp between the device and host libraries
    assert np.allclose(got, want, rtol=4e-7 * 8, atol=1e-6)
    q.close()


@pytest.mark.parametrize("L,h,rho,accel", [((16, 12, 24), 8.0, 4, (0.5, 0.0, 0.0)),
                                           ((12, 10, 16), 5.0, 10, (0.2, 0.1, 0.0))])
@pytest.mark.parametrize("graph", [False, True])
def test_wall_solver_trajectory_bit_exact(ctx, L, h, rho, accel, graph):
    """clist + solvent DPD + wall DPD + body force + kick-drift + bounce-back
    + wrap over 20 steps in exact mode: equal to the oracle bit for bit, also
    when the steps are replayed from a captured graph."""
    say("wall trajectory", L, h, rho, "graph", graph)
    nsteps = 20
    dt = 1e-3
    s = Setup(ctx, L, h, rho, 9)
    params = (25.0, 4.5, O.dpd_sigma(1.0, 4.5, dt), 0.25)
    seeds = O.seeds_table(nsteps)
    wtrunk = O.kiss_fill(42, 42, 42, 42, nsteps)
    seeds28 = np.concatenate([seeds, np.asarray(wtrunk, np.float32)[:, None]], axis=1)
    ref = s.keep.copy()
    pairs_ref = wpairs_ref = 0
    for k in range(nsteps):
        ref, npair, nw = W.flu_wall_step(L, ref, params, seeds28[k], dt, s.ref, s.wpp, s.wstarts, accel=accel,
                                         nthreads=8)
        pairs_ref += npair
        wpairs_ref += nw
    flu = U.Flu(ctx, L, len(s.solvent), U.PairDpd(*params), dt, 1.0, accel, exact=True)
    flu.graph(graph)
    ppd = dev(s.solvent)
    flu.set_pp(ppd)
    flu.set_wall(s.w)
    flu.run(nsteps, seeds28)
    out = torch.zeros_like(ppd)
    n = flu.get_pp(out)
    got = host(out)[:n]
    assert flu.pairs() == pairs_ref
    assert flu.wall_pairs() == wpairs_ref and wpairs_ref > 0
    assert np.array_equal(bits(got), bits(ref))
    flu.close()
    s.close()


# wall velocities (src/wall/wvel): the library's udx_wall_vel and the oracle's
# resolved view at time t; "shear sin" is the shear view with rate gdot sin(w t)
WVEL_CASES = {
    "constant": dict(kind="constant", u=(0.3, -0.2, 0.1)),
    "shear": dict(kind="shear", gdot=0.5, vdir=0, gdir=2),
    "shear half": dict(kind="shear", gdot=-0.4, vdir=1, gdir=2, half=1),
    "shear sin": dict(kind="shear sin", gdot=0.5, vdir=0, gdir=2, w=3.0),
    "hele shaw": dict(kind="hele shaw", hs_u=0.05, hs_h=30.0),
}


def wvel_pair(name, t):
    c = dict(WVEL_CASES[name])
    v = U.wall_vel(**c)
    kind = c.pop("kind")
    if kind == "shear sin":
        ref = W.wvel_shear_sin(c["gdot"], c["w"], t, c["vdir"], c["gdir"], c.get("half", 0))
    elif kind == "hele shaw":
        ref = W.wvel(kind, hu=c["hs_u"], hh=c["hs_h"])
    else:
        ref = W.wvel(kind, **c)
    return v, ref


@pytest.mark.parametrize("name", list(WVEL_CASES))
def test_wall_vel_forces_bit_exact(ctx, name):
    """Exact-mode wall force with moving wall particles, bit for bit."""
    L, h, rho = (16, 12, 24), 8.0, 4
    say("wall force, wall velocity", name)
    s = Setup(ctx, L, h, rho, 1, exact=True)
    starts, counts, srt, _ = O.clist_build(L, s.solvent)
    par = (25.0, 4.5, O.dpd_sigma(0.0444, 4.5, 5e-4), 0.25)
    seed_f = np.float32(0.3141)
    t = np.float32(0.4)
    v, wv = wvel_pair(name, t)
    s.w.set_vel(v)
    s.w.set_time(t)
    want, npair = W.wall_forces(L, srt, s.ref, s.wpp, s.wstarts, par, seed_f, None, wv)
    rest, _ = W.wall_forces(L, srt, s.ref, s.wpp, s.wstarts, par, seed_f)
    assert not np.array_equal(want, rest)
    ffd = torch.zeros((len(srt), 3), dtype=torch.float32, device="cuda")
    s.w.pairs(reset=True)
    s.w.force(U.PairDpd(*par), dev(srt), ffd, seed=seed_f)
    assert s.w.pairs() == npair and npair > 0
    assert np.array_equal(bits(host(ffd)), bits(want))
    s.close()


@pytest.mark.parametrize("name", list(WVEL_CASES))
def test_wall_vel_bounce_bit_exact(ctx, name):
    L = (16, 12, 24)
    say("bounce, wall velocity", name)
    s = Setup(ctx, L, 8.0, 2, 3)
    rng = np.random.default_rng(8)
    n = 4000
    pp = np.zeros((n, 6), np.float32)
    pp[:, 0] = rng.uniform(-8, 8, n)
    pp[:, 1] = rng.uniform(-6, 6, n)
    side = rng.choice([-1, 1], n)
    pp[:, 2] = side * rng.uniform(6.0, 8.3, n)
    pp[:, 3:5] = rng.normal(0, 1, (n, 2))
    pp[:, 5] = side * rng.uniform(-2.0, 2.0, n)
    dt = np.float32(0.05)
    t = np.float32(1.3)
    v, wv = wvel_pair(name, t)
    s.w.set_vel(v)
    s.w.set_time(t)
    ppd = dev(pp)
    s.w.bounce(float(dt), ppd)
    want = s.ref.bounce(pp, dt, wv)
    assert not np.array_equal(want, s.ref.bounce(pp, dt))
    assert np.array_equal(bits(host(ppd)), bits(want))
    s.close()


@pytest.mark.parametrize("name", ["constant", "shear", "shear half", "shear sin"])
def test_wall_vel_trajectory_bit_exact(ctx, name):
    """Moving walls in the solvent step (wall DPD and bounce-back with the
    wall velocity at the step's time; the clock starts at t0 and advances by
    dt in float) over 12 steps in exact mode: equal to the oracle bit for bit."""
    L, h, rho, accel = (16, 12, 24), 8.0, 4, (0.0, 0.0, 0.0)
    say("wall trajectory, wall velocity", name)
    nsteps, dt, t0 = 12, 1e-3, np.float32(0.25)
    s = Setup(ctx, L, h, rho, 9)
    params = (25.0, 4.5, O.dpd_sigma(1.0, 4.5, dt), 0.25)
    seeds = O.seeds_table(nsteps)
    wtrunk = O.kiss_fill(42, 42, 42, 42, nsteps)
    seeds28 = np.concatenate([seeds, np.asarray(wtrunk, np.float32)[:, None]], axis=1)
    ref, t = s.keep.copy(), t0
    pairs_ref = wpairs_ref = 0
    for k in range(nsteps):
        v, wv = wvel_pair(name, t)
        ref, npair, nw = W.flu_wall_step(L, ref, params, seeds28[k], dt, s.ref, s.wpp, s.wstarts, accel=accel,
                                         nthreads=8, wv=wv)
        pairs_ref += npair
        wpairs_ref += nw
        t = np.float32(t + np.float32(dt))
    flu = U.Flu(ctx, L, len(s.solvent), U.PairDpd(*params), dt, 1.0, accel, exact=True)
    ppd = dev(s.solvent)
    flu.set_pp(ppd)
    s.w.set_vel(v)
    flu.set_wall(s.w)
    flu.set_time(t0)
    flu.run(nsteps, seeds28)
    out = torch.zeros_like(ppd)
    n = flu.get_pp(out)
    got = host(out)[:n]
    assert flu.pairs() == pairs_ref
    assert flu.wall_pairs() == wpairs_ref and wpairs_ref > 0
    assert np.array_equal(bits(got), bits(ref))
    flu.close()
    s.close()
