"""Restart through the solvent and the walls (src/flu/imp/start.h,
src/wall/imp/strt.h), through the C ABI.

- The solvent dumped after 3 exact steps and read back into a new solvent
  equals the dumped state bit for bit, and 5 more steps from it equal 5 more
  steps of the original run.
- The wall template holds the oracle's frozen particles in freeze order, and
  a wall built from it has the same particles and cell list as the frozen
  wall, and gives the same wall forces.
"""
import sys

import numpy as np
import pytest

import oracle as O
from oracle import wall as W

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


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_solvent_restart(ctx, tmp_path):
    L, n = (12, 12, 12), 3 * 12 ** 3
    rng = np.random.default_rng(3)
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-6, 6, (n, 3)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    dt = 1e-3
    par = U.PairDpd(75.0, 4.5, O.dpd_sigma(1.0, 4.5, dt), 0.25)
    seeds = O.seeds_table(8)
    (tmp_path / "flu.pp").mkdir()
    say("solvent restart", n)
    a = U.Flu(ctx, L, n, par, dt, 1.0, exact=True)
    a.set_pp(dev(pp))
    a.run(3, seeds[:3])
    a.strt_dump(tmp_path, 3)
    mid = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    a.get_pp(mid)
    b = U.Flu(ctx, L, n, par, dt, 1.0, exact=True)
    b.strt_quants(tmp_path, 3)
    back = torch.zeros_like(mid)
    assert b.get_pp(back) == n
    assert np.array_equal(bits(host(back)), bits(host(mid)))
    assert np.array_equal(U.restart_read_pp(tmp_path, "flu.pp", 3, n).view(np.uint32), bits(host(mid)))
    a.run(5, seeds[3:])
    b.run(5, seeds[3:])
    oa, ob = torch.zeros_like(mid), torch.zeros_like(mid)
    a.get_pp(oa)
    b.get_pp(ob)
    assert np.array_equal(bits(host(oa)), bits(host(ob)))
    a.close()
    b.close()


def test_wall_template_restart(ctx, tmp_path):
    L, h, rho = (16, 12, 24), 8.0, 4
    src = bench.plane_channel(L, 48, h)
    ext = np.array(L, np.float32)
    q = U.Sdf(ctx, L, ext=ext, N=(48, 48, 48), D=src)
    ref = W.Sdf(L, ext, (48, 48, 48), src.ravel())
    rng = np.random.default_rng(5)
    n = rho * L[0] * L[1] * L[2]
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    say("wall template restart", n)
    w = U.Wall(ctx, q, dev(pp), exact=True)
    (tmp_path / "wall.pp").mkdir()
    w.strt_dump_templ(tmp_path)
    _, frozen = ref.freeze(pp)
    templ = U.restart_read_pp(tmp_path, "wall.pp", U.RESTART_TEMPL, n)
    assert np.array_equal(bits(templ), bits(frozen)) and len(frozen) > 0
    w2 = U.Wall(ctx, q, None, exact=True, strt=tmp_path, maxn=n)
    assert w2.counts() == w.counts()
    assert np.array_equal(bits(host(w2.particles())), bits(host(w.particles())))
    assert np.array_equal(host(w2.starts()), host(w.starts()))
    keep = ref.freeze(pp)[0]
    starts, counts, srt, _ = O.clist_build(L, keep)
    par = U.PairDpd(25.0, 4.5, O.dpd_sigma(0.0444, 4.5, 5e-4), 0.25)
    f1 = torch.zeros((len(srt), 3), dtype=torch.float32, device="cuda")
    f2 = torch.zeros_like(f1)
    w.force(par, dev(srt), f1, seed=0.25)
    w2.force(par, dev(srt), f2, seed=0.25)
    assert torch.equal(f1, f2) and bool(f1.abs().sum() > 0)
    with pytest.raises(U.UdxError, match="capacity"):
        U.Wall(ctx, q, None, strt=tmp_path, maxn=len(frozen) - 1)
    w.close()
    w2.close()
    q.close()
