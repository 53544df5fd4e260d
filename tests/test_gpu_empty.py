"""Empty and fully dropped inputs through the C ABI: a cell list of no
particles, or of particles that all lie outside the box (dropped, as
src/clist/dev/main.h:6-11 `inside` drops them), has all-zero starts and
counts equal to the oracle's, and the force pass over no particles returns
without error and writes nothing."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import oracle as O  # noqa: E402
from ctc_amd import udx as U  # noqa: E402

L = (4, 5, 6)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", ["empty", "all_outside"])
def test_clist_without_particles(kind):
    _say("clist", kind)
    ctx = U.Context()
    if kind == "empty":
        pp = np.zeros((0, 6), np.float32)
    else:
        rng = np.random.default_rng(5)
        pp = rng.uniform(-1, 1, (50, 6)).astype(np.float32)
        pp[:, 0] = np.float32(L[0] / 2) + np.abs(pp[:, 0])   # x >= L/2: outside
    s_ref, c_ref, out_ref, _ = O.clist_build(L, pp)
    assert len(out_ref) == 0
    cl = U.Clist(ctx, L, max(len(pp), 1))
    ppd = torch.as_tensor(pp).cuda()
    out = torch.full((max(len(pp), 1), 6), -1.0, dtype=torch.float32, device="cuda")
    nout = cl.build(ppd, out)
    assert nout == 0
    assert np.array_equal(cl.starts().cpu().numpy(), s_ref)
    assert np.array_equal(cl.counts().cpu().numpy(), c_ref)
    assert not cl.starts().any()
    assert bool((out == -1.0).all())          # nothing gathered
    cl.close()


def test_forces_without_particles():
    _say("fluforces empty")
    ctx = U.Context()
    cl = U.Clist(ctx, L, 1)
    empty = torch.zeros((0, 6), dtype=torch.float32, device="cuda")
    assert cl.build(empty, None) == 0
    f = U.FluForces(ctx, L, 1)
    f.prepare(empty)
    ff = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    f.apply(U.dpd_params(25.0, 4.5, 1.0, 1e-3), cl.starts_ptr(), ff, seeds=[0.5] * 27)
    f.apply(U.dpd_params(25.0, 4.5, 1.0, 1e-3), cl.starts_ptr(), ff)
    torch.cuda.synchronize()
    assert f.pairs() == 0
    f.close()
    cl.close()
