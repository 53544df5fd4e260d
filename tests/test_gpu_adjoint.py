"""GPU membrane adjoint (udx_rbc_force_adjoint) through the C ABI.

1. The kernel against the FP64 oracle adjoint on the same float inputs: gx
   within 1e-6 of its largest component (the kernel sums the same double terms
   in the same order and rounds to float once; pow and sqrt may differ by an
   ulp between the device and host libraries), and g . F_k within 1e-9
   relative (double sums in another order).
2. The membrane calibration case (256 cells of 642 vertices, the C4 set-up of
   tests/membranes.py) differentiated by torch through MembraneForce, a
   torch.autograd.Function whose forward is udx_rbc_force_apply and whose
   backward is udx_rbc_force_adjoint: it saves only the positions and the
   coefficients, and the adjoint recomputes every face's geometry.  Against
   the FP64 CPU gradient (oracle adjoint on an FP64 trajectory; that gradient
   is itself checked against central differences on the CPU): dL/dk within
   1e-4 relative per coefficient.  dL/dx0 carries the float32 rounding of the
   trajectory, which the bending term's 1/sin(theta) amplifies at nearly flat
   dihedrals with the reference's spontaneous angle (6.97 degrees): 1e-3 of
   the norm there (3.3e-4 measured, the same on the CPU with float32 states),
   and 2e-4 with angle 0.  Against the same float32 arithmetic on the CPU
   (float32 states and cotangents, production force formulas, FP64 oracle
   adjoint rounded to float): 1e-6 of the norm (1e-9 measured).
3. Central differences (h = 1e-3) of the GPU loss along a random direction of
   the initial positions (spontaneous angle 0, where the bending energy is
   smooth) against the GPU gradient: the float32 state rounds the loss, which
   bounds the agreement at 5e-3 relative.
"""
import sys

import numpy as np
import pytest

from membranes import C4, adj_setup
from oracle import rbc as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from ctc_amd import udx as U  # noqa: E402

COEF = ("ka", "kd", "kv", "ks", "kb")


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


def rbc_for(ctx, m, nc, sf, totArea):
    r = U.Rbc(ctx, m.tt, m.nv, nc)
    if sf is not None:
        r.set_stressfree(m.rr)
    else:
        r.set_stressful(totArea)
    r.set_bending(True)
    return r


@pytest.mark.parametrize("stress_free", [False, True])
def test_adjoint_kernel_matches_oracle(ctx, stress_free):
    say("adjoint kernel vs oracle, stress_free", stress_free)
    m, pp, kw, sf, l0, a0, g = adj_setup(nc=3, seed=5, stress_free=stress_free)
    g = g.astype(np.float32)
    p = R.par(**kw)
    want_gx, _, want_gp = R.adjoint(p, m, pp, 3, g.astype(np.float64), sf, l0, a0)
    r = rbc_for(ctx, m, 3, sf, kw["totArea"])
    gx = torch.zeros((len(pp), 3), dtype=torch.float32, device="cuda")
    gp = r.adjoint(U.rbc_params(**kw), 3, dev(pp), dev(g), gx)
    got = host(gx)
    assert np.abs(got - want_gx).max() <= 1e-6 * np.abs(want_gx).max()
    scale = np.abs(g[None] * R.adjoint(p, m, pp, 3, g.astype(np.float64), sf, l0, a0)[1]).sum(axis=(1, 2))
    assert np.all(np.abs(np.array(gp) - want_gp) <= 1e-9 * scale)
    # a second call on the same handle (scratch reuse) gives the same result
    gx2 = torch.zeros_like(gx)
    assert r.adjoint(U.rbc_params(**kw), 3, dev(pp), dev(g), gx2) == gp
    assert torch.equal(gx, gx2)
    r.close()


class Case:
    """the calibration case on the GPU"""

    def __init__(self, ctx, c):
        self.c, self.nc, self.nv = c, c.nc, c.m.nv
        self.rbc = rbc_for(ctx, c.m, c.nc, c.sf, c.kw["totArea"])
        self.load = torch.zeros((self.nc * self.nv, 3), dtype=torch.float32, device="cuda")
        U.rbc_stretch_apply(ctx, self.nc, self.nv, dev(c.load), self.load)
        self.target = dev(c.target, torch.float64)

    def par(self, k):
        kw = dict(self.c.kw)
        kw.update({n: float(v) for n, v in zip(COEF, k.tolist())})
        return U.rbc_params(**kw)

    def pp(self, x):
        pp = torch.zeros((len(x), 6), dtype=torch.float32, device="cuda")
        pp[:, :3] = x
        return pp


class MembraneForce(torch.autograd.Function):
    """F(x; ka, kd, kv, ks, kb): the conservative membrane force of the
    calibration case's cells"""

    @staticmethod
    def forward(ctx, x, k, case):
        ff = torch.zeros_like(x)
        case.rbc.force(case.par(k), 0.0, case.nc, case.pp(x.detach()), ff)
        ctx.save_for_backward(x, k)
        ctx.case = case
        return ff

    @staticmethod
    def backward(ctx, g):
        x, k = ctx.saved_tensors
        case = ctx.case
        gx = torch.empty_like(x)
        gp = case.rbc.adjoint(case.par(k), case.nc, case.pp(x.detach()), g.contiguous(), gx)
        return gx, torch.tensor(gp, dtype=torch.float64), None


def gpu_loss(case, x0, k):
    c = case.c
    x = x0
    for _ in range(c.STEPS):
        x = x + c.ETA * (MembraneForce.apply(x, k, case) + case.load)
    return ((x.double() - case.target) ** 2).sum()


def gpu_grad(case, x0=None):
    x = dev(case.c.x0 if x0 is None else x0).requires_grad_(True)
    k = torch.tensor(case.c.coef(), dtype=torch.float64, requires_grad=True)
    L = gpu_loss(case, x, k)
    L.backward()
    torch.cuda.synchronize()
    return float(L.detach()), x.grad.double().cpu().numpy(), k.grad.numpy()


def rel_err(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nc,phi,tol", [(256, None, 1e-3), (64, 0.0, 2e-4)])
def test_c4_gradient_matches_fp64(ctx, nc, phi, tol):
    c = C4(nc=nc, seed=1, phi=phi)
    say("calibration case: %d cells x %d vertices, %d steps, phi %s" % (c.nc, c.m.nv, c.STEPS, c.kw["phi"]))
    case = Case(ctx, c)
    L, gx, gk = gpu_grad(case)
    L64, gx64, gk64 = c.grad64()
    e64, e32 = rel_err(gx, gx64), rel_err(gx, c.grad_f32())
    say("loss gpu %.9g fp64 %.9g; dL/dx0 error: %.3g vs fp64, %.3g vs float32 path" % (L, L64, e64, e32))
    say("dL/dk gpu", gk, "fp64", gk64)
    assert abs(L - L64) <= 1e-5 * L64
    assert np.all(np.abs(gk - gk64) <= 1e-4 * np.abs(gk64)), (gk, gk64)
    assert e64 <= tol
    assert e32 <= 1e-6
    case.rbc.close()


def test_c4_gradient_central_difference(ctx):
    c = C4(nc=8, seed=3, phi=0.0)
    say("calibration case, central differences: %d cells" % c.nc)
    case = Case(ctx, c)
    _, gx, _ = gpu_grad(case)
    h = 1e-3
    with torch.no_grad():
        k = torch.tensor(c.coef(), dtype=torch.float64)
        lp = float(gpu_loss(case, dev(c.x0 + h * c.u), k))
        lm = float(gpu_loss(case, dev(c.x0 - h * c.u), k))
    fd = (lp - lm) / (2 * h)
    an = float((gx * c.u).sum())
    say("directional derivative: fd %.9g adjoint %.9g" % (fd, an))
    assert abs(fd - an) <= 5e-3 * abs(an)
    case.rbc.close()


def test_adjoint_rejects_bad_calls(ctx):
    m, pp, kw, sf, l0, a0, g = adj_setup(nc=2, seed=1)
    r = U.Rbc(ctx, m.tt, m.nv, 2)
    gx = torch.zeros((len(pp), 3), dtype=torch.float32, device="cuda")
    with pytest.raises(U.UdxError, match="set_stress"):
        r.adjoint(U.rbc_params(**kw), 2, dev(pp), dev(g), gx)
    r.set_stressful(kw["totArea"])
    with pytest.raises(U.UdxError, match="max_cells"):
        r.adjoint(U.rbc_params(**kw), 3, dev(pp), dev(g), gx)
    r.close()
