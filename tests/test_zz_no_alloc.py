"""udx_clist_build_noalloc (no device allocation, static working space)
against udx_clist_build: the same output order, cell starts and counts and
input order, bit for bit, on two calls in a row (the second smaller, with
particles outside the box, in another box shape); sizes beyond the static
space are rejected."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from ctc_amd import udx as U  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _box(L, rho, seed, outside=0.0):
    rng = np.random.default_rng(seed)
    n = int(rho * np.prod(L))
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L) * (1 + outside)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    return pp


def _with_handle(ctx, L, pp):
    c = U.Clist(ctx, L, len(pp))
    d = torch.as_tensor(pp).cuda()
    out = torch.zeros_like(d)
    m = c.build(d, out)
    res = (U_host(out)[:m], U_host(c.starts()), U_host(c.counts()), U_host(c.order(m)))
    c.close()
    return res


def U_host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _noalloc(ctx, L, pp):
    nc = int(np.prod(L))
    d = torch.as_tensor(pp).cuda()
    out = torch.full_like(d, np.nan)
    starts = torch.full((nc + 1,), -7, dtype=torch.int32, device="cuda")
    counts = torch.full((nc,), -7, dtype=torch.int32, device="cuda")
    order = torch.full((len(pp),), -7, dtype=torch.int32, device="cuda")
    U.clist_build_noalloc(ctx, L, d, out, starts, counts, order)
    s = U_host(starts)
    m = int(s[-1])
    return U_host(out)[:m], s, U_host(counts), U_host(order)[:m]


def test_noalloc_equals_clist_build_twice():
    ctx = U.Context()
    cases = [((24, 20, 16), 10, 1, 0.0), ((9, 7, 11), 3, 2, 0.2)]   # the second smaller, some outside
    for L, rho, seed, outside in cases:
        pp = _box(L, rho, seed, outside)
        say("noalloc clist", L, len(pp))
        want = _with_handle(ctx, L, pp)
        got = _noalloc(ctx, L, pp)
        for a, b, what in zip(got, want, ("pp", "starts", "counts", "order")):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32)), what
        # and the oracle's order (cell, x slab, input index)
        st, cn, srt, order = O.clist_build(L, pp)
        assert np.array_equal(got[1], st) and np.array_equal(got[3], order)
    ctx.close()


def test_noalloc_rejects_sizes_beyond_the_static_space():
    ctx = U.Context()
    d = torch.zeros((8, 6), dtype=torch.float32, device="cuda")
    i = torch.zeros(8, dtype=torch.int32, device="cuda")
    with pytest.raises(U.UdxError, match="cells exceed"):
        U.clist_build_noalloc(ctx, (65, 64, 64), d, None, i, i, i)
    big = torch.zeros((U.NOALLOC_MAXP + 1, 6), dtype=torch.float32, device="cuda")
    with pytest.raises(U.UdxError, match="particles exceed"):
        U.clist_build_noalloc(ctx, (8, 8, 8), big, None, i, i, i)
    ctx.close()
