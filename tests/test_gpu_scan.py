"""The exclusive scan through the C ABI (udx_scan_ini / apply / fin, in place
of scan_ini / scan_apply / scan_fin, src/algo/scan/imp.h:4-6): the
reference's golden vectors (the exclusive scan of `seq n`,
src/algo/scan/cpu/imp.h:4-12), and random counts at the sizes where the
implementation changes (one block up to 16,384 entries, two kernels above,
ragged ends of both) against numpy's cumulative sum; output[size] is never
written and sizes above the capacity are refused before any launch."""
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from ctc_amd import udx as U  # noqa: E402

GOLD = Path(__file__).parent / "golden" / "ref"
SENTINEL = -7


def _scan(ctx, x, cap):
    n = x.numel()
    out = torch.full((n + 1,), SENTINEL, dtype=torch.int32, device="cuda")
    sc = U.Scan(ctx, cap)
    sc.apply(x, out, n)
    torch.cuda.synchronize()
    sc.close()
    return out.cpu().numpy()


@pytest.mark.parametrize("name,n", [("t0", 16), ("t1", 64)])
def test_scan_golden_vectors(name, n):
    print("scan golden", name, n, flush=True)
    sys.stdout.flush()
    ctx = U.Context()
    got = _scan(ctx, torch.arange(1, n + 1, dtype=torch.int32, device="cuda"), n)
    ref = np.loadtxt(GOLD / ("scan.ref.scan.%s.txt" % name), dtype=np.int64)
    assert np.array_equal(got[:n], ref)
    assert got[n] == SENTINEL


@pytest.mark.parametrize("n", [1, 15, 64, 1000, 16383, 16384, 16385, 37 * 1024 + 5, 1 << 20])
def test_scan_matches_cumsum(n):
    print("scan random", n, flush=True)
    sys.stdout.flush()
    ctx = U.Context()
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, n).astype(np.int32)   # counts above 255 too
    got = _scan(ctx, torch.as_tensor(x).cuda(), n + 3)
    ref = np.concatenate([[0], np.cumsum(x.astype(np.int64))[:-1]])
    assert np.array_equal(got[:n], ref)
    assert got[n] == SENTINEL


def test_scan_refuses_size_above_capacity():
    ctx = U.Context()
    x = torch.ones(101, dtype=torch.int32, device="cuda")
    out = torch.zeros(102, dtype=torch.int32, device="cuda")
    sc = U.Scan(ctx, 100)
    with pytest.raises(U.UdxError, match="size = 101 > w->size = 100"):
        sc.apply(x, out, 101)
    sc.apply(x, out, 0)   # empty: nothing to do
    torch.cuda.synchronize()
    assert int(out.sum()) == 0
    sc.close()
