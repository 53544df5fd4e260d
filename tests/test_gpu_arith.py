"""The device arithmetic the force kernels rely on for bit equality with the
CPU restatement, checked over every float of its range on the GPU: the
reciprocal (hardware estimate + one Newton step) of the pair distance, the
correctly rounded square root (distance and w_R roots) and the fractional
part of the branch generator's low-discrepancy sum."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from ctc_amd import udx as U  # noqa: E402

TINY = float(np.ldexp(np.float32(1), -24))


@pytest.mark.parametrize("what,lo,hi", [
    (U.CHECK_RCP, TINY, 1.0),       # r = sqrt(r2) >= 1e-3 in the pair force
    (U.CHECK_SQRT, TINY, 1.0),      # r2 in [1e-6, 1), 1 - r >= 2^-24, its root
    (U.CHECK_FRACT, 0.0, 65536.0),  # the sum stays below 2^15 for indices below 2^27
])
def test_device_arithmetic_is_correctly_rounded(what, lo, hi):
    ctx = U.Context()
    print("selfcheck", what, lo, hi, flush=True)
    sys.stdout.flush()
    assert U.selfcheck(ctx, what, lo, hi) == 0
    ctx.close()


def test_selfcheck_counts_differences():
    """the hardware reciprocal estimate alone is not correctly rounded: the
    check must report differences for it, and reject bad arguments"""
    ctx = U.Context()
    bad = U.selfcheck(ctx, U.CHECK_RCP_ESTIMATE, 1.0, 2.0)
    print("rcp estimate: %d of %d floats in [1, 2) differ" % (bad, 2 ** 23), flush=True)
    assert bad > 0
    with pytest.raises(U.UdxError):
        U.selfcheck(ctx, U.CHECK_RCP, 2.0, 1.0)
    with pytest.raises(U.UdxError):
        U.selfcheck(ctx, 7, 0.5, 1.0)
    ctx.close()
