"""The C-ABI library loads and exports every entry point include/udx.h
declares (no GPU needed: nothing here launches a kernel)."""
import os
import re

import numpy as np

import oracle as O
from ctc_amd import udx as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "udx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(udx_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = U.lib()
    names = _declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(U.EXPORTED)


def test_host_kiss_matches_oracle():
    r = U.RndTrunk(42, 42, 42, 42)
    got = np.array([r.get() for _ in range(1000)], np.float32)
    assert np.array_equal(got.view(np.uint32), O.kiss_fill(42, 42, 42, 42, 1000).view(np.uint32))


def test_sigma_matches_reference_formula():
    p = U.dpd_params(2.5, 8.0, 0.0444, 5e-4)
    assert np.float32(p.s) == np.float32(O.dpd_sigma(0.0444, 8.0, 5e-4))
    assert (p.a, p.g, p.spow) == (2.5, 8.0, 0.25)
