"""The square root the GPU pair force uses (hardware root + fma residual
correction, ctc_amd/csrc/udx_dev.h sqrt_rn) is correctly rounded on the
domains the force evaluates it on, for every value a 1-ulp hardware root can
return: r = sqrt(r2) with r2 in [1e-6, 1), and the w_R roots of 1 - r in
[2^-24, 1].  Exhaustive over the floats of each domain."""
import ctypes as C

import numpy as np

import oracle as O


def _verify(lo, hi, correct=1):
    f = O.lib().orc_verify_sqrt_rn
    f.restype = C.c_long
    f.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int]
    return f(lo, hi, 8, correct)


def test_sqrt_rn_on_the_pair_distance_domain():
    assert _verify(1e-6, 1.0) == 0


def test_sqrt_rn_on_the_weight_domain():
    assert _verify(2.0 ** -24, 1.0) == 0


def test_verifier_catches_uncorrected_roots():
    """control: the raw neighbours of the correctly rounded root fail, two
    per float of the range"""
    lo, hi = np.float32(0.25), np.float32(0.5)
    nfloat = int(hi.view(np.uint32)) - int(lo.view(np.uint32)) + 1
    assert _verify(lo, hi, correct=0) == 2 * nfloat
