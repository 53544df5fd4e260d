"""ctypes bindings of the membrane bounce-back part of the C oracle
(oracle/udx_oracle_mbb.c).  Test infrastructure only."""
import ctypes as C

import numpy as np

from .oracle import F32P, I32P, _f32, _i32, _ptr, lib

U64P = C.POINTER(C.c_uint64)
BB_SUCCESS, BB_NOCROSS, BB_WTRIANGLE, BB_HFAIL = 0, 1, 2, 3
EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)

_declared = False


def _lib():
    global _declared
    L = lib()
    if not _declared:
        sig = {
            "orc_mbb_intersect": (C.c_int, [C.c_float] + [F32P] * 9),
            "orc_mbb_instances": (C.c_int, [I32P, C.c_int, C.c_int, F32P, I32P, F32P]),
            "orc_mbb_find": (None, [C.c_float, I32P, C.c_int, I32P, F32P, C.c_int, C.c_int, I32P, F32P, F32P, I32P,
                                    F32P, F32P, U64P]),
            "orc_mbb_inertia": (None, [C.c_float, F32P, F32P, F32P, F32P]),
            "orc_mbb_inverse": (None, [F32P, F32P]),
            "orc_mbb_nonzero": (C.c_int, [F32P]),
            "orc_mbb_m2v": (None, [C.c_float, F32P, F32P, F32P, F32P, F32P, F32P, F32P]),
            "orc_mbb_bounce": (C.c_long, [C.c_float, C.c_float, C.c_float, I32P, C.c_int, C.c_int, C.c_int, I32P,
                                          F32P, F32P, I32P, C.c_int, F32P, F32P, F32P, I32P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _declared = True
    return L


def _v3(a):
    return _f32(np.asarray(a, np.float32).reshape(3))


def intersect(dt, s1, s2, s3, vs1, vs2, vs3, r, v):
    """(state, h, u, v, side) of a particle r + t v against the moving
    triangle s_i + t vs_i"""
    out = np.zeros(4, np.float32)
    args = [_v3(x) for x in (s1, s2, s3, vs1, vs2, vs3, r, v)]
    st = _lib().orc_mbb_intersect(np.float32(dt), *[_ptr(a, F32P) for a in args], _ptr(out, F32P))
    return int(st), float(out[0]), float(out[1]), float(out[2]), float(out[3])


def instances(L, nm, nv, rc):
    rc = _f32(rc).reshape(-1, 6)
    im = np.zeros(8 * max(nm, 1), np.int32)
    sh = np.zeros((8 * max(nm, 1), 3), np.float32)
    n = _lib().orc_mbb_instances(_ptr(_i32(L), I32P), nm, nv, _ptr(rc, F32P), _ptr(im, I32P), _ptr(sh, F32P))
    return im[:n].copy(), sh[:n].copy()


def inertia(mass, a, b, c):
    out = np.zeros(6, np.float32)
    _lib().orc_mbb_inertia(np.float32(mass), _ptr(_v3(a), F32P), _ptr(_v3(b), F32P), _ptr(_v3(c), F32P),
                           _ptr(out, F32P))
    return out


def inverse(I):
    out = np.zeros(6, np.float32)
    _lib().orc_mbb_inverse(_ptr(_f32(I), F32P), _ptr(out, F32P))
    return out


def nonzero(m):
    return bool(_lib().orc_mbb_nonzero(_ptr(_f32(m), F32P)))


def m2v(mass, m, a, b, c):
    out = np.zeros((3, 3), np.float32)
    _lib().orc_mbb_m2v(np.float32(mass), _ptr(_f32(m), F32P), _ptr(_v3(a), F32P), _ptr(_v3(b), F32P),
                       _ptr(_v3(c), F32P), _ptr(out[0], F32P), _ptr(out[1], F32P), _ptr(out[2], F32P))
    return out


def bounce(dt, flu_mass, rbc_mass, L, nm, tt, rp, rc, starts, pp0, pp):
    """The whole bounce: returns (pp bounced, rc with updated velocities,
    momentum per triangle (nm nt, 6), triangle per particle (-1: none),
    number of bounced particles)."""
    tt = _i32(tt).reshape(-1, 3)
    nt = len(tt)
    rc = _f32(rc).reshape(-1, 6).copy()
    nv = len(rc) // max(nm, 1)
    pp = _f32(pp).reshape(-1, 6).copy()
    pp0 = _f32(pp0).reshape(-1, 6)
    rp = _f32(rp).reshape(-1, 3)
    mm = np.zeros((max(nm * nt, 1), 6), np.float32)
    col = np.zeros(max(len(pp), 1), np.int32)
    ncol = _lib().orc_mbb_bounce(np.float32(dt), np.float32(flu_mass), np.float32(rbc_mass), _ptr(_i32(L), I32P),
                                 nm, nt, nv, _ptr(tt, I32P), _ptr(rp, F32P), _ptr(rc, F32P),
                                 _ptr(_i32(starts), I32P), len(pp), _ptr(pp0, F32P), _ptr(pp, F32P),
                                 _ptr(mm, F32P), _ptr(col, I32P))
    return pp, rc, mm[:nm * nt], col[:len(pp)], int(ncol)
