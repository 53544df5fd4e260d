"""ctypes bindings of the C oracle (oracle/udx_oracle.c).

Test infrastructure only: this module checks the HIP path, it is never part of
it.  All arrays are numpy; particles are AoS float32 rows (x, y, z, vx, vy, vz)
as in the reference's Particle (src/inc/type.h:6-8).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle_udx.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build(ref=True):
    """Compile the oracle (and the reference's own buildable pieces)."""
    targets = ["oracle"]
    if ref and os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def _p(t):
    return C.POINTER(t)


F32P, I32P, U32P, F64P = _p(C.c_float), _p(C.c_int), _p(C.c_uint32), _p(C.c_double)


class Kiss(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("c", C.c_uint32)]


class Dpd(C.Structure):
    _fields_ = [("a", C.c_float), ("g", C.c_float), ("s", C.c_float), ("spow", C.c_float)]


def _declare(L):
    sig = {
        "orc_kiss_ini": (None, [_p(Kiss), C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_kiss_get_int": (C.c_uint32, [_p(Kiss)]),
        "orc_kiss_get": (C.c_float, [_p(Kiss)]),
        "orc_kiss_fill": (None, [C.c_int] * 5 + [F32P]),
        "orc_fma_rz": (C.c_float, [C.c_float] * 3),
        "orc_fma_rz_fenv": (C.c_float, [C.c_float] * 3),
        "orc_fma_rz_batch": (None, [C.c_int, F32P, F32P, F32P, F32P, C.c_int]),
        "orc_mean0var1ii": (C.c_float, [C.c_float, C.c_int, C.c_int]),
        "orc_mean0var1uu": (C.c_float, [C.c_float, C.c_uint32, C.c_uint32]),
        "orc_rng_ii_batch": (None, [C.c_float, C.c_int, I32P, I32P, F32P]),
        "orc_rng_uu_batch": (None, [C.c_float, C.c_int, U32P, U32P, F32P]),
        "orc_seeds_table": (None, [C.c_int, F32P]),
        "orc_frag_i2d3": (None, [C.c_int, I32P]),
        "orc_frag_d2i": (C.c_int, [C.c_int] * 3),
        "orc_frag_anti": (C.c_int, [C.c_int]),
        "orc_frag_table_size": (C.c_int, [I32P]),
        "orc_frag_tables": (None, [I32P, I32P, I32P, I32P]),
        "orc_gen_random": (None, [C.c_int] * 5 + [F32P]),
        "orc_clist_build": (C.c_int, [I32P, C.c_int, F32P, I32P, I32P, F32P, I32P]),
        "orc_clist_build_sub": (C.c_int, [I32P, C.c_int, C.c_int, F32P, I32P, I32P, I32P, F32P, I32P]),
        "orc_slab_of": (C.c_int, [C.c_int, C.c_int, C.c_float]),
        "orc_dpd_sigma": (C.c_float, [C.c_float] * 3),
        "orc_pair_force": (C.c_int, [_p(Dpd), F32P, F32P, C.c_float, F32P]),
        "orc_dpd_forces": (C.c_long, [I32P, C.c_int, F32P, I32P, I32P, Dpd, F32P, C.c_int,
                                      C.c_int, F32P]),
        "orc_dpd_forces_stress": (C.c_long, [I32P, C.c_int, F32P, I32P, I32P, Dpd, F32P, C.c_int,
                                             C.c_int, F32P, F32P]),
        "orc_dpd_forces_half": (C.c_long, [I32P, C.c_int, F32P, I32P, I32P, Dpd, F32P, C.c_int, F32P]),
        "orc_dpd_forces_sum64": (C.c_long, [I32P, C.c_int, F32P, I32P, I32P, Dpd, F32P, C.c_int,
                                            C.c_int, F32P]),
        "orc_dpd_forces_f64": (C.c_long, [I32P, C.c_int, F32P, I32P, I32P, Dpd, F32P, C.c_int,
                                          C.c_int, F64P]),
        "orc_dpd_forces_half_sum64": (C.c_long, [I32P, C.c_int, F32P, I32P, I32P, Dpd, F32P, C.c_int, F32P]),
        "orc_integrate": (None, [I32P, C.c_int, C.c_float, C.c_float, F32P, F32P, F32P]),
        "orc_move": (None, [C.c_int, C.c_float, C.c_float, F32P, F32P]),
        "orc_diag": (None, [C.c_int, F32P, F64P]),
        "orc_flu_step": (C.c_long, [I32P, C.c_int, F32P, F32P, I32P, I32P, F32P, Dpd, F32P,
                                    C.c_float, C.c_float, F32P, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _L(L):
    return _i32(L)


# --------------------------------------------------------------------------
def kiss_fill(x, y, z, c, n):
    out = np.empty(n, np.float32)
    lib().orc_kiss_fill(x, y, z, c, n, _ptr(out, F32P))
    return out


def kiss_ints(x, y, z, c, n):
    k = Kiss()
    lib().orc_kiss_ini(C.byref(k), x, y, z, c)
    return np.array([lib().orc_kiss_get_int(C.byref(k)) for _ in range(n)], np.uint32)


def fma_rz(a, b, c, fenv=False):
    a, b, c = _f32(a), _f32(b), _f32(c)
    out = np.empty_like(a)
    lib().orc_fma_rz_batch(a.size, _ptr(a, F32P), _ptr(b, F32P), _ptr(c, F32P), _ptr(out, F32P),
                           1 if fenv else 0)
    return out


def rng_ii(seed, u, v):
    u, v = _i32(u), _i32(v)
    out = np.empty(u.size, np.float32)
    lib().orc_rng_ii_batch(np.float32(seed), u.size, _ptr(u, I32P), _ptr(v, I32P), _ptr(out, F32P))
    return out


def rng_uu(seed, u, v):
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    out = np.empty(u.size, np.float32)
    lib().orc_rng_uu_batch(np.float32(seed), u.size, _ptr(u, U32P), _ptr(v, U32P), _ptr(out, F32P))
    return out


def seeds_table(nsteps):
    out = np.empty((nsteps, 27), np.float32)
    lib().orc_seeds_table(nsteps, _ptr(out, F32P))
    return out


def frag_i2d3(fid):
    d = np.zeros(3, np.int32)
    lib().orc_frag_i2d3(fid, _ptr(d, I32P))
    return tuple(int(x) for x in d)


def frag_d2i(x, y, z):
    return lib().orc_frag_d2i(x, y, z)


def frag_anti(fid):
    return lib().orc_frag_anti(fid)


def frag_tables(L, counts):
    L = _L(L)
    counts = _i32(counts)
    size = lib().orc_frag_table_size(_ptr(L, I32P))
    fss = np.empty(size, np.int32)
    off = np.empty(27, np.int32)
    lib().orc_frag_tables(_ptr(L, I32P), _ptr(counts, I32P), _ptr(fss, I32P), _ptr(off, I32P))
    return fss, off


def gen_random(X, Y, Z, n, roundtrip=True):
    pp = np.empty((n, 6), np.float32)
    lib().orc_gen_random(X, Y, Z, n, 1 if roundtrip else 0, _ptr(pp, F32P))
    return pp


# x slabs per cell of the solvent's cell lists (ORC_SOLVENT_SUB)
SOLVENT_SUB = 4


def clist_build(L, pp, sub=SOLVENT_SUB, slabs=False):
    """Counting sort by cell, then x slab (sub per cell), then input index.
    Returns (starts, counts, cell-ordered pp, order) and, with slabs=True,
    the slab starts (ncells * sub + 1) as a fifth item.  Solvent cell lists
    use sub = SOLVENT_SUB; the wall particles' margin grid and the FSI
    instances use sub = 1."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    nc = int(L[0] * L[1] * L[2])
    starts = np.empty(nc + 1, np.int32)
    counts = np.empty(nc, np.int32)
    bstarts = np.empty(nc * sub + 1, np.int32)
    out = np.empty_like(pp)
    order = np.empty(max(n, 1), np.int32)
    nout = lib().orc_clist_build_sub(_ptr(L, I32P), sub, n, _ptr(pp, F32P), _ptr(starts, I32P),
                                     _ptr(counts, I32P), _ptr(bstarts, I32P), _ptr(out, F32P), _ptr(order, I32P))
    res = (starts, counts, out[:nout].copy(), order[:nout].copy())
    return res + (bstarts,) if slabs else res


def dpd_sigma(kBT, g, dt):
    return lib().orc_dpd_sigma(kBT, g, dt)


def pair_force(params, a, b, rnd):
    a, b = _f32(a), _f32(b)
    f = np.zeros(3, np.float32)
    p = Dpd(*params)
    is_pair = lib().orc_pair_force(C.byref(p), _ptr(a, F32P), _ptr(b, F32P), np.float32(rnd),
                                   _ptr(f, F32P))
    return f, bool(is_pair)


def dpd_forces(L, pp, starts, counts, params, seeds, periodic=True, nthreads=1, ff=None):
    """Forces on cell-ordered particles; returns (ff (n,3), ordered pair count)."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    ff = np.zeros((n, 3), np.float32) if ff is None else _f32(ff).copy()
    seeds = _f32(seeds)
    np_ = lib().orc_dpd_forces(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(_i32(starts), I32P),
                               _ptr(_i32(counts), I32P), Dpd(*params), _ptr(seeds, F32P),
                               1 if periodic else 0, nthreads, _ptr(ff, F32P))
    return ff, int(np_)


def dpd_forces_sum64(L, pp, starts, counts, params, seeds, periodic=True, nthreads=1, ff=None):
    """dpd_forces with exact sums (pair forces rounded to multiples of 2^-32,
    summed exactly, rounded to float once): the production kernel's forces."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    ff = np.zeros((n, 3), np.float32) if ff is None else _f32(ff).copy()
    np_ = lib().orc_dpd_forces_sum64(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(_i32(starts), I32P),
                                     _ptr(_i32(counts), I32P), Dpd(*params), _ptr(_f32(seeds), F32P),
                                     1 if periodic else 0, nthreads, _ptr(ff, F32P))
    return ff, int(np_)


def dpd_forces_half_sum64(L, pp, starts, counts, params, seeds, nthreads=1):
    """The half-shell traversal with exact sums (periodic box); as dpd_forces."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    ff = np.zeros((n, 3), np.float32)
    np_ = lib().orc_dpd_forces_half_sum64(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(_i32(starts), I32P),
                                          _ptr(_i32(counts), I32P), Dpd(*params), _ptr(_f32(seeds), F32P),
                                          nthreads, _ptr(ff, F32P))
    return ff, int(np_)


def dpd_forces_f64(L, pp, starts, counts, params, seeds, periodic=True, nthreads=1):
    """The float pair forces of dpd_forces summed in double: (ff64 (n,3), ordered pairs)."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    ff = np.zeros((n, 3), np.float64)
    np_ = lib().orc_dpd_forces_f64(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(_i32(starts), I32P),
                                   _ptr(_i32(counts), I32P), Dpd(*params), _ptr(_f32(seeds), F32P),
                                   1 if periodic else 0, nthreads, _ptr(ff, F64P))
    return ff, int(np_)


def dpd_forces_stress(L, pp, starts, counts, params, seeds, periodic=True, nthreads=1):
    """Forces and per-particle stresses (n, 6: xx, xy, xz, yy, yz, zz);
    returns (ff, ss, ordered pair count)."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    ff = np.zeros((n, 3), np.float32)
    ss = np.zeros((n, 6), np.float32)
    np_ = lib().orc_dpd_forces_stress(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(_i32(starts), I32P),
                                      _ptr(_i32(counts), I32P), Dpd(*params), _ptr(_f32(seeds), F32P),
                                      1 if periodic else 0, nthreads, _ptr(ff, F32P), _ptr(ss, F32P))
    return ff, ss, int(np_)


def dpd_forces_half(L, pp, starts, counts, params, seeds, nthreads=1, ff=None):
    """The reference's half-shell traversal (CPU baseline); as dpd_forces."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6)
    n = pp.shape[0]
    ff = np.zeros((n, 3), np.float32) if ff is None else _f32(ff).copy()
    np_ = lib().orc_dpd_forces_half(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(_i32(starts), I32P),
                                    _ptr(_i32(counts), I32P), Dpd(*params), _ptr(_f32(seeds), F32P), nthreads,
                                    _ptr(ff, F32P))
    return ff, int(np_)


def integrate(L, pp, ff, dt, mass=1.0, accel=(0.0, 0.0, 0.0), wrap=True):
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6).copy()
    ff = _f32(ff)
    acc = _f32(accel)
    f = lib().orc_integrate_ex
    f.restype = None
    f.argtypes = [I32P, C.c_int, C.c_float, C.c_float, F32P, F32P, F32P, C.c_int]
    f(_ptr(L, I32P), pp.shape[0], np.float32(dt), np.float32(mass), _ptr(acc, F32P), _ptr(ff, F32P),
      _ptr(pp, F32P), 1 if wrap else 0)
    return pp


def wrap(L, pp):
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6).copy()
    f = lib().orc_wrap
    f.restype = None
    f.argtypes = [I32P, C.c_int, F32P]
    f(_ptr(L, I32P), pp.shape[0], _ptr(pp, F32P))
    return pp


def move(pp, ff, dt, mass):
    """kick-drift without the periodic wrap (membrane vertices)"""
    pp = _f32(pp).reshape(-1, 6).copy()
    lib().orc_move(pp.shape[0], np.float32(dt), np.float32(mass), _ptr(_f32(ff), F32P), _ptr(pp, F32P))
    return pp


def diag(pp):
    pp = _f32(pp).reshape(-1, 6)
    out = np.zeros(5, np.float64)
    lib().orc_diag(pp.shape[0], _ptr(pp, F32P), _ptr(out, F64P))
    return out


def flu_step(L, pp, params, seeds, dt, mass=1.0, accel=(0.0, 0.0, 0.0), nthreads=1):
    """One step (clist, forces, integrate); returns (new cell-ordered pp, pairs)."""
    L = _L(L)
    pp = _f32(pp).reshape(-1, 6).copy()
    n = pp.shape[0]
    nc = int(L[0] * L[1] * L[2])
    scratch = np.empty_like(pp)
    starts = np.empty(nc + 1, np.int32)
    counts = np.empty(nc, np.int32)
    ff = np.empty((n, 3), np.float32)
    acc = _f32(accel)
    npairs = lib().orc_flu_step(_ptr(L, I32P), n, _ptr(pp, F32P), _ptr(scratch, F32P),
                                _ptr(starts, I32P), _ptr(counts, I32P), _ptr(ff, F32P),
                                Dpd(*params), _ptr(_f32(seeds), F32P), np.float32(dt),
                                np.float32(mass), _ptr(acc, F32P), nthreads)
    return pp, int(npairs)


# --------------------------------------------------------------------------
# the reference's own code (oracle/_ref), where it was built
def ref_rnd_available():
    return os.path.exists(os.path.join(REF_DIR, "librefrnd.so"))


def ref_rnd_fill(x, y, z, c, n):
    """rnd_ini/rnd_get of src/math/rnd/imp.cpp, compiled from the reference."""
    L = C.CDLL(os.path.join(REF_DIR, "librefrnd.so"))
    ini = getattr(L, "_Z7rnd_iniiiiiPP7RNDunif")
    get = getattr(L, "_Z7rnd_getP7RNDunif")
    fin = getattr(L, "_Z7rnd_finP7RNDunif")
    ini.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_void_p)]
    get.argtypes = [C.c_void_p]
    get.restype = C.c_float
    fin.argtypes = [C.c_void_p]
    h = C.c_void_p()
    ini(x, y, z, c, C.byref(h))
    out = np.array([get(h) for _ in range(n)], np.float32)
    fin(h)
    return out


def ref_gen_random_available():
    return os.path.exists(os.path.join(REF_DIR, "gen_random"))


def ref_gen_random_text(X, Y, Z, n):
    exe = os.path.join(REF_DIR, "gen_random")
    return subprocess.run([exe, str(X), str(Y), str(Z), str(n)], check=True,
                          capture_output=True, text=True).stdout
