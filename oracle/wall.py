"""ctypes bindings of the wall part of the C oracle (oracle/udx_oracle_wall.c).

Test infrastructure only.
"""
import ctypes as C

import numpy as np

from .oracle import Dpd, F32P, I32P, _f32, _i32, _ptr, lib

MARGIN = 6


class SdfStruct(C.Structure):
    _fields_ = [("L", C.c_int * 3), ("T", C.c_int * 3), ("o", C.c_float * 3), ("s", C.c_float * 3),
                ("cheap_threshold", C.c_float), ("tex", C.POINTER(C.c_float))]


class Wvel(C.Structure):
    """wall velocity (orc_wvel): type 0 resting, 1 constant u, 2 shear, 3 Hele-Shaw"""
    _fields_ = [("type", C.c_int), ("u", C.c_float * 3), ("gdot", C.c_float), ("vdir", C.c_int), ("gdir", C.c_int),
                ("half", C.c_int), ("hu", C.c_float), ("hh", C.c_float)]


def wvel(kind="none", u=(0.0, 0.0, 0.0), gdot=0.0, vdir=0, gdir=2, half=0, hu=0.0, hh=1.0):
    w = Wvel()
    w.type = {"none": 0, "constant": 1, "shear": 2, "hele shaw": 3}[kind]
    for c in range(3):
        w.u[c] = u[c]
    w.gdot, w.vdir, w.gdir, w.half, w.hu, w.hh = gdot, vdir, gdir, half, hu, hh
    return w


def wvel_shear_sin(gdot, w, t, vdir=0, gdir=2, half=0):
    """the shear view of a "shear sin" wall at time t (float32 t)"""
    rate = _lib().orc_wvel_sin_rate(np.float32(gdot), np.float32(w), np.float32(t))
    return wvel("shear", gdot=rate, vdir=vdir, gdir=gdir, half=half)


def _wv(w):
    return C.byref(w) if w is not None else None


_declared = False


def _lib():
    global _declared
    L = lib()
    if not _declared:
        P = C.POINTER(SdfStruct)
        sig = {
            "orc_sdf_texture_size": (None, [I32P, I32P]),
            "orc_sdf_gen": (None, [I32P, F32P, I32P, F32P, P]),
            "orc_sdf_sample_box": (None, [I32P, F32P, I32P, F32P, F32P]),
            "orc_sdf_eval": (C.c_float, [P, C.c_float, C.c_float, C.c_float]),
            "orc_sdf_grad": (None, [P, C.c_float, C.c_float, C.c_float, F32P]),
            "orc_sdf_cheap": (C.c_float, [P, C.c_float, C.c_float, C.c_float]),
            "orc_sdf_far": (C.c_int, [P, C.c_float, C.c_float, C.c_float]),
            "orc_wall_freeze": (C.c_int, [P, I32P, F32P, F32P]),
            "orc_wall_images": (C.c_int, [I32P, C.c_int, F32P, F32P]),
            "orc_wall_forces": (C.c_long, [I32P, C.c_int, F32P, P, C.c_int, F32P, I32P, Dpd, C.c_float, F32P,
                                           C.POINTER(Wvel)]),
            "orc_wall_bounce": (None, [P, C.c_float, C.c_int, F32P, C.POINTER(Wvel)]),
            "orc_wvel_at": (None, [C.POINTER(Wvel), F32P, F32P]),
            "orc_wvel_sin_rate": (C.c_float, [C.c_float, C.c_float, C.c_float]),
            "orc_wall_repulse": (None, [P, C.c_float, C.c_int, F32P, F32P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _declared = True
    return L


def read_sdf_dat(path):
    """sdf.dat: "ex ey ez\\n" "nx ny nz\\n" then nx*ny*nz float32, x fastest
    (src/wall/sdf/field/imp/ini.h:19-43)."""
    with open(path, "rb") as f:
        ext = np.array([float(t) for t in f.readline().split()], np.float32)
        N = np.array([int(t) for t in f.readline().split()], np.int32)
        D = np.frombuffer(f.read(4 * int(np.prod(N))), dtype=np.float32).copy()
    return ext, N, D


def write_sdf_dat(path, ext, D):
    """D indexed [z, y, x]; writes the sdf.dat layout."""
    D = np.ascontiguousarray(D, np.float32)
    nz, ny, nx = D.shape
    with open(path, "wb") as f:
        f.write(("%f %f %f\n" % tuple(ext)).encode())
        f.write(("%d %d %d\n" % (nx, ny, nz)).encode())
        f.write(D.tobytes())


def plane_channel_sdf(L, N, half_width):
    """s(z) = |z| - h over the box (positive inside the walls), sampled at the
    cell centres of an N^3 grid over [0, L] (global coordinates)."""
    z = (np.arange(N) + 0.5) * (L[2] / N) - 0.5 * L[2]
    s = (np.abs(z) - half_width).astype(np.float32)
    return np.broadcast_to(s[:, None, None], (N, N, N)).copy()


class Sdf:
    def __init__(self, L, ext, N, D):
        self.L = _i32(L)
        T = np.zeros(3, np.int32)
        _lib().orc_sdf_texture_size(_ptr(self.L, I32P), _ptr(T, I32P))
        self.tex = np.zeros(int(np.prod(T)), np.float32)
        self.st = SdfStruct()
        self.st.tex = self.tex.ctypes.data_as(C.POINTER(C.c_float))
        _lib().orc_sdf_gen(_ptr(self.L, I32P), _ptr(_f32(ext), F32P), _ptr(_i32(N), I32P), _ptr(_f32(D), F32P),
                           C.byref(self.st))
        self.T = tuple(T)

    @property
    def cheap_threshold(self):
        return self.st.cheap_threshold

    def eval(self, r):
        r = np.asarray(r, np.float32).reshape(-1, 3)
        return np.array([_lib().orc_sdf_eval(C.byref(self.st), *map(float, p)) for p in r], np.float32)

    def cheap(self, r):
        r = np.asarray(r, np.float32).reshape(-1, 3)
        return np.array([_lib().orc_sdf_cheap(C.byref(self.st), *map(float, p)) for p in r], np.float32)

    def far(self, r):
        r = np.asarray(r, np.float32).reshape(-1, 3)
        return np.array([_lib().orc_sdf_far(C.byref(self.st), *map(float, p)) for p in r], bool)

    def grad(self, r):
        r = np.asarray(r, np.float32).reshape(-1, 3)
        out = np.zeros((len(r), 3), np.float32)
        for k, p in enumerate(r):
            g = np.zeros(3, np.float32)
            _lib().orc_sdf_grad(C.byref(self.st), *map(float, p), _ptr(g, F32P))
            out[k] = g
        return out

    def freeze(self, pp):
        pp = _f32(pp).reshape(-1, 6).copy()
        n = np.array([len(pp)], np.int32)
        wall = np.zeros_like(pp)
        nw = _lib().orc_wall_freeze(C.byref(self.st), _ptr(n, I32P), _ptr(pp, F32P), _ptr(wall, F32P))
        return pp[:int(n[0])].copy(), wall[:nw].copy()

    def bounce(self, pp, dt, wv=None):
        pp = _f32(pp).reshape(-1, 6).copy()
        _lib().orc_wall_bounce(C.byref(self.st), np.float32(dt), len(pp), _ptr(pp, F32P), _wv(wv))
        return pp

    def repulse(self, pp, lam, ff=None):
        pp = _f32(pp).reshape(-1, 6)
        ff = np.zeros((len(pp), 3), np.float32) if ff is None else _f32(ff).copy()
        _lib().orc_wall_repulse(C.byref(self.st), np.float32(lam), len(pp), _ptr(pp, F32P), _ptr(ff, F32P))
        return ff


def sample_box(L, ext, N, D):
    out = np.zeros(int(np.prod(L)), np.float32)
    _lib().orc_sdf_sample_box(_ptr(_i32(L), I32P), _ptr(_f32(ext), F32P), _ptr(_i32(N), I32P), _ptr(_f32(D), F32P),
                              _ptr(out, F32P))
    return out


def wall_images(L, wall):
    wall = _f32(wall).reshape(-1, 6)
    out = np.zeros((27 * max(len(wall), 1), 6), np.float32)
    k = _lib().orc_wall_images(_ptr(_i32(L), I32P), len(wall), _ptr(wall, F32P), _ptr(out, F32P))
    return out[:k].copy()


def margin_box(L):
    return tuple(int(x) + 2 * MARGIN for x in L)


def wall_forces(L, pp, sdf, wpp, wstarts, params, seed, ff=None, wv=None):
    pp = _f32(pp).reshape(-1, 6)
    wpp = _f32(wpp).reshape(-1, 6)
    ff = np.zeros((len(pp), 3), np.float32) if ff is None else _f32(ff).copy()
    npair = _lib().orc_wall_forces(_ptr(_i32(L), I32P), len(pp), _ptr(pp, F32P), C.byref(sdf.st), len(wpp),
                                   _ptr(wpp, F32P), _ptr(_i32(wstarts), I32P), Dpd(*params), np.float32(seed),
                                   _ptr(ff, F32P), _wv(wv))
    return ff, int(npair)


def wvel_at(wv, r):
    r = _f32(r).reshape(-1, 3)
    out = np.zeros_like(r)
    for k in range(len(r)):
        _lib().orc_wvel_at(_wv(wv), _ptr(r[k], F32P), _ptr(out[k], F32P))
    return out


def integrate_wall(L, pp, ff, dt, sdf=None, mass=1.0, accel=(0.0, 0.0, 0.0), wv=None, wrap=True):
    """Kick-drift, bounce-back (when sdf is given), periodic wrap (optional)."""
    L_ = _lib()
    if not hasattr(L_.orc_integrate_wall_ex, "_udx_decl"):
        L_.orc_integrate_wall_ex.restype = None
        L_.orc_integrate_wall_ex.argtypes = [I32P, C.c_int, C.c_float, C.c_float, F32P, F32P, F32P,
                                             C.POINTER(SdfStruct), C.POINTER(Wvel), C.c_int]
        L_.orc_integrate_wall_ex._udx_decl = True
    pp = _f32(pp).reshape(-1, 6).copy()
    L_.orc_integrate_wall_ex(_ptr(_i32(L), I32P), len(pp), np.float32(dt), np.float32(mass),
                             _ptr(_f32(accel), F32P), _ptr(_f32(ff), F32P), _ptr(pp, F32P),
                             C.byref(sdf.st) if sdf is not None else None, _wv(wv), 1 if wrap else 0)
    return pp


def flu_wall_step(L, pp, params, seeds28, dt, sdf, wpp, wstarts, mass=1.0, accel=(0.0, 0.0, 0.0), nthreads=1,
                  wv=None, half=False):
    """One wall-bounded solvent step: clist, DPD forces, wall forces, kick-drift,
    bounce-back, wrap.  Returns (cell-ordered new pp, solvent pairs, wall pairs).
    half=True: the solvent forces by the reference's half-shell traversal
    (the CPU baseline; not a parity reference)."""
    from . import oracle as O
    starts, counts, srt, _ = O.clist_build(L, pp)
    forces = O.dpd_forces_half if half else O.dpd_forces
    ff, npair = forces(L, srt, starts, counts, params, seeds28[:27], nthreads=nthreads)
    nw = 0
    if sdf is not None:
        ff, nw = wall_forces(L, srt, sdf, wpp, wstarts, params, seeds28[27], ff, wv)
    out = integrate_wall(L, srt, ff, dt, sdf, mass, accel, wv)
    return out, npair // 2, nw
