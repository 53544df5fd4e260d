"""ctypes bindings of libudx.so and thin Python mirrors of the reference's
module interfaces (clist, fluforces, scheme, diag, solvent solver).

Device memory and streams come from torch (ROCm); the arithmetic is entirely
in the HIP kernels of libudx.  Particle tensors are float32 (n, 6) rows
(x, y, z, vx, vy, vz), forces float32 (n, 3), as in src/inc/type.h:6-22.
"""
import ctypes as C
import os

from . import LIB_PATH

_lib = None


class UdxError(RuntimeError):
    pass


class Rnd(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("c", C.c_uint32)]


class PairDpd(C.Structure):
    _fields_ = [("a", C.c_float), ("g", C.c_float), ("s", C.c_float), ("spow", C.c_float)]


class PairDpdLj(C.Structure):
    """udx_pair_dpd_lj (PairDPDLJ, src/pair/type.h:22-25)"""
    _fields_ = [(k, C.c_float) for k in ("a", "g", "s", "spow", "ljs", "lje")]


class FluConf(C.Structure):
    _fields_ = [("L", C.c_int * 3), ("maxp", C.c_int), ("dpd", PairDpd), ("dt", C.c_float),
                ("mass", C.c_float), ("accel", C.c_float * 3), ("mode", C.c_int)]


class EnsBox(C.Structure):
    _fields_ = [("L", C.c_int * 3), ("n", C.c_int), ("dpd", PairDpd), ("dt", C.c_float), ("mass", C.c_float),
                ("accel", C.c_float * 3)]


class RbcParams(C.Structure):
    """src/rbc/params/imp.h:3-9"""
    _fields_ = [(k, C.c_float) for k in ("gammaC", "gammaT", "kBT", "kb", "phi", "ks", "x0", "mpow", "ka", "kd",
                                          "kv", "totArea", "totVolume")]


class WallVel(C.Structure):
    """udx_wall_vel: the wall velocity types of src/wall/wvel/imp/type.h"""
    _fields_ = [("type", C.c_int), ("u", C.c_float * 3), ("gdot", C.c_float), ("vdir", C.c_int), ("gdir", C.c_int),
                ("half", C.c_int), ("w", C.c_float), ("hs_u", C.c_float), ("hs_h", C.c_float)]


WVEL_TYPES = {"none": 0, "constant": 1, "shear": 2, "shear sin": 3, "hele shaw": 4}


def wall_vel(kind="none", u=(0.0, 0.0, 0.0), gdot=0.0, vdir=0, gdir=2, half=0, w=0.0, hs_u=0.0, hs_h=1.0):
    """wvel_set_cste / wvel_set_shear / wvel_set_shear_sin / wvel_set_hs
    (src/wall/wvel/imp.h); kind "none" keeps the walls at rest"""
    v = WallVel()
    v.type = WVEL_TYPES[kind]
    for c in range(3):
        v.u[c] = u[c]
    v.gdot, v.vdir, v.gdir, v.half, v.w, v.hs_u, v.hs_h = gdot, vdir, gdir, half, w, hs_u, hs_h
    return v


_VP, _I, _F = C.c_void_p, C.c_int, C.c_float
_SIGS = {
    "udx_last_error": (C.c_char_p, []),
    "udx_rnd_ini": (None, [_I, _I, _I, _I, C.POINTER(Rnd)]),
    "udx_rnd_get": (_F, [C.POINTER(Rnd)]),
    "udx_pair_compute_dpd_sigma": (None, [_F, _F, C.POINTER(PairDpd)]),
    "udx_ctx_ini": (_I, [_VP, C.POINTER(_VP)]),
    "udx_ctx_fin": (_I, [_VP]),
    "udx_ctx_sync": (_I, [_VP]),
    "udx_rnd_branch": (_I, [_VP, _F, _I, _VP, _VP, _I, _VP]),
    "udx_selfcheck": (_I, [_VP, _I, _F, _F, C.POINTER(C.c_ulonglong)]),
    "udx_clist_ini": (_I, [_VP, _I, _I, _I, _I, C.POINTER(_VP)]),
    "udx_clist_fin": (_I, [_VP]),
    "udx_clist_build": (_I, [_VP, _I, _VP, _VP, C.POINTER(_I)]),
    "udx_clist_starts": (_VP, [_VP]),
    "udx_clist_counts": (_VP, [_VP]),
    "udx_clist_order": (_VP, [_VP]),
    "udx_clist_build_noalloc": (_I, [_VP, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
    "udx_scan_ini": (_I, [_VP, _I, C.POINTER(_VP)]),
    "udx_scan_apply": (_I, [_VP, _VP, _I, _VP]),
    "udx_scan_fin": (_I, [_VP]),
    "udx_fluforces_ini": (_I, [_VP, _I, _I, _I, _I, C.POINTER(_VP)]),
    "udx_fluforces_fin": (_I, [_VP]),
    "udx_fluforces_prepare": (_I, [_VP, _I, _VP]),
    "udx_fluforces_apply": (_I, [_VP, C.POINTER(PairDpd), _VP, _I, _VP]),
    "udx_fluforces_apply_seeds": (_I, [_VP, C.POINTER(PairDpd), _VP, _I, C.POINTER(_F), _VP]),
    "udx_fluforces_set_mode": (_I, [_VP, _I]),
    "udx_fluforces_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong), _I]),
    "udx_scheme_move_apply": (_I, [_VP, _F, _F, _I, _VP, _VP]),
    "udx_bforce_apply_cste": (_I, [_VP, _F, C.POINTER(_F), _I, _VP]),
    "udx_wrap_pp": (_I, [_VP, _I, _I, _I, _I, _VP]),
    "udx_diag_part": (_I, [_VP, _I, _VP, C.POINTER(C.c_double)]),
    "udx_flu_ini": (_I, [_VP, C.POINTER(FluConf), C.POINTER(_VP)]),
    "udx_flu_fin": (_I, [_VP]),
    "udx_flu_set_pp": (_I, [_VP, _I, _VP]),
    "udx_flu_get_pp": (_I, [_VP, _VP, C.POINTER(_I)]),
    "udx_flu_get_ff": (_I, [_VP, _VP]),
    "udx_flu_run": (_I, [_VP, _I, C.POINTER(_F)]),
    "udx_flu_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong), _I]),
    "udx_flu_timing": (_I, [_VP, _I, C.POINTER(_F)]),
    "udx_flu_wall_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong)]),
    "udx_flu_set_wall": (_I, [_VP, _VP]),
    "udx_sdf_ini": (_I, [_VP, _I, _I, _I, C.POINTER(_VP)]),
    "udx_sdf_fin": (_I, [_VP]),
    "udx_sdf_gen": (_I, [_VP, C.POINTER(_F), C.POINTER(_I), C.POINTER(_F)]),
    "udx_sdf_gen_file": (_I, [_VP, C.c_char_p]),
    "udx_sdf_size": (_I, [_VP, C.POINTER(_I), C.POINTER(_F)]),
    "udx_sdf_texture": (_VP, [_VP]),
    "udx_sdf_probe": (_I, [_VP, _I, _VP, _I, _VP]),
    "udx_wall_gen": (_I, [_VP, _VP, C.POINTER(_I), _VP, C.POINTER(_VP)]),
    "udx_wall_fin": (_I, [_VP]),
    "udx_wall_counts": (_I, [_VP, C.POINTER(_I), C.POINTER(_I)]),
    "udx_wall_pp": (_VP, [_VP]),
    "udx_wall_starts": (_VP, [_VP]),
    "udx_wall_set_mode": (_I, [_VP, _I]),
    "udx_wall_force": (_I, [_VP, C.POINTER(PairDpd), _I, _VP, _VP]),
    "udx_wall_force_seed": (_I, [_VP, C.POINTER(PairDpd), _I, _VP, _F, _VP]),
    "udx_wall_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong), _I]),
    "udx_wall_bounce": (_I, [_VP, _F, _I, _VP]),
    "udx_wall_set_vel": (_I, [_VP, C.POINTER(WallVel)]),
    "udx_wall_set_time": (_I, [_VP, _F]),
    "udx_wall_bounce_wall": (_I, [_VP, _F, _I, _VP]),
    "udx_flu_set_time": (_I, [_VP, _F]),
    "udx_wall_vel_set_conf": (_I, [_VP, C.POINTER(WallVel)]),
    "udx_wall_repulse_set_conf": (_I, [_VP, C.c_char_p, C.POINTER(_F)]),
    "udx_restart_write_pp": (_I, [C.c_char_p, C.c_char_p, _I, C.c_long, _VP]),
    "udx_restart_read_pp": (_I, [C.c_char_p, C.c_char_p, _I, _I, C.POINTER(_I), _VP]),
    "udx_restart_write_ii": (_I, [C.c_char_p, C.c_char_p, _I, C.c_long, _VP]),
    "udx_restart_read_ii": (_I, [C.c_char_p, C.c_char_p, _I, _I, C.POINTER(_I), _VP]),
    "udx_flu_strt_dump": (_I, [_VP, C.c_char_p, _I]),
    "udx_flu_strt_quants": (_I, [_VP, C.c_char_p, _I]),
    "udx_wall_strt_dump_templ": (_I, [_VP, C.c_char_p]),
    "udx_wall_strt_gen": (_I, [_VP, _VP, C.c_char_p, _I, C.POINTER(_VP)]),
    "udx_mbb_ini": (_I, [_VP, _I, _I, C.POINTER(_I), _I, _I, C.POINTER(_VP)]),
    "udx_mbb_fin": (_I, [_VP]),
    "udx_mbb_bounce": (_I, [_VP, _F, _F, _F, _I, _I, _I, _I, _VP, _VP, _VP, _I, _VP, _VP, _VP]),
    "udx_mbb_count": (_I, [_VP, C.POINTER(C.c_ulonglong)]),
    "udx_mbb_collisions": (_VP, [_VP]),
    "udx_mbb_set_pair_capacity": (_I, [_VP, _I]),
    "udx_flu_set_rbc_bounce": (_I, [_VP, _VP]),
    "udx_cnt_ini": (_I, [_VP, _I, _I, _I, _I, C.POINTER(_VP)]),
    "udx_cnt_fin": (_I, [_VP]),
    "udx_cnt_apply": (_I, [_VP, C.POINTER(PairDpdLj), _I, _VP, _VP]),
    "udx_cnt_apply_seed": (_I, [_VP, C.POINTER(PairDpdLj), _I, _VP, _F, _F, _VP]),
    "udx_cnt_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong), _I]),
    "udx_flu_set_rbc_contact": (_I, [_VP, _VP, C.POINTER(PairDpdLj)]),
    "udx_flu_set_rbc_repulsion": (_I, [_VP, _F]),
    "udx_wall_repulse": (_I, [_VP, _F, _I, _VP, _VP]),
    "udx_rbc_ini": (_I, [_VP, _I, _I, C.POINTER(_I), _I, C.POINTER(_VP)]),
    "udx_rbc_fin": (_I, [_VP]),
    "udx_rbc_md": (_I, [_VP, C.POINTER(_I)]),
    "udx_rbc_adjacency": (_I, [_VP, C.POINTER(_I), C.POINTER(_I)]),
    "udx_rbc_set_stressful": (_I, [_VP, _F]),
    "udx_rbc_set_stressfree": (_I, [_VP, C.POINTER(_F)]),
    "udx_rbc_set_bending": (_I, [_VP, _I]),
    "udx_rbc_set_bending_type": (_I, [_VP, _I]),
    "udx_rbc_area_volume": (_I, [_VP, _I, _VP, _VP]),
    "udx_rbc_av": (_VP, [_VP]),
    "udx_rbc_force_apply": (_I, [_VP, C.POINTER(RbcParams), _F, _I, _VP, _VP]),
    "udx_rbc_stretch_apply": (_I, [_VP, _I, _I, _VP, _VP]),
    "udx_rbc_force_adjoint": (_I, [_VP, C.POINTER(RbcParams), _I, _VP, _VP, _VP, C.POINTER(C.c_double)]),
    "udx_flu_set_graph": (_I, [_VP, _I]),
    "udx_fluforces_apply_stress": (_I, [_VP, C.POINTER(PairDpd), _VP, _I, C.POINTER(_F), _VP, _VP]),
    "udx_flu_set_rbc": (_I, [_VP, _VP, _VP, C.POINTER(RbcParams), C.POINTER(PairDpd), _F, _I, _I, _VP]),
    "udx_rbc_distribute": (_I, [_VP, _I, _I, _I, _I, _VP]),
    "udx_fsi_ini": (_I, [_VP, _I, _I, _I, _I, C.POINTER(_VP)]),
    "udx_fsi_fin": (_I, [_VP]),
    "udx_fsi_set_mode": (_I, [_VP, _I]),
    "udx_fsi_bind_solvent": (_I, [_VP, _VP, _VP, _I, _VP]),
    "udx_fsi_apply": (_I, [_VP, C.POINTER(PairDpd), _I, _VP, _VP]),
    "udx_fsi_apply_seeds": (_I, [_VP, C.POINTER(PairDpd), _I, _VP, _VP, C.POINTER(_F)]),
    "udx_fsi_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong), _I]),
    "udx_conf_ini": (_I, [C.POINTER(_VP)]),
    "udx_conf_fin": (_I, [_VP]),
    "udx_conf_set_include_dir": (_I, [_VP, C.c_char_p]),
    "udx_conf_read": (_I, [_VP, C.c_char_p, C.c_char_p, C.c_char_p]),
    "udx_conf_has": (_I, [_VP, C.c_char_p]),
    "udx_conf_lookup_int": (_I, [_VP, C.c_char_p, C.POINTER(_I)]),
    "udx_conf_lookup_float": (_I, [_VP, C.c_char_p, C.POINTER(_F)]),
    "udx_conf_lookup_bool": (_I, [_VP, C.c_char_p, C.POINTER(_I)]),
    "udx_conf_lookup_string": (_I, [_VP, C.c_char_p, C.POINTER(C.c_char_p)]),
    "udx_conf_lookup_vint": (_I, [_VP, C.c_char_p, _I, C.POINTER(_I), C.POINTER(_I)]),
    "udx_conf_lookup_vfloat": (_I, [_VP, C.c_char_p, _I, C.POINTER(_I), C.POINTER(_F)]),
    "udx_conf_lookup_vstring": (_I, [_VP, C.c_char_p, _I, C.POINTER(_I), C.POINTER(C.c_char_p)]),
    "udx_conf_lookup_int3": (_I, [_VP, C.c_char_p, C.POINTER(_I)]),
    "udx_conf_lookup_float3": (_I, [_VP, C.c_char_p, C.POINTER(_F)]),
    "udx_conf_set_int": (_I, [_VP, C.c_char_p, _I]),
    "udx_conf_set_float": (_I, [_VP, C.c_char_p, _F]),
    "udx_conf_set_bool": (_I, [_VP, C.c_char_p, _I]),
    "udx_conf_set_string": (_I, [_VP, C.c_char_p, C.c_char_p]),
    "udx_pair_set_conf": (_I, [_VP, C.c_char_p, _F, _F, C.POINTER(PairDpd)]),
    "udx_rbc_params_set_conf": (_I, [_VP, C.c_char_p, C.POINTER(RbcParams)]),
    "udx_ens_ini": (_I, [_VP, _I, C.POINTER(EnsBox), C.POINTER(_VP)]),
    "udx_ens_fin": (_I, [_VP]),
    "udx_ens_set_mode": (_I, [_VP, _I]),
    "udx_ens_set_pp": (_I, [_VP, _VP]),
    "udx_ens_get_pp": (_I, [_VP, _VP]),
    "udx_ens_run": (_I, [_VP, _I, C.POINTER(_F)]),
    "udx_ens_set_graph": (_I, [_VP, _I]),
    "udx_ens_autotune": (_I, [_VP, _I, C.POINTER(_I), C.POINTER(_F)]),
    "udx_dpd_set_split_edges": (_I, [C.c_double]),
    "udx_ens_set_variant": (_I, [_VP, _I]),
    "udx_ens_pairs": (_I, [_VP, C.POINTER(C.c_ulonglong), _I]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libudx.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UdxError("libudx.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ok(status):
    if status != 0:
        raise UdxError(lib().udx_last_error().decode())


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------- host RNG
class RndTrunk:
    """KISS trunk (src/math/rnd/imp.cpp): rnd_ini / rnd_get."""

    def __init__(self, x, y, z, c):
        self.s = Rnd()
        lib().udx_rnd_ini(x, y, z, c, C.byref(self.s))

    def get(self):
        return lib().udx_rnd_get(C.byref(self.s))


MODE_FAST, MODE_EXACT, MODE_HALF = 0, 1, 2


def _mode(exact, mode):
    """Force kernel mode: an explicit MODE_* wins, else exact selects
    MODE_EXACT over MODE_FAST."""
    if mode is not None:
        return int(mode)
    return MODE_EXACT if exact else MODE_FAST


def dpd_params(a, g, kBT, dt, spow=0.25):
    """pair_set_dpd + pair_compute_dpd_sigma (src/pair/imp/main.h:54-62)."""
    p = PairDpd(a, g, 0.0, spow)
    lib().udx_pair_compute_dpd_sigma(kBT, dt, C.byref(p))
    return p


def rnd_branch(ctx, seed, u, v, out, uu=False):
    """rnd::mean0var1ii/uu on the GPU for int32 device tensors u, v."""
    _ok(lib().udx_rnd_branch(ctx.h, seed, u.shape[0], _ptr(u), _ptr(v), 1 if uu else 0, _ptr(out)))


CHECK_RCP, CHECK_SQRT, CHECK_FRACT, CHECK_RCP_ESTIMATE = 0, 1, 2, 3


def selfcheck(ctx, what, lo, hi):
    """floats x in [lo, hi] where the device routine `what` (CHECK_RCP,
    CHECK_SQRT, CHECK_FRACT, CHECK_RCP_ESTIMATE) differs from the correctly
    rounded result"""
    v = C.c_ulonglong()
    _ok(lib().udx_selfcheck(ctx.h, what, lo, hi, C.byref(v)))
    return v.value


# ---------------------------------------------------------------- context
class Context:
    def __init__(self, stream=None):
        self.h = C.c_void_p()
        _ok(lib().udx_ctx_ini(stream if stream is not None else _stream(), C.byref(self.h)))

    def sync(self):
        _ok(lib().udx_ctx_sync(self.h))

    def close(self):
        if self.h:
            lib().udx_ctx_fin(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- clist
class Clist:
    """clist_ini / clist_build / clists_get_ss / clists_get_cc (src/clist/imp.h)."""

    def __init__(self, ctx, L, maxp):
        self.ctx, self.L, self.maxp = ctx, tuple(L), maxp
        self.h = C.c_void_p()
        _ok(lib().udx_clist_ini(ctx.h, L[0], L[1], L[2], maxp, C.byref(self.h)))

    @property
    def ncells(self):
        return self.L[0] * self.L[1] * self.L[2]

    def build(self, pp, ppout):
        n = C.c_int()
        _ok(lib().udx_clist_build(self.h, pp.shape[0], _ptr(pp), _ptr(ppout), C.byref(n)))
        return n.value

    def _view(self, fn, count):
        import torch
        out = torch.empty(count, dtype=torch.int32, device="cuda")
        _copy_dev(out, fn(self.h), count * 4)
        return out

    def starts(self):
        return self._view(lib().udx_clist_starts, self.ncells + 1)

    def counts(self):
        return self._view(lib().udx_clist_counts, self.ncells)

    def order(self, n):
        return self._view(lib().udx_clist_order, n)

    def starts_ptr(self):
        return C.c_void_p(lib().udx_clist_starts(self.h))

    def close(self):
        if self.h:
            lib().udx_clist_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- scan
class Scan:
    """scan_ini / scan_apply / scan_fin (src/algo/scan/imp.h:4-6): exclusive
    scan of device int32 tensors of at most `size` entries."""

    def __init__(self, ctx, size):
        self.ctx, self.size = ctx, size
        self.h = C.c_void_p()
        _ok(lib().udx_scan_ini(ctx.h, size, C.byref(self.h)))

    def apply(self, inp, out, size=None):
        n = inp.numel() if size is None else size
        _ok(lib().udx_scan_apply(self.h, _ptr(inp), n, _ptr(out)))

    def close(self):
        if self.h:
            lib().udx_scan_fin(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# static working space of udx_clist_build_noalloc (include/udx.h)
NOALLOC_MAXCELLS, NOALLOC_MAXP = 1 << 18, 1 << 20


def clist_build_noalloc(ctx, L, pp, ppout, starts, counts, order):
    """udx_clist_build_noalloc: the cell list with caller-owned outputs and no
    device allocation (torch device tensors; ppout may be None)."""
    _ok(lib().udx_clist_build_noalloc(ctx.h, int(L[0]), int(L[1]), int(L[2]), int(pp.shape[0]), _ptr(pp),
                                      _ptr(ppout), _ptr(starts), _ptr(counts), _ptr(order)))


def _copy_dev(dst, src_ptr, nbytes):
    """Device-to-device copy from a raw pointer owned by libudx."""
    import torch
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipMemcpyAsync.restype = C.c_int
    st = hip.hipMemcpyAsync(C.c_void_p(dst.data_ptr()), C.c_void_p(src_ptr), nbytes, 3, _stream())
    if st != 0:
        raise UdxError("hipMemcpyAsync failed: %d" % st)
    torch.cuda.current_stream().synchronize()


# ---------------------------------------------------------------- fluforces
class FluForces:
    """fluforces_bulk_* + fluforces_halo_* (src/fluforces/imp.h) on one
    periodic domain."""

    def __init__(self, ctx, L, maxp, exact=False, mode=None):
        self.ctx, self.L = ctx, tuple(L)
        self.h = C.c_void_p()
        _ok(lib().udx_fluforces_ini(ctx.h, L[0], L[1], L[2], maxp, C.byref(self.h)))
        _ok(lib().udx_fluforces_set_mode(self.h, _mode(exact, mode)))

    def prepare(self, pp):
        _ok(lib().udx_fluforces_prepare(self.h, pp.shape[0], _ptr(pp)))

    def apply(self, params, starts_ptr, ff, periodic=True, seeds=None):
        if seeds is None:
            _ok(lib().udx_fluforces_apply(self.h, C.byref(params), starts_ptr, 1 if periodic else 0, _ptr(ff)))
        else:
            arr = (C.c_float * 27)(*[float(x) for x in seeds])
            _ok(lib().udx_fluforces_apply_seeds(self.h, C.byref(params), starts_ptr, 1 if periodic else 0,
                                                arr, _ptr(ff)))

    def apply_stress(self, params, starts_ptr, ff, ss, periodic=True, seeds=None):
        """forces and per-particle stresses (ss: torch (n, 6) float32), added"""
        arr = None if seeds is None else (C.c_float * 27)(*[float(x) for x in seeds])
        _ok(lib().udx_fluforces_apply_stress(self.h, C.byref(params), starts_ptr, 1 if periodic else 0, arr,
                                             _ptr(ff), _ptr(ss)))

    def pairs(self, reset=False):
        v = C.c_ulonglong()
        _ok(lib().udx_fluforces_pairs(self.h, C.byref(v), 1 if reset else 0))
        return v.value

    def close(self):
        if self.h:
            lib().udx_fluforces_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- scheme / diag
def scheme_move_apply(ctx, dt, mass, ff, pp):
    _ok(lib().udx_scheme_move_apply(ctx.h, dt, mass, pp.shape[0], _ptr(ff), _ptr(pp)))


def bforce_apply_cste(ctx, mass, accel, ff):
    a = (C.c_float * 3)(*accel)
    _ok(lib().udx_bforce_apply_cste(ctx.h, mass, a, ff.shape[0], _ptr(ff)))


def wrap_pp(ctx, L, pp):
    _ok(lib().udx_wrap_pp(ctx.h, L[0], L[1], L[2], pp.shape[0], _ptr(pp)))


def diag_part(ctx, pp):
    out = (C.c_double * 5)()
    _ok(lib().udx_diag_part(ctx.h, pp.shape[0], _ptr(pp), out))
    return list(out)


# ---------------------------------------------------------------- solver
class Flu:
    """Solvent-only step pipeline (src/sim/imp/run.h:8-68): clist + reorder,
    DPD forces, body force, kick-drift, periodic wrap."""

    def __init__(self, ctx, L, maxp, params, dt, mass=1.0, accel=(0.0, 0.0, 0.0), exact=False, mode=None):
        self.ctx = ctx
        conf = FluConf()
        for k in range(3):
            conf.L[k] = L[k]
            conf.accel[k] = accel[k]
        conf.maxp, conf.dpd, conf.dt, conf.mass = maxp, params, dt, mass
        conf.mode = _mode(exact, mode)
        self.h = C.c_void_p()
        _ok(lib().udx_flu_ini(ctx.h, C.byref(conf), C.byref(self.h)))

    def set_pp(self, pp):
        _ok(lib().udx_flu_set_pp(self.h, pp.shape[0], _ptr(pp)))

    def get_pp(self, out):
        n = C.c_int()
        _ok(lib().udx_flu_get_pp(self.h, _ptr(out), C.byref(n)))
        return n.value

    def get_ff(self, out):
        _ok(lib().udx_flu_get_ff(self.h, _ptr(out)))

    def run(self, nsteps, seeds=None):
        """seeds: None (internal trunks) or an (nsteps, 27|28) table; a
        27-column table gets the wall trunk (42,42,42,42) as column 27."""
        if seeds is None:
            _ok(lib().udx_flu_run(self.h, nsteps, None))
        else:
            import numpy as np
            s = np.ascontiguousarray(seeds, dtype=np.float32)
            if s.shape == (nsteps, 27):
                if not hasattr(self, "_wall_trunk"):
                    self._wall_trunk = RndTrunk(42, 42, 42, 42)
                w = np.array([self._wall_trunk.get() for _ in range(nsteps)], np.float32)
                s = np.ascontiguousarray(np.concatenate([s, w[:, None]], axis=1))
            assert s.shape == (nsteps, 28)
            _ok(lib().udx_flu_run(self.h, nsteps, s.ctypes.data_as(C.POINTER(C.c_float))))

    def set_wall(self, wall):
        self.wall = wall
        _ok(lib().udx_flu_set_wall(self.h, wall.h if wall is not None else None))

    def graph(self, enable=True):
        """replay captured step pairs (udx_flu_set_graph)"""
        _ok(lib().udx_flu_set_graph(self.h, 1 if enable else 0))

    def set_time(self, t):
        """the step clock (time of the next step), advanced by dt per step"""
        _ok(lib().udx_flu_set_time(self.h, float(t)))

    def set_rbc_repulsion(self, lam):
        """wall repulsion of the membrane vertices (0: none)"""
        _ok(lib().udx_flu_set_rbc_repulsion(self.h, float(lam)))

    def set_rbc_contact(self, cnt, params=None):
        """membrane contact in the step (udx_flu_set_rbc_contact); None: off"""
        self._cnt = cnt
        _ok(lib().udx_flu_set_rbc_contact(self.h, cnt.h if cnt is not None else None,
                                          C.byref(params) if params is not None else None))

    def set_rbc_bounce(self, mbb):
        """membrane bounce-back in the step (udx_flu_set_rbc_bounce); None: off"""
        self._mbb = mbb
        _ok(lib().udx_flu_set_rbc_bounce(self.h, mbb.h if mbb is not None else None))

    def strt_dump(self, base, id):
        _ok(lib().udx_flu_strt_dump(self.h, str(base).encode(), id))

    def strt_quants(self, base, id):
        _ok(lib().udx_flu_strt_quants(self.h, str(base).encode(), id))

    def set_rbc(self, rbc, fsi=None, rbc_params=None, fsi_params=None, mass=0.5, substeps=1, nc=0, pp=None):
        """membranes in the step (udx_flu_set_rbc); pp: torch (nc * nv, 6)
        float32 on the device, advanced in place; rbc=None detaches"""
        if rbc is None:
            _ok(lib().udx_flu_set_rbc(self.h, None, None, None, None, 0.0, 0, 0, None))
            self._rbc = None
            return
        self._rbc = (rbc, fsi, pp)
        _ok(lib().udx_flu_set_rbc(self.h, rbc.h, fsi.h, C.byref(rbc_params), C.byref(fsi_params), float(mass),
                                  int(substeps), int(nc), _ptr(pp) if pp is not None else None))

    def wall_pairs(self):
        v = C.c_ulonglong()
        _ok(lib().udx_flu_wall_pairs(self.h, C.byref(v)))
        return v.value

    def pairs(self, reset=False):
        v = C.c_ulonglong()
        _ok(lib().udx_flu_pairs(self.h, C.byref(v), 1 if reset else 0))
        return v.value

    def timing(self, enable=True):
        out = (C.c_float * 5)()
        _ok(lib().udx_flu_timing(self.h, 1 if enable else 0, out))
        return list(out)

    def close(self):
        if self.h:
            lib().udx_flu_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- walls
class Sdf:
    """sdf_ini / sdf_gen (src/wall/sdf/imp.h): signed distance of the walls."""

    def __init__(self, ctx, L, ext=None, N=None, D=None, path=None):
        self.ctx, self.L = ctx, tuple(L)
        self.h = C.c_void_p()
        _ok(lib().udx_sdf_ini(ctx.h, L[0], L[1], L[2], C.byref(self.h)))
        if path is not None:
            _ok(lib().udx_sdf_gen_file(self.h, os.fsencode(path)))
        elif D is not None:
            import numpy as np
            e = (C.c_float * 3)(*[float(x) for x in ext])
            n = (C.c_int * 3)(*[int(x) for x in N])
            d = np.ascontiguousarray(D, np.float32)
            _ok(lib().udx_sdf_gen(self.h, e, n, d.ctypes.data_as(C.POINTER(C.c_float))))

    def size(self):
        T = (C.c_int * 3)()
        thr = C.c_float()
        _ok(lib().udx_sdf_size(self.h, T, C.byref(thr)))
        return tuple(T), thr.value

    def probe(self, r3, what=0):
        """what 0: sdf_eval, 1: nearest texel, 2: gradient."""
        import torch
        n = r3.shape[0]
        out = torch.empty((n, 3) if what == 2 else (n,), dtype=torch.float32, device="cuda")
        _ok(lib().udx_sdf_probe(self.h, n, _ptr(r3), what, _ptr(out)))
        return out

    def bounce(self, dt, pp):
        _ok(lib().udx_wall_bounce(self.h, dt, pp.shape[0], _ptr(pp)))

    def repulse(self, lam, pp, ff):
        _ok(lib().udx_wall_repulse(self.h, lam, pp.shape[0], _ptr(pp), _ptr(ff)))

    def close(self):
        if self.h:
            lib().udx_sdf_fin(self.h)
            self.h = C.c_void_p()


class Wall:
    """wall_gen + wall_force (src/wall/imp.h): frozen wall particles."""

    def __init__(self, ctx, sdf, pp, exact=False, mode=None, strt=None, maxn=0):
        """pp: solvent (n, 6) cuda tensor; on return the first self.n_solvent
        rows hold the solvent that stays (order kept).  With strt=base the
        frozen particles come from the restart template instead (pp unused)."""
        self.ctx, self.sdf = ctx, sdf
        self.h = C.c_void_p()
        if strt is not None:
            _ok(lib().udx_wall_strt_gen(ctx.h, sdf.h, str(strt).encode(), maxn, C.byref(self.h)))
            self.n_solvent = None
        else:
            n = C.c_int(pp.shape[0])
            _ok(lib().udx_wall_gen(ctx.h, sdf.h, C.byref(n), _ptr(pp), C.byref(self.h)))
            self.n_solvent = n.value
        _ok(lib().udx_wall_set_mode(self.h, _mode(exact, mode)))

    def strt_dump_templ(self, base):
        _ok(lib().udx_wall_strt_dump_templ(self.h, str(base).encode()))

    def counts(self):
        a, b = C.c_int(), C.c_int()
        _ok(lib().udx_wall_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def particles(self):
        import torch
        _, n = self.counts()
        out = torch.empty((n, 6), dtype=torch.float32, device="cuda")
        _copy_dev(out, lib().udx_wall_pp(self.h), n * 24)
        return out

    def starts(self):
        import torch
        W = [x + 12 for x in self.sdf.L]
        nc = W[0] * W[1] * W[2]
        out = torch.empty(nc + 1, dtype=torch.int32, device="cuda")
        _copy_dev(out, lib().udx_wall_starts(self.h), (nc + 1) * 4)
        return out

    def force(self, params, pp, ff, seed=None):
        if seed is None:
            _ok(lib().udx_wall_force(self.h, C.byref(params), pp.shape[0], _ptr(pp), _ptr(ff)))
        else:
            _ok(lib().udx_wall_force_seed(self.h, C.byref(params), pp.shape[0], _ptr(pp), float(seed), _ptr(ff)))

    def pairs(self, reset=False):
        v = C.c_ulonglong()
        _ok(lib().udx_wall_pairs(self.h, C.byref(v), 1 if reset else 0))
        return v.value

    def set_vel(self, v):
        """v: a WallVel (wall_vel(...)); the walls rest until this is called"""
        _ok(lib().udx_wall_set_vel(self.h, C.byref(v)))

    def set_time(self, t):
        _ok(lib().udx_wall_set_time(self.h, float(t)))

    def bounce(self, dt, pp):
        """bounce-back off this wall's surface with its velocity"""
        _ok(lib().udx_wall_bounce_wall(self.h, dt, pp.shape[0], _ptr(pp)))

    def close(self):
        if self.h:
            lib().udx_wall_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- membrane contact
class Cnt:
    """cnt_* (src/cnt/imp.h): contact between membrane vertices"""

    def __init__(self, ctx, L, max_n):
        self.ctx = ctx
        self.h = C.c_void_p()
        _ok(lib().udx_cnt_ini(ctx.h, L[0], L[1], L[2], max_n, C.byref(self.h)))

    def apply(self, params, pp, ff, seeds=None):
        """pp: (n, 6) cuda tensor; ff: (n, 3) cuda tensor, added to; seeds:
        (bulk, halo) or None to draw them from the handle's trunk"""
        if seeds is None:
            _ok(lib().udx_cnt_apply(self.h, C.byref(params), pp.shape[0], _ptr(pp), _ptr(ff)))
        else:
            _ok(lib().udx_cnt_apply_seed(self.h, C.byref(params), pp.shape[0], _ptr(pp), float(seeds[0]),
                                         float(seeds[1]), _ptr(ff)))

    def pairs(self, reset=False):
        v = C.c_ulonglong()
        _ok(lib().udx_cnt_pairs(self.h, C.byref(v), 1 if reset else 0))
        return v.value

    def close(self):
        if self.h:
            lib().udx_cnt_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- membrane bounce-back
class Mbb:
    """mesh_bounce_* (src/mesh_bounce/imp.h): solvent bounce-back on moving
    membrane meshes of one template"""

    def __init__(self, ctx, nv, tt, max_meshes, maxp):
        import numpy as np
        self.ctx, self.nv = ctx, nv
        tt = np.ascontiguousarray(tt, np.int32)
        self.nt = len(tt)
        self.h = C.c_void_p()
        _ok(lib().udx_mbb_ini(ctx.h, nv, self.nt, tt.ctypes.data_as(C.POINTER(C.c_int)), max_meshes, maxp,
                              C.byref(self.h)))

    def bounce(self, dt, flu_mass, rbc_mass, L, nm, rp, rc, starts, pp0, pp, mm=None):
        """rp, rc: (nm nv, 6) cuda tensors (rc's velocities updated); starts:
        int32 cell starts of pp0; pp0, pp: (n, 6) cuda tensors (pp bounced)"""
        _ok(lib().udx_mbb_bounce(self.h, dt, flu_mass, rbc_mass, L[0], L[1], L[2], nm, _ptr(rp), _ptr(rc),
                                 _ptr(starts), pp.shape[0], _ptr(pp0), _ptr(pp),
                                 _ptr(mm) if mm is not None else None))

    def count(self):
        v = C.c_ulonglong()
        _ok(lib().udx_mbb_count(self.h, C.byref(v)))
        return v.value

    def set_pair_capacity(self, cap):
        _ok(lib().udx_mbb_set_pair_capacity(self.h, cap))

    def collisions(self, n):
        import torch
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        _copy_dev(out, lib().udx_mbb_collisions(self.h), n * 4)
        return out

    def close(self):
        if self.h:
            lib().udx_mbb_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- restart
RESTART_TEMPL, RESTART_FINAL = -1, -2


def restart_write_pp(base, code, id, pp):
    import numpy as np
    pp = np.ascontiguousarray(pp, np.float32).reshape(-1, 6)
    _ok(lib().udx_restart_write_pp(str(base).encode(), code.encode(), id, len(pp), pp.ctypes.data))


def restart_read_pp(base, code, id, maxn):
    import numpy as np
    out = np.zeros((max(maxn, 1), 6), np.float32)
    n = C.c_int()
    _ok(lib().udx_restart_read_pp(str(base).encode(), code.encode(), id, maxn, C.byref(n), out.ctypes.data))
    return out[:n.value].copy()


def restart_write_ii(base, code, id, ii):
    import numpy as np
    ii = np.ascontiguousarray(ii, np.int32).ravel()
    _ok(lib().udx_restart_write_ii(str(base).encode(), code.encode(), id, len(ii), ii.ctypes.data))


def restart_read_ii(base, code, id, maxn):
    import numpy as np
    out = np.zeros(max(maxn, 1), np.int32)
    n = C.c_int()
    _ok(lib().udx_restart_read_ii(str(base).encode(), code.encode(), id, maxn, C.byref(n), out.ctypes.data))
    return out[:n.value].copy()


# ---------------------------------------------------------------- RBC
def read_off(path):
    """OFF mesh (src/io/mesh_read/imp/off.h): vertices (nv, 3) float32 and
    triangles (nt, 3) int32."""
    import numpy as np
    with open(path) as f:
        toks = f.read().split()
    if not toks or toks[0] != "OFF":
        raise UdxError("%s: not an OFF file" % path)
    nv, nt = int(toks[1]), int(toks[2])
    k = 4
    rr = np.array(toks[k:k + 3 * nv], np.float32).reshape(nv, 3)
    k += 3 * nv
    tt = np.zeros((nt, 3), np.int32)
    for t in range(nt):
        if toks[k] != "3":
            raise UdxError("%s: face %d is not a triangle" % (path, t))
        tt[t] = [int(toks[k + 1]), int(toks[k + 2]), int(toks[k + 3])]
        k += 4
    return rr, tt


def rbc_params(**kw):
    p = RbcParams()
    for k, v in kw.items():
        setattr(p, k, float(v))
    return p


class Rbc:
    """Membrane forces of nc copies of one mesh (src/rbc/force/imp.h:9-28)."""

    def __init__(self, ctx, tt, nv, max_cells):
        import numpy as np
        self.ctx = ctx
        t = np.ascontiguousarray(tt, np.int32).reshape(-1, 3)
        self.nv, self.nt = int(nv), len(t)
        self.h = C.c_void_p()
        _ok(lib().udx_rbc_ini(ctx.h, self.nv, self.nt, t.ctypes.data_as(C.POINTER(_I)), max_cells,
                              C.byref(self.h)))
        md = C.c_int()
        _ok(lib().udx_rbc_md(self.h, C.byref(md)))
        self.md = md.value

    def adjacency(self):
        import numpy as np
        a0 = np.zeros(self.md * self.nv, np.int32)
        a1 = np.zeros(self.md * self.nv, np.int32)
        _ok(lib().udx_rbc_adjacency(self.h, a0.ctypes.data_as(C.POINTER(_I)), a1.ctypes.data_as(C.POINTER(_I))))
        return a0, a1

    def set_stressful(self, totArea):
        _ok(lib().udx_rbc_set_stressful(self.h, float(totArea)))

    def set_stressfree(self, rr):
        import numpy as np
        r = np.ascontiguousarray(rr, np.float32)
        _ok(lib().udx_rbc_set_stressfree(self.h, r.ctypes.data_as(C.POINTER(_F))))

    def set_bending(self, kantor=True):
        _ok(lib().udx_rbc_set_bending(self.h, 1 if kantor else 0))

    def set_bending_type(self, kind):
        """rbc_force_set_bending: one of none, kantor, juelicher"""
        _ok(lib().udx_rbc_set_bending_type(self.h, {"none": 0, "kantor": 1, "juelicher": 2}[kind]))

    def area_volume(self, nc, pp):
        import torch
        av = torch.empty(2 * nc, dtype=torch.float64, device="cuda")
        _ok(lib().udx_rbc_area_volume(self.h, nc, _ptr(pp), _ptr(av)))
        return av

    def force(self, params, dt, nc, pp, ff):
        _ok(lib().udx_rbc_force_apply(self.h, C.byref(params), float(dt), nc, _ptr(pp), _ptr(ff)))

    def distribute(self, L, nc, pp):
        """whole-cell periodic redistribution of pp (device, in place)"""
        _ok(lib().udx_rbc_distribute(self.h, L[0], L[1], L[2], nc, _ptr(pp)))

    def adjoint(self, params, nc, pp, g, gx):
        """gx = (dF/dx)^T g; returns [dL/dka, dL/dkd, dL/dkv, dL/dks, dL/dkb]"""
        gp = (C.c_double * 5)()
        _ok(lib().udx_rbc_force_adjoint(self.h, C.byref(params), nc, _ptr(pp), _ptr(g), _ptr(gx), gp))
        return list(gp)

    def close(self):
        if self.h:
            lib().udx_rbc_fin(self.h)
            self.h = C.c_void_p()


def rbc_stretch_apply(ctx, nc, nv, f, ff):
    _ok(lib().udx_rbc_stretch_apply(ctx.h, nc, nv, _ptr(f), _ptr(ff)))


# ---------------------------------------------------------------- FSI
class Fsi:
    """Solvent-membrane DPD (src/fsi/imp.h:6-13)."""

    def __init__(self, ctx, L, max_vertices, exact=False, mode=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        _ok(lib().udx_fsi_ini(ctx.h, L[0], L[1], L[2], max_vertices, C.byref(self.h)))
        _ok(lib().udx_fsi_set_mode(self.h, _mode(exact, mode)))

    def bind_solvent(self, pp, ff, starts_ptr):
        self._bound = (pp, ff)
        _ok(lib().udx_fsi_bind_solvent(self.h, _ptr(pp), _ptr(ff), pp.shape[0], starts_ptr))

    def apply(self, params, opp, off, seeds=None):
        if seeds is None:
            _ok(lib().udx_fsi_apply(self.h, C.byref(params), opp.shape[0], _ptr(opp), _ptr(off)))
        else:
            s = (C.c_float * 2)(*[float(x) for x in seeds])
            _ok(lib().udx_fsi_apply_seeds(self.h, C.byref(params), opp.shape[0], _ptr(opp), _ptr(off), s))

    def pairs(self, reset=False):
        v = C.c_ulonglong()
        _ok(lib().udx_fsi_pairs(self.h, C.byref(v), 1 if reset else 0))
        return v.value

    def close(self):
        if self.h:
            lib().udx_fsi_fin(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- configuration
class Conf:
    """Layered libconfig-subset configuration (src/conf/imp.h)."""

    def __init__(self, default=None, opt=None, args=None, include_dir=None):
        self.h = C.c_void_p()
        _ok(lib().udx_conf_ini(C.byref(self.h)))
        if include_dir is not None:
            _ok(lib().udx_conf_set_include_dir(self.h, os.fsencode(include_dir)))
        enc = (lambda x: None if x is None else (os.fsencode(x) if not isinstance(x, bytes) else x))
        if default is not None or opt is not None or args is not None:
            _ok(lib().udx_conf_read(self.h, enc(default), enc(opt), None if args is None else args.encode()))

    def has(self, key):
        return bool(lib().udx_conf_has(self.h, key.encode()))

    def int(self, key):
        v = C.c_int()
        _ok(lib().udx_conf_lookup_int(self.h, key.encode(), C.byref(v)))
        return v.value

    def float(self, key):
        v = C.c_float()
        _ok(lib().udx_conf_lookup_float(self.h, key.encode(), C.byref(v)))
        return v.value

    def bool(self, key):
        v = C.c_int()
        _ok(lib().udx_conf_lookup_bool(self.h, key.encode(), C.byref(v)))
        return bool(v.value)

    def string(self, key):
        v = C.c_char_p()
        _ok(lib().udx_conf_lookup_string(self.h, key.encode(), C.byref(v)))
        return v.value.decode()

    def vint(self, key, maxn=64):
        a, n = (C.c_int * maxn)(), C.c_int()
        _ok(lib().udx_conf_lookup_vint(self.h, key.encode(), maxn, C.byref(n), a))
        return list(a[:n.value])

    def vfloat(self, key, maxn=64):
        a, n = (C.c_float * maxn)(), C.c_int()
        _ok(lib().udx_conf_lookup_vfloat(self.h, key.encode(), maxn, C.byref(n), a))
        return list(a[:n.value])

    def vstring(self, key, maxn=64):
        a, n = (C.c_char_p * maxn)(), C.c_int()
        _ok(lib().udx_conf_lookup_vstring(self.h, key.encode(), maxn, C.byref(n), a))
        return [x.decode() for x in a[:n.value]]

    def int3(self, key):
        a = (C.c_int * 3)()
        _ok(lib().udx_conf_lookup_int3(self.h, key.encode(), a))
        return tuple(a)

    def float3(self, key):
        a = (C.c_float * 3)()
        _ok(lib().udx_conf_lookup_float3(self.h, key.encode(), a))
        return tuple(a)

    def set(self, key, value):
        k = key.encode()
        if isinstance(value, bool):
            _ok(lib().udx_conf_set_bool(self.h, k, int(value)))
        elif isinstance(value, int):
            _ok(lib().udx_conf_set_int(self.h, k, value))
        elif isinstance(value, float):
            _ok(lib().udx_conf_set_float(self.h, k, value))
        else:
            _ok(lib().udx_conf_set_string(self.h, k, str(value).encode()))

    def pair(self, ns, kBT, dt):
        p = PairDpd()
        _ok(lib().udx_pair_set_conf(self.h, ns.encode(), kBT, dt, C.byref(p)))
        return p

    def rbc_params(self, ns="rbc"):
        p = RbcParams()
        _ok(lib().udx_rbc_params_set_conf(self.h, ns.encode(), C.byref(p)))
        return p

    def wall_vel(self):
        v = WallVel()
        _ok(lib().udx_wall_vel_set_conf(self.h, C.byref(v)))
        return v

    def wall_repulse(self, ns):
        """the wall repulsion coefficient of object type ns (0 for "none")"""
        lam = _F()
        _ok(lib().udx_wall_repulse_set_conf(self.h, ns.encode(), C.byref(lam)))
        return lam.value

    def close(self):
        if self.h:
            lib().udx_conf_fin(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- ensemble
class Ensemble:
    """Independent periodic boxes advanced together (one launch per kernel
    per step for all boxes)."""

    def __init__(self, ctx, boxes, exact=False, mode=None):
        """boxes: list of dicts with L, n, params (PairDpd), dt, mass, accel"""
        self.ctx = ctx
        arr = (EnsBox * len(boxes))()
        for k, b in enumerate(boxes):
            for c in range(3):
                arr[k].L[c] = b["L"][c]
                arr[k].accel[c] = b.get("accel", (0.0, 0.0, 0.0))[c]
            arr[k].n = b["n"]
            arr[k].dpd = b["params"]
            arr[k].dt = b["dt"]
            arr[k].mass = b.get("mass", 1.0)
        self.nbox = len(boxes)
        self.n = sum(b["n"] for b in boxes)
        self.h = C.c_void_p()
        _ok(lib().udx_ens_ini(ctx.h, self.nbox, arr, C.byref(self.h)))
        _ok(lib().udx_ens_set_mode(self.h, _mode(exact, mode)))

    def set_pp(self, pp):
        _ok(lib().udx_ens_set_pp(self.h, _ptr(pp)))

    def get_pp(self, out):
        _ok(lib().udx_ens_get_pp(self.h, _ptr(out)))

    def run(self, nsteps, seeds=None):
        if seeds is None:
            _ok(lib().udx_ens_run(self.h, nsteps, None))
        else:
            import numpy as np
            s = np.ascontiguousarray(seeds, np.float32)
            assert s.shape == (nsteps, self.nbox, 27)
            _ok(lib().udx_ens_run(self.h, nsteps, s.ctypes.data_as(C.POINTER(C.c_float))))

    def graph(self, enable=True):
        _ok(lib().udx_ens_set_graph(self.h, 1 if enable else 0))

    def autotune(self, steps=3):
        best = C.c_int()
        ms = (C.c_float * 16)()
        _ok(lib().udx_ens_autotune(self.h, steps, C.byref(best), ms))
        return best.value, list(ms[:6])

    def set_variant(self, v):
        _ok(lib().udx_ens_set_variant(self.h, v))

    def pairs(self, reset=False):
        v = C.c_ulonglong()
        _ok(lib().udx_ens_pairs(self.h, C.byref(v), 1 if reset else 0))
        return v.value

    def close(self):
        if self.h:
            lib().udx_ens_fin(self.h)
            self.h = C.c_void_p()


def ens_trunk_seeds(nbox, nsteps):
    """the boxes' default seed stream (nsteps, nbox, 27): box b's trunks are
    the solvent trunks with the first KISS word offset by 7919 b"""
    import numpy as np
    tr = []
    for b in range(nbox):
        row = []
        for f in range(27):
            if f == 26:
                x, y, z, c = 0, 0, 0, 0
            else:
                d = ((f + 2) % 3 - 1, (f // 3 + 2) % 3 - 1, (f // 9 + 2) % 3 - 1)
                s = sum(d)
                plus = (s > 0) or (s == 0 and d[0] > 0) or (d[0] == 0 and d[1] > 0) or (d[1] == 0 and d[2] > 0)
                sign = 2 * int(plus) - 1
                v = [1 + sign * dc for dc in d]
                seed = v[0] + 3 * (v[1] + 3 * v[2])
                x, y, z, c = 390 + seed, seed + 615, 12309, 23094
            row.append(RndTrunk((x + 7919 * b) & 0xFFFFFFFF, y, z, c))
        tr.append(row)
    out = np.zeros((nsteps, nbox, 27), np.float32)
    for k in range(nsteps):
        for b in range(nbox):
            for f in range(27):
                out[k, b, f] = tr[b][f].get()
    return out


def set_split_edges(n):
    """udx_dpd_set_split_edges: the expected boundary particles from which the
    gather's halo pass is a launch of its own (same forces either way)."""
    _ok(lib().udx_dpd_set_split_edges(float(n)))
