"""ctypes bindings of the RBC membrane part of the C oracle
(oracle/udx_oracle_rbc.c).  Test infrastructure only."""
import ctypes as C

import numpy as np

from .oracle import Dpd, F32P, F64P, I32P, _f32, _i32, _ptr, lib


class RbcPar(C.Structure):
    """src/rbc/params/imp.h:3-9"""
    _fields_ = [(k, C.c_float) for k in ("gammaC", "gammaT", "kBT", "kb", "phi", "ks", "x0", "mpow", "ka", "kd",
                                          "kv", "totArea", "totVolume")]


class DpdLj(C.Structure):
    """DPD + Lennard-Jones pair (src/pair/type.h:22-25)"""
    _fields_ = [(k, C.c_float) for k in ("a", "g", "s", "spow", "ljs", "lje")]


# conf/rbc.test.cfg
RBC_TEST = dict(gammaC=30.0, gammaT=90.0, kBT=0.0444, kb=17.78, phi=6.97, ks=14.8, x0=0.5, mpow=2.0, ka=2450.0,
                kd=100.0, kv=3750.0)

_declared = False


def _lib():
    global _declared
    L = lib()
    if not _declared:
        P = C.POINTER(RbcPar)
        sig = {
            "orc_mesh_md": (C.c_int, [C.c_int, C.c_int, I32P]),
            "orc_adj_ini": (C.c_int, [C.c_int, C.c_int, C.c_int, I32P, I32P, I32P]),
            "orc_adj_map": (C.c_int, [C.c_int, C.c_int, I32P, I32P, C.c_int, I32P]),
            "orc_rbc_area_volume": (None, [C.c_int, C.c_int, C.c_int, I32P, F32P, F64P]),
            "orc_rbc_shape": (None, [C.c_int, C.c_int, I32P, I32P, F32P, F32P, F32P]),
            "orc_rbc_stressful": (None, [C.c_int, C.c_float, F32P, F32P]),
            "orc_rbc_forces": (None, [P, C.c_int, C.c_int, C.c_int, I32P, I32P, F32P, F64P, F32P, F32P, C.c_float,
                                      C.c_float, F64P]),
            "orc_rbc_bending": (None, [P, C.c_int, C.c_int, C.c_int, I32P, I32P, F32P, F64P]),
            "orc_rbc_energy": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int, I32P, I32P, I32P, F64P, F32P,
                                            F32P, C.c_float, C.c_float, C.c_int]),
            "orc_rbc_stretch": (None, [C.c_int, C.c_int, F32P, F32P]),
            "orc_fsi_forces": (C.c_long, [I32P, C.c_int, F32P, I32P, C.c_int, F32P, Dpd, F32P, F32P, F32P]),
            "orc_fsi_seeds": (None, [C.c_int, F32P]),
            "orc_rbc_adjoint": (None, [P, C.c_int, C.c_int, C.c_int, C.c_int, I32P, I32P, I32P, F64P, F32P, F32P,
                                       C.c_float, C.c_float, F64P, F64P, F64P, F64P]),
            "orc_jul_edges": (C.c_int, [C.c_int, C.c_int, I32P, I32P]),
            "orc_cnt_pair": (C.c_int, [C.POINTER(DpdLj), F32P, F32P, C.c_float, F32P]),
            "orc_cnt_images": (C.c_int, [I32P, C.c_int, F32P, F32P, I32P]),
            "orc_cnt_forces": (C.c_long, [I32P, C.c_int, F32P, DpdLj, C.c_float, C.c_float, F32P]),
            "orc_jul_theta": (C.c_double, [F64P, F64P, F64P, F64P]),
            "orc_jul_forces": (None, [C.c_int, C.c_int, C.c_int, I32P, C.c_int, I32P, F32P, F32P, F32P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _declared = True
    return L


def read_off(path):
    """OFF mesh: (vertices (nv, 3) float32, triangles (nt, 3) int32)
    (src/io/mesh_read/imp/off.h)."""
    with open(path) as f:
        toks = f.read().split()
    assert toks[0] == "OFF"
    nv, nt = int(toks[1]), int(toks[2])
    k = 4
    rr = np.array(toks[k:k + 3 * nv], np.float32).reshape(nv, 3)
    k += 3 * nv
    tt = np.zeros((nt, 3), np.int32)
    for t in range(nt):
        assert toks[k] == "3"
        tt[t] = [int(toks[k + 1]), int(toks[k + 2]), int(toks[k + 3])]
        k += 4
    return rr, tt


class Mesh:
    """Template mesh with its ring adjacency."""

    def __init__(self, rr, tt):
        self.rr = _f32(rr).reshape(-1, 3)
        self.tt = _i32(tt).reshape(-1, 3)
        self.nv, self.nt = len(self.rr), len(self.tt)
        self.md = _lib().orc_mesh_md(self.nv, self.nt, _ptr(self.tt, I32P))
        self.adj0 = np.zeros(self.md * self.nv, np.int32)
        self.adj1 = np.zeros(self.md * self.nv, np.int32)
        _lib().orc_adj_ini(self.md, self.nt, self.nv, _ptr(self.tt, I32P), _ptr(self.adj0, I32P),
                           _ptr(self.adj1, I32P))

    def map(self, i):
        m = np.zeros(6, np.int32)
        ok = _lib().orc_adj_map(self.md, self.nv, _ptr(self.adj0, I32P), _ptr(self.adj1, I32P), i, _ptr(m, I32P))
        return m if ok else None

    def shape(self):
        """stress-free (ll, aa) per slot of the template itself"""
        n = self.md * self.nv
        ll, aa = np.zeros(n, np.float32), np.zeros(n, np.float32)
        _lib().orc_rbc_shape(self.md, self.nv, _ptr(self.adj0, I32P), _ptr(self.adj1, I32P), _ptr(self.rr, F32P),
                             _ptr(ll, F32P), _ptr(aa, F32P))
        return ll, aa

    def stressful(self, totArea):
        l0, a0 = np.zeros(1, np.float32), np.zeros(1, np.float32)
        _lib().orc_rbc_stressful(self.nt, np.float32(totArea), _ptr(l0, F32P), _ptr(a0, F32P))
        return float(l0[0]), float(a0[0])


def par(**kw):
    p = RbcPar()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def area_volume(mesh, pp, nc):
    pp = _f32(pp).reshape(-1, 6)
    av = np.zeros(2 * nc, np.float64)
    _lib().orc_rbc_area_volume(mesh.nt, mesh.nv, nc, _ptr(mesh.tt, I32P), _ptr(pp, F32P), _ptr(av, F64P))
    return av


def forces(p, mesh, pp, nc, av, stress_free=None, l0=0.0, a0=0.0, bending=True):
    """membrane (+ Kantor bending) forces, double (nc*nv, 3)"""
    pp = _f32(pp).reshape(-1, 6)
    ff = np.zeros((nc * mesh.nv, 3), np.float64)
    ll = aa = None
    if stress_free is not None:
        ll, aa = (_f32(x) for x in stress_free)
    _lib().orc_rbc_forces(C.byref(p), mesh.md, mesh.nv, nc, _ptr(mesh.adj0, I32P), _ptr(mesh.adj1, I32P),
                          _ptr(pp, F32P), _ptr(_f64(av), F64P), _ptr(ll, F32P) if ll is not None else None,
                          _ptr(aa, F32P) if aa is not None else None, np.float32(l0), np.float32(a0),
                          _ptr(ff, F64P))
    if bending:
        _lib().orc_rbc_bending(C.byref(p), mesh.md, mesh.nv, nc, _ptr(mesh.adj0, I32P), _ptr(mesh.adj1, I32P),
                               _ptr(pp, F32P), _ptr(ff, F64P))
    return ff


def bending(p, mesh, pp, nc):
    pp = _f32(pp).reshape(-1, 6)
    ff = np.zeros((nc * mesh.nv, 3), np.float64)
    _lib().orc_rbc_bending(C.byref(p), mesh.md, mesh.nv, nc, _ptr(mesh.adj0, I32P), _ptr(mesh.adj1, I32P),
                           _ptr(pp, F32P), _ptr(ff, F64P))
    return ff


def energy(p, mesh, r, nc, terms, stress_free=None, l0=0.0, a0=0.0):
    r = np.ascontiguousarray(r, np.float64).reshape(-1, 3)
    ll = aa = None
    if stress_free is not None:
        ll, aa = (_f32(x) for x in stress_free)
    return _lib().orc_rbc_energy(C.byref(p), mesh.md, mesh.nv, nc, mesh.nt, _ptr(mesh.tt, I32P),
                                 _ptr(mesh.adj0, I32P), _ptr(mesh.adj1, I32P), _ptr(r, F64P),
                                 _ptr(ll, F32P) if ll is not None else None,
                                 _ptr(aa, F32P) if aa is not None else None, np.float32(l0), np.float32(a0), terms)


def stretch(nc, nv, f, ff):
    ff = _f32(ff).reshape(-1, 3).copy()
    _lib().orc_rbc_stretch(nc, nv, _ptr(_f32(f), F32P), _ptr(ff, F32P))
    return ff


def place(rr, matrices, L):
    """cells from the template by 4x4 affine matrices (row major), in double,
    shifted to local coordinates (src/mesh/gen/imp/main.h:1-56)"""
    out = []
    for A in matrices:
        A = np.asarray(A, np.float64).reshape(4, 4)
        h = np.concatenate([rr.astype(np.float64), np.ones((len(rr), 1))], axis=1)
        g = (h @ A.T)[:, :3].astype(np.float32)
        g = g - (np.array(L, np.float32) / np.float32(2))
        out.append(g)
    r = np.concatenate(out).astype(np.float32)
    pp = np.zeros((len(r), 6), np.float32)
    pp[:, :3] = r
    return pp


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def fsi_forces(L, spp, starts, opp, params, seeds, fo=None, fs=None):
    """solvent-membrane DPD; returns (fo, fs, pairs)"""
    spp = _f32(spp).reshape(-1, 6)
    opp = _f32(opp).reshape(-1, 6)
    fo = np.zeros((len(opp), 3), np.float32) if fo is None else _f32(fo).copy()
    fs = np.zeros((len(spp), 3), np.float32) if fs is None else _f32(fs).copy()
    n = _lib().orc_fsi_forces(_ptr(_i32(L), I32P), len(spp), _ptr(spp, F32P), _ptr(_i32(starts), I32P), len(opp),
                              _ptr(opp, F32P), Dpd(*params), _ptr(_f32(seeds), F32P), _ptr(fo, F32P),
                              _ptr(fs, F32P))
    return fo, fs, int(n)


def fsi_seeds(nsteps):
    out = np.zeros(2 * nsteps, np.float32)
    _lib().orc_fsi_seeds(nsteps, _ptr(out, F32P))
    return out.reshape(nsteps, 2)


def adjoint(p, mesh, pp, nc, g, stress_free=None, l0=0.0, a0=0.0):
    """(gx, terms (5, n, 3), gp (5,)): the membrane force's vector-Jacobian
    product with g, the unit-coefficient term forces and g . term; pp is
    (n, 6) particles or (n, 3) positions, taken in double"""
    pp = np.asarray(pp)
    x = _f64(pp.reshape(len(pp), -1)[:, :3])
    g = _f64(g).reshape(-1, 3)
    n = nc * mesh.nv
    gx = np.zeros((n, 3), np.float64)
    terms = np.zeros((5, n, 3), np.float64)
    gp = np.zeros(5, np.float64)
    ll = aa = None
    if stress_free is not None:
        ll, aa = (_f32(x) for x in stress_free)
    _lib().orc_rbc_adjoint(C.byref(p), mesh.md, mesh.nv, nc, mesh.nt, _ptr(mesh.tt, I32P), _ptr(mesh.adj0, I32P),
                           _ptr(mesh.adj1, I32P), _ptr(x, F64P), _ptr(ll, F32P) if ll is not None else None,
                           _ptr(aa, F32P) if aa is not None else None, np.float32(l0), np.float32(a0),
                           _ptr(g, F64P), _ptr(gx, F64P), _ptr(terms, F64P), _ptr(gp, F64P))
    return gx, terms, gp


def redistribute(L, opp, nc, nv):
    """Whole-cell periodic redistribution: a cell whose bounding-box centre
    0.5 (min + max) left [-L/2, L/2) on an axis is shifted by -L d
    (src/distr/rbc/dev.h:3-20, src/distr/map/dev.h:17-26,
    src/distr/common/dev.h:20-32); the kept cells come first, then fragments
    0..25 (src/distr/rbc/imp/unpack.h:30-51), ascending within each group.
    Returns (pp, fragment id of each input cell)."""
    x = _f32(opp).reshape(nc, nv, 6).copy()
    Li = np.array(L, np.int64)
    r = np.float32(0.5) * (x[:, :, :3].min(axis=1) + x[:, :, :3].max(axis=1))
    d = -1 + (r >= (-(Li // 2)).astype(np.float32)).astype(np.int64) + (r >= (Li // 2).astype(np.float32))
    fid = (d[:, 0] + 2) % 3 + 3 * ((d[:, 1] + 2) % 3) + 9 * ((d[:, 2] + 2) % 3)
    x[:, :, :3] = x[:, :, :3] + (-d * Li).astype(np.float32)[:, None, :]
    order = [c for c in range(nc) if fid[c] == 26] + [c for f in range(26) for c in range(nc) if fid[c] == f]
    return x[order].reshape(-1, 6), fid


def suspension_step(L, spp, opp, mesh, nc, dpd, seeds28, fsi_par, fsi_seed2, p, dt, mass=1.0, rbc_mass=0.5,
                    substeps=1, stress_free=None, l0=0.0, a0=0.0, nthreads=1, bounce=False, cnt_par=None,
                    cnt_seeds=None):
    """One step of solvent + membranes on one periodic domain in the order of
    src/sim/imp/run.h:8-68: solvent cell list, membrane redistribution,
    solvent DPD, FSI (+f vertices, -f solvent), solvent kick-drift + wrap,
    then per sub-step the vertex force = FSI force + float(membrane force)
    and the vertex kick-drift (src/sim/objects/imp/main.h:40-61).  Returns
    (cell-ordered solvent, vertices, solvent pairs, FSI pairs).  With bounce,
    the solvent bounces back off the membranes after the advance
    (src/sim/imp/run.h:54, src/sim/objects/imp/bounce.h:236-248) before the
    wrap, and the number of bounced particles is returned fifth."""
    from . import oracle as O
    starts, counts, srt, _ = O.clist_build(L, spp)
    opp, _ = redistribute(L, opp, nc, mesh.nv)
    start = opp.copy()
    ff, npair = O.dpd_forces(L, srt, starts, counts, dpd, seeds28[:27], nthreads=nthreads)
    fo, ff, nfsi = fsi_forces(L, srt, starts, opp, fsi_par, fsi_seed2, fs=ff)
    if cnt_par is not None:   # contact with the FSI force (src/sim/objinter/imp/main.h:42-80)
        fo, _ = cnt_forces(L, opp, cnt_par, cnt_seeds[0], cnt_seeds[1], ff=fo)
    sout = O.integrate(L, srt, ff, dt, mass, wrap=not bounce)
    sub_dt = np.float32(np.float32(dt) / np.float32(substeps))
    for _ in range(substeps):
        F = forces(p, mesh, opp, nc, area_volume(mesh, opp, nc), stress_free=stress_free, l0=l0, a0=a0)
        ffast = fo + F.astype(np.float32)
        opp = O.move(opp, ffast, sub_dt, rbc_mass)
    if not bounce:
        return sout, opp, npair // 2, nfsi
    from . import mbb as B
    sout, opp, _, _, ncol = B.bounce(dt, mass, rbc_mass, L, nc, mesh.tt, start[:, :3], opp, starts, srt, sout)
    return O.wrap(L, sout), opp, npair // 2, nfsi, ncol


def jul_edges(mesh):
    """(ne, 4) edges (i, j, k, l) in the oracle's numbering"""
    tt = _i32(mesh.tt)
    dih = np.zeros((3 * len(tt), 4), np.int32)
    ne = _lib().orc_jul_edges(mesh.nv, len(tt), _ptr(tt, I32P), _ptr(dih, I32P))
    return dih[:ne].copy()


def jul_theta(a, b, c, d):
    return _lib().orc_jul_theta(*[_ptr(_f64(x), F64P) for x in (a, b, c, d)])


def jul_forces(mesh, pp, nc):
    """Juelicher bending (f) and area-difference (fad) forces, 3 floats per
    vertex each (src/rbc/force/bending/juelicher)"""
    dih = jul_edges(mesh)
    pp = _f32(pp).reshape(-1, 6)
    f = np.zeros((len(pp), 3), np.float32)
    fad = np.zeros((len(pp), 3), np.float32)
    tt = _i32(mesh.tt)
    _lib().orc_jul_forces(nc, mesh.nv, len(tt), _ptr(tt, I32P), len(dih), _ptr(dih, I32P), _ptr(pp, F32P),
                          _ptr(f, F32P), _ptr(fad, F32P))
    return f, fad


def cnt_pair(params, a, b, rnd):
    f = np.zeros(3, np.float32)
    p = DpdLj(*params)
    ok = _lib().orc_cnt_pair(C.byref(p), _ptr(_f32(a), F32P), _ptr(_f32(b), F32P), np.float32(rnd), _ptr(f, F32P))
    return f, bool(ok)


def cnt_images(L, pp):
    pp = _f32(pp).reshape(-1, 6)
    img = np.zeros((26 * max(len(pp), 1), 6), np.float32)
    iv = np.zeros(26 * max(len(pp), 1), np.int32)
    n = _lib().orc_cnt_images(_ptr(_i32(L), I32P), len(pp), _ptr(pp, F32P), _ptr(img, F32P), _ptr(iv, I32P))
    return img[:n].copy(), iv[:n].copy()


def cnt_forces(L, pp, params, seed_bulk, seed_halo, ff=None):
    """membrane contact forces (src/cnt) added to ff; returns (ff, pairs)"""
    pp = _f32(pp).reshape(-1, 6)
    ff = np.zeros((len(pp), 3), np.float32) if ff is None else _f32(ff).copy()
    n = _lib().orc_cnt_forces(_ptr(_i32(L), I32P), len(pp), _ptr(pp, F32P), DpdLj(*params), np.float32(seed_bulk),
                              np.float32(seed_halo), _ptr(ff, F32P))
    return ff, int(n)


def suspension_wall_step(L, spp, opp, mesh, nc, dpd, seeds28, fsi_par, fsi_seed2, mwall_seed, p, dt, sdf, wpp, wstarts,
                         lam=0.0, mass=1.0, rbc_mass=0.5, substeps=1, l0=0.0, a0=0.0, nthreads=1,
                         accel=(0.0, 0.0, 0.0), bounce=False):
    """One step of solvent + membranes between walls, in the order of
    src/sim/imp/run.h:8-68 and src/sim/imp/force.h:122-141: solvent cell list,
    membrane redistribution, solvent DPD, wall DPD of the solvent (seed
    seeds28[27]) and of the vertices (FSI pair, seed mwall_seed), wall
    repulsion of the vertices (lam > 0), FSI, solvent kick-drift + wall
    bounce-back + wrap, membrane sub-steps, the vertices' wall bounce-back.
    Returns (cell-ordered solvent, vertices, solvent pairs, FSI pairs,
    solvent-wall pairs, vertex-wall pairs).  With bounce, the solvent moves
    without the wrap, bounces back off the membranes after the vertices' wall
    bounce-back (src/sim/imp/run.h:54-56) and is wrapped after it; the number
    of bounced particles is returned seventh."""
    from . import oracle as O
    from . import wall as W
    starts, counts, srt, _ = O.clist_build(L, spp)
    opp, _ = redistribute(L, opp, nc, mesh.nv)
    start = opp.copy()
    ff, npair = O.dpd_forces(L, srt, starts, counts, dpd, seeds28[:27], nthreads=nthreads)
    ff, nw = W.wall_forces(L, srt, sdf, wpp, wstarts, dpd, seeds28[27], ff)
    fo, nwm = W.wall_forces(L, opp, sdf, wpp, wstarts, fsi_par, mwall_seed)
    if lam > 0:
        fo = sdf.repulse(opp, lam, fo)
    fo, ff, nfsi = fsi_forces(L, srt, starts, opp, fsi_par, fsi_seed2, fo=fo, fs=ff)
    sout = W.integrate_wall(L, srt, ff, dt, sdf, mass, accel, wrap=not bounce)
    sub_dt = np.float32(np.float32(dt) / np.float32(substeps))
    for _ in range(substeps):
        F = forces(p, mesh, opp, nc, area_volume(mesh, opp, nc), l0=l0, a0=a0)
        opp = O.move(opp, fo + F.astype(np.float32), sub_dt, rbc_mass)
    opp = sdf.bounce(opp, dt)
    if not bounce:
        return sout, opp, npair // 2, nfsi, nw, nwm
    from . import mbb as B
    sout, opp, _, _, ncol = B.bounce(dt, mass, rbc_mass, L, nc, mesh.tt, start[:, :3], opp, starts, srt, sout)
    return O.wrap(L, sout), opp, npair // 2, nfsi, nw, nwm, ncol
