"""Membrane test cases shared by the oracle and GPU tests: the reference's
meshes, an icosahedron with closed-form area and volume, the adjoint set-up
and the membrane calibration case (C4) with its FP64 CPU gradient.

The calibration case: nc copies of the 642-vertex biconcave cell, vertices
perturbed by Gaussian noise (sigma 0.05), opposed point loads along +-x on 5%
of the vertices (the 2.5% of largest and of smallest template x, applied as
the reference's stretch loads, src/rbc/stretch/imp/main.h), K explicit
relaxation steps x <- x + eta (F(x; k) + f_load) with the conservative
membrane force F (the viscous term is zero at zero velocities), and the loss
L = sum |x_K - x*|^2 against the target x*: each cell's noise-free template
stretched by 1.1 along x and by 1/sqrt(1.1) across it about its centre.
"""
import os

import numpy as np

from oracle import rbc as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref")
COEF = ("ka", "kd", "kv", "ks", "kb")


def mesh(nv):
    return R.Mesh(*R.read_off(os.path.join(GOLD, "cells", "%d.off" % nv)))


def icosahedron():
    """unit-circumradius icosahedron, outward-oriented faces"""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v[0])
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
                  [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int32)
    return v, f


def adj_setup(nc=2, seed=1, stress_free=False, ico=False):
    """noisy 642-vertex cells (sigma 0.05) off their targets; ico: a scaled
    icosahedron instead (all dihedrals far from flat)"""
    if ico:
        v, f = icosahedron()
        m = R.Mesh((2.0 * v).astype(np.float32), f)
    else:
        m = mesh(642)
    rng = np.random.default_rng(seed)
    T = [np.eye(4) for _ in range(nc)]
    for k, A in enumerate(T):
        A[:3, 3] = [3.0 * k, 0.0, 0.0]
    pp = R.place(m.rr, T, (0, 0, 0))
    pp[:, :3] += rng.normal(0, 0.05, (len(pp), 3)).astype(np.float32)
    tmpl = np.zeros((m.nv, 6), np.float32)
    tmpl[:, :3] = m.rr
    A0, V0 = R.area_volume(m, tmpl, 1)
    kw = dict(R.RBC_TEST, gammaC=0.0, gammaT=0.0, totArea=np.float32(A0 * 0.97), totVolume=np.float32(V0 * 1.03))
    sf = m.shape() if stress_free else None
    l0, a0 = m.stressful(kw["totArea"])
    g = rng.normal(0, 1, (len(pp), 3))
    return m, pp, kw, sf, l0, a0, g


class C4:
    """The membrane calibration case; positions are float32 (n, 3)."""
    ETA = 5e-7          # relaxation step: eta * max|F| about 0.02 at the noisy start
    STEPS = 4
    LOAD = 50.0         # total load per side

    def __init__(self, nc=256, seed=1, phi=None, stress_free=True):
        m = self.m = mesh(642)
        self.nc = nc
        T = [np.eye(4) for _ in range(nc)]
        for k, A in enumerate(T):
            A[:3, 3] = [12.0 * (k % 16), 12.0 * (k // 16), 12.0 * (k // 256)]
        clean = R.place(m.rr, T, (0, 0, 0))[:, :3].astype(np.float64)
        rng = np.random.default_rng(seed)
        self.x0 = (clean + rng.normal(0, 0.05, clean.shape)).astype(np.float32)
        s = np.array([1.1, 1.1 ** -0.5, 1.1 ** -0.5])
        c = clean.reshape(nc, m.nv, 3).mean(axis=1, keepdims=True)
        self.target = (c + (clean.reshape(nc, m.nv, 3) - c) * s).reshape(-1, 3)
        tmpl = np.zeros((m.nv, 6), np.float32)
        tmpl[:, :3] = m.rr
        A0, V0 = R.area_volume(m, tmpl, 1)
        self.kw = dict(R.RBC_TEST, gammaC=0.0, gammaT=0.0, totArea=np.float32(A0), totVolume=np.float32(V0))
        if phi is not None:
            self.kw["phi"] = phi
        self.sf = m.shape() if stress_free else None
        self.l0, self.a0 = m.stressful(self.kw["totArea"])
        order = np.argsort(m.rr[:, 0], kind="stable")
        k = int(round(0.025 * m.nv))
        self.load = np.zeros((m.nv, 3), np.float32)
        self.load[order[:k], 0] = -self.LOAD / k
        self.load[order[-k:], 0] = self.LOAD / k
        self.u = rng.normal(0, 1, self.x0.shape)          # a direction for directional derivatives

    def coef(self):
        return np.array([self.kw[c] for c in COEF], np.float64)

    def par(self, coef=None):
        kw = dict(self.kw)
        if coef is not None:
            kw.update({c: float(v) for c, v in zip(COEF, coef)})
        return R.par(**kw)

    def forces64(self, x, coef=None):
        """FP64 membrane force at double positions: sum of coefficient times
        unit-coefficient term force"""
        k = self.coef() if coef is None else np.asarray(coef, np.float64)
        _, terms, _ = R.adjoint(self.par(), self.m, x, self.nc, np.zeros_like(x), self.sf, self.l0, self.a0)
        return np.tensordot(k, terms, 1)

    def loss64(self, x0=None, coef=None):
        """(L, trajectory) in FP64"""
        x = np.asarray(self.x0 if x0 is None else x0, np.float64)
        load = np.tile(self.load, (self.nc, 1)).astype(np.float64)
        xs = [x]
        for _ in range(self.STEPS):
            x = x + self.ETA * (self.forces64(x, coef) + load)
            xs.append(x)
        return ((x - self.target) ** 2).sum(), xs

    def grad64(self, coef=None):
        """FP64 CPU gradient (L, dL/dx0, dL/dk) through the oracle adjoint"""
        p = self.par(coef)
        L, xs = self.loss64(coef=coef)
        lam = 2 * (xs[-1] - self.target)
        gk = np.zeros(5)
        for t in range(self.STEPS - 1, -1, -1):
            gx, _, gp = R.adjoint(p, self.m, xs[t], self.nc, self.ETA * lam, self.sf, self.l0, self.a0)
            gk += gp
            lam = lam + gx
        return L, lam, gk

    def grad_f32(self):
        """dL/dx0 with the arithmetic of the float32 path: float32 states and
        cotangents, the production force formulas (oracle forces, rounded to
        float), the FP64 oracle adjoint rounded to float"""
        p = self.par()
        eta = np.float32(self.ETA)
        load = np.tile(self.load, (self.nc, 1))
        x = self.x0.copy()
        xs = [x]
        for _ in range(self.STEPS):
            pp = np.zeros((len(x), 6), np.float32)
            pp[:, :3] = x
            F = R.forces(p, self.m, pp, self.nc, R.area_volume(self.m, pp, self.nc), stress_free=self.sf,
                         l0=self.l0, a0=self.a0).astype(np.float32)
            x = (x + eta * (F + load)).astype(np.float32)
            xs.append(x)
        lam = (2 * (xs[-1].astype(np.float64) - self.target)).astype(np.float32)
        for t in range(self.STEPS - 1, -1, -1):
            g = (eta * lam).astype(np.float64)
            gx, _, _ = R.adjoint(p, self.m, xs[t], self.nc, g, self.sf, self.l0, self.a0)
            lam = (lam + gx.astype(np.float32)).astype(np.float32)
        return lam.astype(np.float64)


def rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def bounce_case(seed=1, L=(16, 16, 16), rho=10, dt=0.02, vsig=1.0, wobble=0.01, centres=None):
    """A membrane bounce-back set-up: 498-vertex cells (one straddling the +x
    face, so that its periodic image acts too) moving rigidly plus vertex noise
    over the step (rp -> rc), and solvent moving in straight lines (pp0 -> pp,
    cell-ordered by the start positions)."""
    import oracle as O
    m = mesh(498)
    rng = np.random.default_rng(seed)
    if centres is None:
        centres = [(L[0] / 2 - 1.0, 0.5, -0.3), (-3.0, -4.0, 3.5), (-2.5, 4.5, -4.0)]
    nm, nv = len(centres), m.nv
    rp = np.zeros((nm * nv, 3), np.float32)
    rc = np.zeros((nm * nv, 6), np.float32)
    for k, c in enumerate(centres):
        x = m.rr.astype(np.float64) @ rotation(rng).T + np.array(c)
        rp[k * nv:(k + 1) * nv] = x.astype(np.float32)
        shift = rng.normal(0, 0.01, 3)
        small = np.eye(3) + 0.004 * rng.normal(size=(3, 3))
        y = (x - np.array(c)) @ small.T + np.array(c) + shift + rng.normal(0, wobble, x.shape)
        rc[k * nv:(k + 1) * nv, :3] = y.astype(np.float32)
    rc[:, 3:] = rng.normal(0, 0.1, (nm * nv, 3)).astype(np.float32)
    n = int(rho * np.prod(L))
    pp0 = np.zeros((n, 6), np.float32)
    pp0[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    starts, counts, srt0, order = O.clist_build(L, pp0)
    v = rng.normal(0, vsig, (len(srt0), 3)).astype(np.float32)
    pp = np.zeros_like(srt0)
    pp[:, :3] = srt0[:, :3] + v * np.float32(dt)
    pp[:, 3:] = v
    srt0[:, 3:] = v
    return dict(L=L, dt=dt, nm=nm, nv=nv, tt=m.tt, rp=rp, rc=rc, starts=starts, pp0=srt0, pp=pp)
