"""Benchmark of the DPD per-timestep hot path on one MI355X.

Workload: wall-bounded Poiseuille flow in a 96^3 box at number density 10
(a = 2.5, gamma = 8, kBT = 0.0444, dt = 5e-4, spow = 0.25).  The walls come
from a 128^3 sdf.dat-style field of a plane channel, s(z) = |z| - 40 (positive
inside the walls), resampled onto the wall texture.  The box is filled with
uniform random positions and Maxwell velocities, equilibrated as a periodic
fluid, then the wall particles are frozen (0 <= s <= 2), the particles inside
the walls dropped and the body force 0.01 x switched on.

One step = cell list (counting sort + reorder + fragment tables), solvent DPD
(conservative + dissipative + random with the step-seeded hash RNG), wall
DPD, body force + kick-drift, bounce-back on the walls and periodic wrap.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DPD pair-interactions/s & steps/s on 1 MI355X; batched vs per-box launches"
HBM_PEAK_GBS = 8000.0
# vector FP32 peak without packed (two-per-lane) instructions, which the
# library is built without: 256 CUs x 4 SIMDs x 16 lanes x 2 flop x 2.4 GHz
FP32_PEAK_TFLOPS = 78.6
FP32_PACKED_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_dpd_gather_96cube_wall.md")
# the sources the force kernel is built from; the committed traffic figure
# is used only while their hash matches the one it was measured with
KERNEL_SOURCES = [os.path.join(ROOT, "ctc_amd", "csrc", f) for f in ("dpd.hip", "udx_dev.h", "clist.hip")]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--phase-steps", type=int, default=50, help="steps of the per-phase timing run after the timed one")
    ap.add_argument("--L", type=int, default=96)
    ap.add_argument("--rho", type=float, default=10.0)
    ap.add_argument("--half-width", type=float, default=40.0, help="channel half-width (0: periodic, no walls)")
    ap.add_argument("--sdf-grid", type=int, default=128)
    ap.add_argument("--equil", type=int, default=200, help="periodic equilibration steps before the freeze")
    ap.add_argument("--cpu-steps", type=int, default=4,
                    help="timed steps of the CPU baseline after one warm-up step (0: skip)")
    ap.add_argument("--c1-steps", type=int, default=1000, help="steps per timed run of the 16^3 leg (0: skip)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="MiB copied by one plain device copy before the workload (calibrates HBM counters)")
    ap.add_argument("--ens-boxes", type=int, default=64, help="boxes of the batched-ensemble leg (0: skip)")
    ap.add_argument("--ens-steps", type=int, default=40, help="timed steps of the ensemble leg")
    ap.add_argument("--adj-cells", type=int, default=256, help="cells of the membrane adjoint leg (0: skip)")
    ap.add_argument("--susp-steps", type=int, default=20, help="timed steps of the RBC suspension leg (0: skip)")
    ap.add_argument("--half-steps", type=int, default=20,
                    help="timed steps of the same wall-bounded box in the half-shell force mode (0: skip)")
    ap.add_argument("--adj-reps", type=int, default=20, help="timed forward + adjoint batches")
    return ap.parse_args()


def plane_channel(L, N, h):
    """s(z) = |z| - h at the cell centres of an N^3 grid over the box, x
    fastest (the sdf.dat layout)."""
    import numpy as np
    z = (np.arange(N) + 0.5) * (L[2] / N) - 0.5 * L[2]
    s = (np.abs(z) - h).astype(np.float32)
    return np.ascontiguousarray(np.broadcast_to(s[:, None, None], (N, N, N)))


def kernel_hash():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def read_traffic():
    """HBM bytes per k_dpd launch (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated) from the committed rocprofv3 counter summary, if it was
    measured on the current kernel sources (its kernel_sources_sha16 line);
    None when it is stale or missing."""
    val = stamp = None
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                val = float(line.split(":")[1])
            if line.startswith("kernel_sources_sha16:"):
                stamp = line.split(":")[1].strip()
    except OSError:
        return None
    return val if stamp == kernel_hash() else None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the path would shard by "
              "3-D domain decomposition)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ctc_amd import udx as U

    torch.cuda.set_device(0)
    if args.copy_calib > 0:
        # a plain copy of known size: reads and writes copy_calib MiB each
        src = torch.ones(args.copy_calib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    L = (args.L, args.L, args.L)
    ncells = L[0] * L[1] * L[2]
    n0 = int(round(args.rho * ncells))
    a, g, kBT, dt, fx = 2.5, 8.0, 0.0444, 5e-4, 0.01
    params = U.dpd_params(a, g, kBT, dt, 0.25)
    walls = args.half_width > 0

    # synthetic state: uniform positions, Maxwell velocities with zero net momentum
    gen = torch.Generator(device="cuda").manual_seed(1234)
    pp = torch.empty((n0, 6), dtype=torch.float32, device="cuda")
    pp[:, :3] = (torch.rand((n0, 3), generator=gen, device="cuda") - 0.5) * float(args.L)
    pp[:, :3].clamp_(-0.5 * args.L, 0.5 * args.L - 1e-4)
    v = torch.randn((n0, 3), generator=gen, device="cuda", dtype=torch.float64) * np.sqrt(kBT)
    v -= v.mean(0)
    pp[:, 3:] = v.float()

    ctx = U.Context()
    t_setup = time.perf_counter()
    # equilibration as a periodic fluid, no body force
    eq = U.Flu(ctx, L, n0, params, dt, 1.0, (0.0, 0.0, 0.0))
    eq.set_pp(pp)
    eq.run(args.equil)
    eq.get_pp(pp)
    eq.close()

    sdf = wall = None
    n = n0
    D = None
    if walls:
        N = args.sdf_grid
        D = plane_channel(L, N, args.half_width)
        sdf = U.Sdf(ctx, L, ext=[float(x) for x in L], N=(N, N, N), D=D)
        wall = U.Wall(ctx, sdf, pp)             # freeze: pp[:n_solvent] is the solvent
        n = wall.n_solvent
    flu = U.Flu(ctx, L, n, params, dt, 1.0, (fx, 0.0, 0.0))
    flu.set_pp(pp[:n].contiguous())
    if wall is not None:
        flu.set_wall(wall)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    flu.run(args.warmup)
    torch.cuda.synchronize()

    # timed region: the steps alone (no timing events)
    flu.pairs(reset=True)
    w0 = flu.wall_pairs() if walls else 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    flu.run(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pairs = flu.pairs()
    wpairs = (flu.wall_pairs() - w0) if walls else 0
    # per-phase kernel times from HIP events recorded on the library's stream
    # around each phase of the next steps (read after the run)
    flu.timing(True)
    flu.run(args.phase_steps)
    torch.cuda.synchronize()
    t_phase = flu.timing(False)          # ms: clist, solvent DPD, wall DPD, integrate, sum
    sec = t1 - t0
    ms_step = 1e3 * sec / args.steps
    steps_s = args.steps / sec
    pairs_step = (pairs + wpairs) / args.steps
    pairs_s = (pairs + wpairs) / sec

    # dominant kernel: the solvent force gather k_dpd, one launch per step;
    # in this step (no objects) it also evaluates the wall forces, moves the
    # particles and counts them into the next step's slabs.  Algorithmic bytes
    # per launch = n * (36 + 60 + 16) B: forces (24 B particle in, 12 B force
    # out), integrate (36 B in, 24 B out), slab count (16 B)
    t_force = t_phase[1] * 1e-3
    force_bytes = n * (36.0 + 60.0 + 16.0)
    force_gbs = force_bytes / t_force / 1e9
    # algorithmic FLOPs: 100 per solvent or wall pair (RNG included) + 9 per
    # half-shell distance test
    flops = ((pairs + wpairs) / args.steps) * 100.0 + n * 13.5 * args.rho * 9.0
    step_bytes = n * 164.0 + ncells * 16.0
    traffic = read_traffic()

    out = {
        "metric": METRIC,
        "value": pairs_s,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "steps_per_s": steps_s,
        "pairs_per_step": pairs_step,
        "wall_pairs_per_step": wpairs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform positions, Maxwell velocities, seed 1234; %d periodic equilibration steps "
                "before the wall freeze)" % args.equil,
        "config": {
            "workload": ("wall_poiseuille_%d^3_rho%g_plane_channel_h%g" % (args.L, args.rho, args.half_width)
                         if walls else "dpd_%d^3_rho%g_periodic" % (args.L, args.rho)),
            "box": list(L), "particles_before_freeze": n0, "solvent_particles": n,
            "wall_particles": wall.counts()[1] if walls else 0,
            "a": a, "gamma": g, "kBT": kBT, "dt": dt, "spow": 0.25, "body_force_x": fx,
            "walls": walls, "sdf_grid": args.sdf_grid if walls else None,
            "parallelism": "single GPU",
        },
        # HIP-event phases of args.phase_steps further steps; the gather also
        # does the walls and the move, so wall_dpd and integrate are gaps
        "phase_ms": {"clist": t_phase[0], "dpd": t_phase[1], "wall_dpd": t_phase[2], "integrate": t_phase[3]},
        "phase_steps": args.phase_steps,
        "setup_s": t_setup,
        "roofline": {
            "bound": "hbm", "kernel": "k_dpd (solvent force gather with the wall forces, the move and the slab "
                                      "count)",
            "achieved": force_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": force_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "step_roofline": {
            "achieved": step_bytes / (ms_step * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_step": step_bytes,
        },
        "valu": {
            "kernel": "k_dpd",
            "achieved": flops / t_force / 1e12, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": flops / t_force / 1e12 / FP32_PEAK_TFLOPS, "flop_per_launch": flops,
            "frac_of_packed_peak": flops / t_force / 1e12 / FP32_PACKED_PEAK_TFLOPS,
        },
        "kernel_sources_sha16": kernel_hash(),
    }

    if args.half_steps > 0:
        out["half_shell_mode"] = half_leg(ctx, flu, L, n, params, dt, fx, wall, args.half_steps)

    if args.cpu_steps > 0:
        cur = torch.empty((n, 6), dtype=torch.float32, device="cuda")
        flu.get_pp(cur)
        wdata = None
        if walls:
            wdata = (D, host(wall.particles()), host(wall.starts()))
        cb = cpu_baseline(host(cur), L, (a, g, params.s, 0.25), dt, fx, args.cpu_steps, wdata, args.sdf_grid)
        cb["gpu_over_cpu"] = pairs_s / cb["value"] if cb["value"] else None
        out["cpu_baseline"] = cb

    flu.close()
    if wall is not None:
        wall.close()
        sdf.close()
    del pp
    torch.cuda.empty_cache()
    if args.c1_steps > 0:
        out["c1"] = c1_leg(ctx, args.c1_steps, args.cpu_steps > 0)
    if args.ens_boxes > 0:
        out["ensemble"] = ensemble_leg(ctx, args.ens_boxes, args.ens_steps)
    if args.adj_cells > 0:
        out["adjoint"] = adjoint_leg(ctx, args.adj_cells, args.adj_reps, args.cpu_steps > 0)
    if args.susp_steps > 0:
        out["suspension"] = suspension_leg(ctx, args.susp_steps, 5, args.cpu_steps > 0)
    print(json.dumps(out))


def half_leg(ctx, flu, L, n, params, dt, fx, wall, nsteps):
    """The same state stepped with the half-shell force mode (each pair inside
    the box evaluated once, exact FP64 sums): per-phase times of nsteps steps."""
    import torch

    from ctc_amd import udx as U

    cur = torch.empty((n, 6), dtype=torch.float32, device="cuda")
    flu.get_pp(cur)
    h = U.Flu(ctx, L, n, params, dt, 1.0, (fx, 0.0, 0.0), mode=U.MODE_HALF)
    h.set_pp(cur)
    if wall is not None:
        h.set_wall(wall)
    h.run(3)
    h.pairs(reset=True)
    h.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    h.run(nsteps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    t = h.timing(False)
    pairs = h.pairs()
    h.close()
    del cur
    return {"steps": nsteps, "ms_per_step": sec / nsteps * 1e3, "pairs_per_step": pairs / nsteps,
            "phase_ms": {"clist": t[0], "dpd": t[1], "wall_dpd": t[2], "integrate": t[3]},
            "kernels": "k_dpd_half (half shell, wave-level hit queue, FP64 reactions) + k_acc_to_ff"}


def ensemble_leg(ctx, nbox, nsteps):
    """Batched ensemble vs launching each box on its own: nbox periodic
    boxes, edges uniform in 8..48, density in {3, 5, 10}, gamma in [4, 10],
    kBT in [0.04, 1], a = 25, dt = 1e-3, one seed stream per box.  Both runs
    start from the same state with the same seeds and must end identical."""
    import numpy as np
    import torch

    from ctc_amd import udx as U

    rng = np.random.default_rng(2024)
    boxes, pps = [], []
    for b in range(nbox):
        L = tuple(int(x) for x in rng.integers(8, 49, 3))
        rho = int(rng.choice([3, 5, 10]))
        g, kBT = float(rng.uniform(4, 10)), float(rng.uniform(0.04, 1.0))
        n = rho * L[0] * L[1] * L[2]
        boxes.append(dict(L=L, n=n, params=U.dpd_params(25.0, g, kBT, 1e-3, 0.25), dt=1e-3, mass=1.0,
                          accel=(0.0, 0.0, 0.0)))
        pp = torch.empty((n, 6), dtype=torch.float32, device="cuda")
        pp[:, :3] = (torch.rand((n, 3), device="cuda") - 0.5) * torch.tensor(L, device="cuda", dtype=torch.float32)
        for c in range(3):
            pp[:, c].clamp_(-0.5 * L[c], 0.5 * L[c] - 1e-4)
        pp[:, 3:] = torch.randn((n, 3), device="cuda") * float(np.sqrt(kBT))
        pps.append(pp)
    ntot = sum(b["n"] for b in boxes)
    seeds = U.ens_trunk_seeds(nbox, 2 * nsteps)   # warm-up run, then the timed run
    allpp = torch.cat(pps)

    # batched: autotuned tile, graph-captured step pairs
    ens = U.Ensemble(ctx, boxes)
    ens.set_pp(allpp)
    best, ms = ens.autotune(2)
    ens.set_pp(allpp)
    ens.graph(True)
    ens.run(nsteps, seeds[:nsteps])             # warm-up (captures the graph)
    ens.pairs(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(nsteps, seeds[nsteps:])
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    pairs = ens.pairs()
    got = torch.empty_like(allpp)
    ens.get_pp(got)
    ens.close()

    # each box on its own: its full run of warm-up + timed steps, one box after another
    off = np.cumsum([0] + [b["n"] for b in boxes])
    t_s = 0.0
    same = True
    for b in range(nbox):
        x = boxes[b]
        flu = U.Flu(ctx, x["L"], x["n"], x["params"], x["dt"], 1.0, x["accel"])
        flu.set_pp(pps[b])
        s28 = np.concatenate([seeds[:, b], np.zeros((2 * nsteps, 1), np.float32)], axis=1)
        flu.run(nsteps, s28[:nsteps])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        flu.run(nsteps, s28[nsteps:])
        torch.cuda.synchronize()
        t_s += time.perf_counter() - t0
        mine = torch.empty_like(pps[b])
        flu.get_pp(mine)
        torch.cuda.synchronize()
        same = same and bool(torch.equal(mine.view(torch.int32), got[off[b]:off[b + 1]].view(torch.int32)))
        flu.close()
    return {
        "boxes": nbox, "particles": ntot, "steps": nsteps,
        "batched_steps_per_s": nsteps / t_b, "per_box_steps_per_s": nsteps / t_s,
        "speedup": t_s / t_b, "batched_pairs_per_s": pairs / t_b,
        "tile_variant": best, "tile_ms": ms, "identical_to_per_box": same,
        "workload": "64 periodic boxes, edges U{8..48}, density {3,5,10}, gamma U[4,10], kBT U[0.04,1], "
                    "a = 25, dt = 1e-3, graph-captured, autotuned" if nbox == 64 else "%d boxes" % nbox,
    }


def adjoint_leg(ctx, nc, reps, cpu):
    """Membrane forward + adjoint per batch: nc copies of the 642-vertex
    biconcave cell with Gaussian vertex noise (sigma 0.05), stress-free
    springs, Kantor bending, the reference's RBC test coefficients; a random
    cotangent.  The CPU leg is the FP64 oracle gradient of the same batch."""
    import numpy as np
    import torch

    from ctc_amd import udx as U

    rr, tt = U.read_off(os.path.join(ROOT, "tests", "golden", "ref", "cells", "642.off"))
    nv = len(rr)
    rng = np.random.default_rng(7)
    off = np.stack([12.0 * (np.arange(nc) % 16), 12.0 * (np.arange(nc) // 16 % 16), 12.0 * (np.arange(nc) // 256)], 1)
    x = (rr[None].astype(np.float64) + off[:, None]).reshape(-1, 3) + rng.normal(0, 0.05, (nc * nv, 3))
    pp_h = np.zeros((nc * nv, 6), np.float32)
    pp_h[:, :3] = x
    g_h = rng.normal(0, 1, (nc * nv, 3)).astype(np.float32)
    kw = dict(gammaC=0.0, gammaT=0.0, kBT=0.0444, kb=17.78, phi=6.97, ks=14.8, x0=0.5, mpow=2.0, ka=2450.0,
              kd=100.0, kv=3750.0, totArea=80.0594, totVolume=43.3586)
    r = U.Rbc(ctx, tt, nv, nc)
    r.set_stressfree(rr)
    r.set_bending(True)
    par = U.rbc_params(**kw)
    pp = torch.as_tensor(pp_h).cuda()
    g = torch.as_tensor(g_h).cuda()
    ff = torch.zeros((nc * nv, 3), dtype=torch.float32, device="cuda")
    gx = torch.empty_like(ff)
    for _ in range(3):
        r.force(par, 0.0, nc, pp, ff)
        r.adjoint(par, nc, pp, g, gx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ff.zero_()
        r.force(par, 0.0, nc, pp, ff)
        gp = r.adjoint(par, nc, pp, g, gx)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    r.close()
    out = {"cells": nc, "vertices": nc * nv, "faces": nc * len(tt), "ms_forward_adjoint": ms,
           "dL_dk": gp,
           "workload": "%d x 642-vertex membranes, sigma 0.05 vertex noise, stress-free springs, Kantor "
                       "bending; forward force + adjoint (vertex and coefficient gradients) per batch" % nc}
    if cpu:
        from oracle import rbc as R
        m = R.Mesh(rr, tt)
        p = R.par(**kw)
        sf = m.shape()
        t0 = time.perf_counter()
        R.adjoint(p, m, pp_h, nc, g_h.astype(np.float64), sf)
        sec = time.perf_counter() - t0
        out["cpu_fp64_ms"] = sec * 1e3
        out["cpu_cores"] = 1
        out["gpu_over_cpu"] = sec * 1e3 / ms
    return out


def inside_mesh(pts, rr, tt):
    """points inside a closed triangle mesh: parity of the crossings of a ray
    along +z (Moller-Trumbore against every triangle)"""
    import numpy as np
    v0 = rr[tt[:, 0]].astype(np.float64)
    e1, e2 = rr[tt[:, 1]] - v0, rr[tt[:, 2]] - v0
    d = np.array([0.0, 0.0, 1.0])
    pv = np.cross(d, e2)
    det = (e1 * pv).sum(1)
    cnt = np.zeros(len(pts), np.int64)
    for k in range(0, len(pts), 256):
        q = pts[k:k + 256, None, :] - v0[None]
        u = (q * pv[None]).sum(2) / det
        qv = np.cross(q, e1[None])
        v = (qv * d).sum(2) / det
        t = (qv * e2[None]).sum(2) / det
        cnt[k:k + 256] = ((u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)).sum(1)
    return cnt % 2 == 1


def place_cells(rr, tt, L, lattice, seed, h=0.5):
    """one cell per lattice site, each in a random orientation (and, after
    200 failed tries, a random shift of up to half a cell) tried until its
    solid - the template's interior points on an h grid and its vertices -
    shares no h-voxel of the periodic box with an earlier cell"""
    import numpy as np
    rng = np.random.default_rng(seed)
    lo, hi = rr.min(0), rr.max(0)
    g = np.stack(np.meshgrid(*[np.arange(lo[c] + h / 2, hi[c], h) for c in range(3)], indexing="ij"), -1)
    g = g.reshape(-1, 3) + 1e-3
    solid = np.concatenate([g[inside_mesh(g, rr, tt)], rr.astype(np.float64)])
    L = np.array(L, np.float64)
    G = np.ceil(L / h).astype(int)
    occ = np.zeros(G, bool)
    sp = L / np.array(lattice)
    out = []
    for k in range(lattice[2]):
        for j in range(lattice[1]):
            for i in range(lattice[0]):
                base = (np.array([i, j, k]) + 0.5) * sp - L / 2
                for tr in range(2000):
                    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
                    q *= np.sign(np.linalg.det(q))
                    c = base + rng.uniform(-0.5, 0.5, 3) * (0.3 if tr < 200 else 3.0)
                    idx = np.floor((solid @ q.T + c + L / 2) / h).astype(int) % G
                    if occ[idx[:, 0], idx[:, 1], idx[:, 2]].any():
                        continue
                    occ[idx[:, 0], idx[:, 1], idx[:, 2]] = True
                    out.append(rr.astype(np.float64) @ q.T + c)
                    break
                else:
                    raise RuntimeError("place_cells: no room at lattice site %d %d %d" % (i, j, k))
    return np.concatenate(out).astype(np.float32)


def suspension_leg(ctx, steps, warmup, cpu):
    """RBC suspension: 64^3 periodic box, density-10 solvent (uniform,
    Maxwell kBT), 567 cells of the 498-vertex mesh scaled to a diameter of
    7.8 (hematocrit 0.20) in random orientations on a 9 x 9 x 7 lattice;
    parameters through the configuration layer: conf/default.cfg (pair_flu,
    fsi, rbc.mass, rbc.substeps) with conf/rbc.test.cfg included in the rbc
    group and kb = 40; kBT = 0.0444, dt = 5e-4.  One step = solvent cell
    list, membrane redistribution, solvent DPD, FSI, solvent kick-drift,
    membrane forces and vertex kick-drift."""
    import numpy as np
    import torch

    from ctc_amd import udx as U

    L, lattice, kBT, dt = (64, 64, 64), (9, 9, 7), 0.0444, 5e-4
    rr, tt = U.read_off(os.path.join(ROOT, "tests", "golden", "ref", "cells", "498.off"))
    rr = (rr * np.float32(1.5)).astype(np.float32)
    nv = len(rr)
    vert_h = np.zeros((np.prod(lattice) * nv, 6), np.float32)
    vert_h[:, :3] = place_cells(rr, tt, L, lattice, 2024)
    nc = len(vert_h) // nv
    tmpl = np.zeros((nv, 6), np.float32)
    tmpl[:, :3] = rr
    r1 = U.Rbc(ctx, tt, nv, 1)
    A0, V0 = (float(x) for x in host(r1.area_volume(1, torch.as_tensor(tmpl).cuda())))
    r1.close()
    cdir = os.path.join(ROOT, "tests", "golden", "ref", "conf")
    conf = U.Conf(default=os.path.join(cdir, "default.cfg"),
                  args='rbc = { @include "rbc.test.cfg" totArea = %.6f totVolume = %.6f }' % (A0, V0),
                  include_dir=cdir)
    conf.set("rbc.kb", 40.0)
    sp, fp = conf.pair("pair_flu", kBT, dt), conf.pair("fsi", kBT, dt)
    rp = conf.rbc_params("rbc")
    mass, substeps = conf.float("rbc.mass"), conf.int("rbc.substeps")
    conf.close()
    rng = np.random.default_rng(99)
    n = 10 * L[0] * L[1] * L[2]
    spp_h = np.zeros((n, 6), np.float32)
    spp_h[:, :3] = (rng.uniform(-0.5, 0.5, (n, 3)) * np.array(L)).astype(np.float32)
    spp_h[:, 3:] = rng.normal(0, np.sqrt(kBT), (n, 3)).astype(np.float32)
    flu = U.Flu(ctx, L, n, sp, dt, 1.0)
    flu.set_pp(torch.as_tensor(spp_h).cuda())
    r = U.Rbc(ctx, tt, nv, nc)
    r.set_stressful(rp.totArea)
    r.set_bending(True)
    fsi = U.Fsi(ctx, L, nc * nv)
    vert = torch.as_tensor(vert_h).cuda()
    flu.set_rbc(r, fsi, rp, fp, mass, substeps, nc, vert)
    flu.run(warmup)
    torch.cuda.synchronize()
    flu.pairs(reset=True)
    fsi.pairs(reset=True)
    t0 = time.perf_counter()
    flu.run(steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    npair, nfsi = flu.pairs(), fsi.pairs()
    flu.timing(True)
    flu.run(3)
    ph = flu.timing(False)
    # the same steps with the solvent bounced back off the membranes
    bb = U.Mbb(ctx, nv, tt, nc, n)
    flu.set_rbc_bounce(bb)
    flu.run(2)
    torch.cuda.synchronize()
    bounced = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        flu.run(1)
        bounced += bb.count()
    torch.cuda.synchronize()
    sec_bb = time.perf_counter() - t0
    flu.set_rbc_bounce(None)
    bb.close()
    out = {"cells": nc, "vertices": nc * nv, "solvent": n, "steps": steps, "ms_per_step": sec / steps * 1e3,
           "ms_per_step_with_bounce": sec_bb / steps * 1e3, "bounced_per_step": bounced / steps,
           "steps_per_s": steps / sec, "pairs_per_s": (npair + nfsi) / sec,
           "solvent_pairs_per_step": npair / steps, "fsi_pairs_per_step": nfsi / steps,
           "phase_ms": {"clist_and_redistribution": ph[0], "dpd": ph[1], "fsi": ph[2],
                        "integrate_and_membranes": ph[3]},
           "workload": "RBC suspension: 64^3 periodic, density-10 solvent, %d x 498-vertex cells (D = 7.8, "
                       "hematocrit 0.20), rbc.test.cfg with kb = 40, default.cfg fsi and pair_flu" % nc}
    if cpu:
        from oracle import rbc as R
        import oracle as O
        cur = torch.empty((n, 6), dtype=torch.float32, device="cuda")
        flu.get_pp(cur)
        sh, vh = host(cur), host(vert)
        m = R.Mesh(rr, tt)
        l0, a0 = m.stressful(rp.totArea)
        seeds = O.seeds_table(1)
        p = R.par(**{k: getattr(rp, k) for k, _ in U.RbcParams._fields_})
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        cores = max(1, min(cores, 16))
        t0 = time.perf_counter()
        _, _, cp, cf = R.suspension_step(L, sh, vh, m, nc, (sp.a, sp.g, sp.s, sp.spow), seeds[0],
                                         (fp.a, fp.g, fp.s, fp.spow), R.fsi_seeds(1)[0], p, dt, 1.0, mass,
                                         substeps, None, l0, a0, nthreads=cores)
        csec = time.perf_counter() - t0
        out["cpu_baseline"] = {"steps_per_s": 1 / csec, "pairs_per_s": (cp + cf) / csec, "cores": cores,
                               "kind": "port", "sample": "1 step of the same suspension (solvent DPD on %d "
                               "threads, FSI and membranes on one)" % cores}
    flu.set_rbc(None)
    flu.close()
    r.close()
    return out


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def cpu_threads():
    """host threads of the CPU baseline: the cores this process may run on,
    at most 16 (the host's CPU share per GPU on MI355X nodes)"""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def cpu_baseline(pp, L, params, dt, fx, nsteps, wdata, sdf_grid):
    """The reference's formula on the host cores: the oracle's cell list, the
    reference's half-shell traversal with reactions (one-sided wrapped
    pairs), wall DPD, kick-drift and bounce-back, timed on the same state and
    walls after one warm-up step."""
    import numpy as np

    import oracle as O
    from oracle import wall as W

    cores = cpu_threads()
    seeds = O.seeds_table(nsteps + 1)
    seeds = np.concatenate([seeds, O.kiss_fill(42, 42, 42, 42, nsteps + 1)[:, None]], axis=1)
    q = wpp = wst = None
    if wdata is not None:
        D, wpp, wst = wdata
        q = W.Sdf(L, np.array(L, np.float32), (sdf_grid,) * 3, D.ravel())
    pairs, t0 = 0, None
    for k in range(nsteps + 1):
        if k == 1:
            pairs, t0 = 0, time.perf_counter()
        pp, npair, nw = W.flu_wall_step(L, pp, params, seeds[k], dt, q, wpp, wst, accel=(fx, 0.0, 0.0),
                                        nthreads=cores, half=True)
        pairs += npair + nw
    sec = time.perf_counter() - t0
    return {"value": pairs / sec, "unit": "pair-interactions/s", "steps_per_s": nsteps / sec,
            "cores": cores, "kind": "port",
            "sample": "%d timed step(s) after 1 warm-up of the same %dx%dx%d %s box (clist + half-shell solvent "
                      "DPD + wall DPD + integrate + bounce), %d threads" % (nsteps, *L,
                                                                           "wall-bounded" if wdata is not None
                                                                           else "periodic", cores)}


def c1_leg(ctx, nsteps, cpu):
    """The 16^3, density-3 periodic fluid (a = 75, gamma = 4.5, kBT = 1,
    dt = 1e-3, spow = 0.25): 50 warm-up steps, then the median of 5 timed
    runs of nsteps steps; the CPU restatement (half-shell traversal) runs the
    same nsteps steps from the same state."""
    import numpy as np
    import torch

    import oracle as O
    from ctc_amd import udx as U

    L, rho, dt = (16, 16, 16), 3, 1e-3
    n = rho * 16 ** 3
    rng = np.random.default_rng(1)
    pp = np.zeros((n, 6), np.float32)
    pp[:, :3] = rng.uniform(-8, 8, (n, 3)).astype(np.float32)
    pp[:, 3:] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    pp[:, 3:] -= pp[:, 3:].mean(0)
    params = U.dpd_params(75.0, 4.5, 1.0, dt, 0.25)
    flu = U.Flu(ctx, L, n, params, dt, 1.0, (0.0, 0.0, 0.0))
    flu.graph(True)                      # step pairs replayed from a captured HIP graph
    d = torch.as_tensor(pp).cuda()
    flu.set_pp(d)
    flu.run(50)
    torch.cuda.synchronize()
    runs, pairs = [], []
    for _ in range(5):
        flu.pairs(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        flu.run(nsteps)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
        pairs.append(flu.pairs())
    flu.close()
    k = int(np.argsort(runs)[2])
    out = {"workload": "periodic 16^3, density 3 (12288 particles), a = 75, gamma = 4.5, kBT = 1, dt = 1e-3; "
                       "graph-replayed step pairs",
           "steps": nsteps, "steps_per_s": nsteps / runs[k], "pairs_per_s": pairs[k] / runs[k],
           "pairs_per_step": pairs[k] / nsteps, "ms_per_step": runs[k] / nsteps * 1e3,
           "runs_s": runs}
    if cpu:
        cores = cpu_threads()
        seeds = O.seeds_table(nsteps)
        ref = pp.copy()
        ncell = 16 ** 3
        t0 = time.perf_counter()
        cp = 0
        for s in range(nsteps):
            starts, counts, srt, _ = O.clist_build(L, ref)
            ff, npair = O.dpd_forces_half(L, srt, starts, counts, (75.0, 4.5, params.s, 0.25), seeds[s],
                                          nthreads=cores)
            ref = O.integrate(L, srt, ff, dt)
            cp += npair // 2
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = {"steps_per_s": nsteps / sec, "pairs_per_s": cp / sec, "cores": cores,
                               "kind": "port", "sample": "%d steps of the same box from the same state "
                               "(%d cells), half-shell traversal" % (nsteps, ncell)}
    return out


if __name__ == "__main__":
    main()
