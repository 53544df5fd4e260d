"""Benchmark of the DPD per-timestep hot path on one MI355X.

One step = cell list (counting sort + reorder), DPD forces (conservative +
dissipative + random with the step-seeded hash RNG), constant body force,
kick-drift and periodic wrap, for the 96^3, density-10 box with the channel
flow parameters (a = 2.5, gamma = 8, kBT = 0.0444, dt = 5e-4, body force 0.01
x).  Prints one JSON line.
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
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--L", type=int, default=96)
    ap.add_argument("--rho", type=float, default=10.0)
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=20, help="steps timed per kernel phase")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the path would shard by "
              "3-D domain decomposition, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ctc_amd import udx as U

    torch.cuda.set_device(0)
    L = (args.L, args.L, args.L)
    ncells = L[0] * L[1] * L[2]
    n = int(round(args.rho * ncells))
    a, g, kBT, dt, fx = 2.5, 8.0, 0.0444, 5e-4, 0.01
    params = U.dpd_params(a, g, kBT, dt, 0.25)

    # synthetic state: uniform positions, Maxwell velocities with zero net momentum
    gen = torch.Generator(device="cuda").manual_seed(1234)
    pp = torch.empty((n, 6), dtype=torch.float32, device="cuda")
    pp[:, :3] = (torch.rand((n, 3), generator=gen, device="cuda") - 0.5) * float(args.L)
    pp[:, :3].clamp_(-0.5 * args.L, 0.5 * args.L - 1e-4)
    v = torch.randn((n, 3), generator=gen, device="cuda", dtype=torch.float64) * np.sqrt(kBT)
    v -= v.mean(0)
    pp[:, 3:] = v.float()

    ctx = U.Context()
    flu = U.Flu(ctx, L, n, params, dt, 1.0, (fx, 0.0, 0.0))
    flu.set_pp(pp)

    # warm-up
    flu.run(args.warmup)
    torch.cuda.synchronize()

    # per-phase kernel times (HIP events on the bound stream, one step at a time)
    flu.timing(True)
    flu.run(args.profile_steps)
    t_phase = flu.timing(False)          # ms: clist, forces, integrate, sum
    torch.cuda.synchronize()

    # timed region
    flu.pairs(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    flu.run(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pairs = flu.pairs()
    sec = t1 - t0
    ms_step = 1e3 * sec / args.steps
    steps_s = args.steps / sec
    pairs_step = pairs / args.steps
    pairs_s = pairs / sec

    # roofline of the dominant kernel (the force gather): algorithmic bytes
    # per launch = N * 36 B (24 B particle in, 12 B force out)
    t_force = t_phase[1] * 1e-3
    force_bytes = n * 36.0
    force_gbs = force_bytes / t_force / 1e9
    flops = pairs_step * 100.0 + n * 13.5 * args.rho * 9.0
    step_bytes = n * 164.0 + ncells * 16.0

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
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform positions, Maxwell velocities, seed 1234)",
        "config": {
            "workload": "dpd_%d^3_rho%g_channel_params_periodic" % (args.L, args.rho),
            "box": list(L), "particles": n, "a": a, "gamma": g, "kBT": kBT, "dt": dt, "body_force_x": fx,
            "walls": False,
            "parallelism": "single GPU",
        },
        "phase_ms": {"clist": t_phase[0], "forces": t_phase[1], "integrate": t_phase[2]},
        "roofline": {
            "bound": "hbm", "kernel": "k_dpd (force gather)",
            "achieved": force_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": force_gbs / HBM_PEAK_GBS,
            "traffic": None,
        },
        "step_roofline": {
            "achieved": step_bytes / (ms_step * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_step": step_bytes,
        },
        "valu": {
            "achieved": flops / t_force / 1e12, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": flops / t_force / 1e12 / FP32_PEAK_TFLOPS, "flop_per_step": flops,
        },
    }

    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(pp, L, (a, g, params.s, 0.25), dt, fx, args.cpu_steps)
        cb = out["cpu_baseline"]
        out["cpu_baseline"]["gpu_over_cpu"] = pairs_s / cb["value"] if cb["value"] else None

    flu.close()
    print(json.dumps(out))


def cpu_baseline(pp_dev, L, params, dt, fx, nsteps):
    """The oracle (CPU restatement, OpenMP over particles) on the same state."""
    import numpy as np

    import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    pp = pp_dev.cpu().numpy()
    seeds = O.seeds_table(nsteps)
    t0 = time.perf_counter()
    pairs = 0
    for k in range(nsteps):
        pp, npair = O.flu_step(L, pp, params, seeds[k], dt, accel=(fx, 0.0, 0.0), nthreads=cores)
        pairs += npair
    sec = time.perf_counter() - t0
    return {"value": pairs / sec, "unit": "pair-interactions/s", "steps_per_s": nsteps / sec,
            "cores": cores, "kind": "port",
            "sample": "%d full step(s) of the same %dx%dx%d box (clist + forces + integrate)" % (nsteps, *L)}


if __name__ == "__main__":
    main()
