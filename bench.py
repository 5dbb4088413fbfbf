"""Throughput benchmark of the differentiable DFSPH fluid-rigid step on one MI355X.

Prints ONE JSON line.  The timed region is a K-step rollout of the 1M-particle
tank (neighbour search, density/factor, divergence and pressure Jacobi solves
with Akinci coupling, rigid update, a checkpoint every `--checkpoint-every`
steps) followed by its exact adjoint (every step re-run once recording its
Jacobi fields, then the reverse kernels), i.e. forward + adjoint particle-steps
per second.  Inputs are resident in HBM; the rollout starts from a state
settled by W untimed steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--checkpoint-every k] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fluid particle-steps/s (fwd+adjoint) @1M, learned vs plain at 0.01% density err"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--workload", default="tank_1m", choices=["tank_1m", "tank_250k", "tank_4m"])
    ap.add_argument("--checkpoint-every", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--seed", type=int, default=0)
    return ap.parse_args()


def own_bytes_per_particle():
    """Own (non-gathered) bytes each gather kernel reads + writes per fluid particle,
    excluding the neighbour count (4) and index stream (4 per neighbour)."""
    return {
        # x(12) v(12) k(4) kappa(4) -> v(12) kappa(4)
        "velocity_update": 12 + 12 + 4 + 4 + 12 + 4,
        # x(12) v(12) rho(4) alpha(4) -> b(4) k(4)
        "density_recompute": 12 + 12 + 4 + 4 + 4 + 4,
        # x(12) v(12) kappa_v(4) -> rho(4) alpha(4) s(12) d(4) kappa_v(4)
        "density_factor": 12 + 12 + 4 + 4 + 4 + 12 + 4 + 4,
        # x(12) v(12) rho(4) kappa(4) -> b(4) kappa(4)
        "pressure_warmstart": 12 + 12 + 4 + 4 + 4 + 4,
        # x(12) k(4) b(4) alpha(4) kacc(4) vbar(12) -> xbar(12 r + 12 w) out(4) alpha_bar(4+4) rho_bar(4+4)
        "adj_velocity_update": 12 + 4 + 4 + 4 + 4 + 12 + 24 + 4 + 8 + 8,
        # x(12) v(12) dbar(4) -> vbar(12 r + 12 w) xbar(12 r + 12 w)
        "adj_density_change": 12 + 12 + 4 + 24 + 24,
        # x(12) alpha(4) alpha_bar(4) rho_bar(4) G(12) -> xbar(12 r + 12 w)
        "adj_density_factor": 12 + 4 + 4 + 4 + 12 + 24,
    }


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy kernel), committed under profiles/."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            d = json.load(f)
        return d.get("bytes_per_launch")
    return None


def cpu_baseline(scene, steps: int):
    """Forward + exact adjoint of the CPU oracle (FP32 forward as in the reference,
    OpenMP; FP64 scatter adjoint with per-thread accumulators) on `steps` steps of
    the same scene."""
    from oracle.pyoracle import OracleSim
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = OracleSim(scene, "f32", {"compute_gradient": 0})
    o.step(1)  # warm-up step (allocations, first neighbour search)
    o.set_record(True)
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    seed = np.zeros((o.n_bodies, 7))
    seed[-1, :] = 1.0
    o.adjoint(seed)
    t2 = time.perf_counter()
    return dict(value=scene.n_fluid * steps / (t2 - t0), unit="fluid particle-steps/s", cores=threads, kind="port",
                forward_s=t1 - t0, adjoint_s=t2 - t1,
                sample=f"{steps} steps of the same {scene.n_fluid}-particle scene after 1 warm-up step, "
                       f"forward + exact adjoint of the CPU oracle (FP32 forward, FP64 adjoint, OpenMP)")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from difffr_amd import scenes
    from difffr_amd.simulation import Simulation

    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    torch.cuda.init()
    build = {"tank_1m": scenes.tank_1m, "tank_250k": scenes.tank_250k, "tank_4m": scenes.tank_4m}[args.workload]
    sc = build(seed=args.seed, steps=args.warmup + args.steps)
    # the exact adjoint replaces DiffFR's forward-mode gradient on this path
    sim = Simulation(sc, params_override={"compute_gradient": 0})
    sim.step(args.warmup)
    sim.reset_stats()
    seed = np.zeros((sim.n_bodies, 7))
    seed[-1] = np.random.default_rng(args.seed).normal(size=7)

    # Timed region: K-step rollout with checkpoints + its exact adjoint, bracketed
    # by device synchronisation; kernel-class durations recorded with HIP events
    # on the library's stream.
    sim.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.adjoint_begin(args.checkpoint_every)
    sim.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    grad = sim.adjoint_backward(seed)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sim.set_profiling(False)
    elapsed = t2 - t0
    st = sim.stats()
    times = sim.kernel_times()
    nst = sim.neighbor_stats()
    sim.adjoint_end()
    assert np.all(np.isfinite(grad)), "non-finite adjoint"

    n = sc.n_fluid
    value = n * args.steps / elapsed
    per_kernel_total = {k: v[0] * v[1] for k, v in times.items()}
    dominant = max(per_kernel_total, key=per_kernel_total.get)
    nbar = nst["mean_fluid"] + nst["mean_boundary"]
    own = own_bytes_per_particle()
    if dominant in own:
        bytes_per_launch = n * (4 + 4 * nbar + own[dominant])
    else:
        bytes_per_launch = None
    avg_ms = times[dominant][0]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if bytes_per_launch else None
    traffic = load_traffic(dominant, args.workload)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                "kernel": dominant, "avg_launch_ms": avg_ms, "launches": times[dominant][1],
                "algorithmic_bytes_per_launch": bytes_per_launch}
    steps_done = max(st["steps_since_reset"], 1)
    line = {
        "metric": METRIC, "value": value, "unit": "fluid particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "forward_ms_per_step": 1e3 * (t1 - t0) / args.steps, "adjoint_ms_per_step": 1e3 * (t2 - t1) / args.steps,
        "forward_value": n * args.steps / (t1 - t0),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, procedural rigid body)",
        "config": {
            "workload": f"{sc.name}: {n} fluid particles + {sc.n_boundary} boundary samples, "
                        f"dt=1e-3, 0.01% density / 0.1% divergence tolerance",
            "path": f"plain solver (learned warm start not in this measurement), forward rollout with a "
                    f"checkpoint every {args.checkpoint_every} step(s) + exact adjoint (record-replay of every step "
                    f"+ reverse kernels) of dL/d(v0, omega0)",
            "mean_pressure_iterations": st["total_iterations"] / steps_done,
            "mean_divergence_iterations": st["total_iterations_v"] / steps_done,
            "final_avg_density_err": st["avg_density_err"],
            "mean_fluid_neighbors": nst["mean_fluid"], "mean_boundary_neighbors": nst["mean_boundary"],
        },
        "roofline": roofline,
        "kernel_ms": {k: {"avg_ms": v[0], "launches": v[1]} for k, v in times.items()},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
