"""Throughput benchmark of the differentiable DFSPH fluid-rigid step on one MI355X.

Prints ONE JSON line.  The workload (default: the 1M-particle tank with one
sphere) is fixed before any timing:

  1. settle: the tank runs SETTLE_STEPS plain steps from its jittered lattice
     (untimed); that state is kept in device memory;
  2. drop: every dynamic body is placed just above the fluid surface at a
     seeded position, rotation, velocity and angular velocity (the held-out
     evaluation seed; the warm-start network never saw it).

From that state the timed region is a K-step rollout (neighbour search,
density/factor, divergence and pressure Jacobi solves to 0.1 % / 0.01 %,
Akinci coupling, rigid update) that keeps every step's tape in HBM
(`--checkpoint-every 0`; k >= 1 checkpoints every k steps and re-runs the steps
in the backward pass), followed by its exact adjoint: forward + adjoint
particle-steps per second, inputs resident in HBM.  It is timed twice from the
same state: with the reference's warm start (halved previous kappa) and with
the learned warm start, whose network is first trained (untimed) on pairs
captured from 64 other seeded drops from the settled state (three rounds: each
later one captures rollouts warm-started by the network so far).  `value` is the
learned run; both iterate to the same tolerance.  W warm-up steps of the same
window (recorded + reversed, every kernel class timed, untimed) precede each
timed run.  The reference's own gradient (DiffFR forward mode, gradient mode 1)
is timed over the same window on the GPU, and the CPU oracle runs both the
reference path and the exact adjoint on a bounded sample of the same state.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tank_1m|tank_250k|tank_4m]
                  [--checkpoint-every k] [--no-learned] [--no-cpu-baseline]
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
SETTLE_STEPS = 300      # plain steps from the lattice before the drop (part of the workload)
STATE_SLOT = 0          # device slot of the settled state


def progress(msg: str):
    """One line per finished phase on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="tank_1m", choices=["tank_1m", "tank_250k", "tank_4m"])
    ap.add_argument("--checkpoint-every", type=int, default=None,
                    help="0: keep every step's tape in HBM (no recomputation); k >= 1: checkpoint every k steps "
                         "and re-run the steps during the backward pass (default: 0, 20 for tank_4m)")
    ap.add_argument("--no-learned", action="store_true", help="measure the plain warm start only")
    ap.add_argument("--train-drops", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=25, help="captured steps per training drop")
    ap.add_argument("--train-iters", type=int, default=6000, help="Adam iterations per training round")
    ap.add_argument("--train-rounds", type=int, default=3)
    ap.add_argument("--train-pairs", type=int, default=4_000_000)
    ap.add_argument("--train-lr", type=float, default=2e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    if a.checkpoint_every is None:
        a.checkpoint_every = 20 if a.workload == "tank_4m" else 0
    return a


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
    calibrated on a copy kernel), committed under profiles/ with the command that
    measured them."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            d = json.load(f)
        return d.get("bytes_per_launch"), d.get("state", d.get("command"))
    return None, None


def cpu_info() -> dict:
    """Host CPU model, physical cores (unique (package, core) pairs) and the CPUs
    this process may run on."""
    model, pairs, logical = "", set(), 0
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and not model:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    pairs.add((phys, core))
                    phys = core = None
        if phys is not None and core is not None:
            pairs.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "physical_cores": len(pairs) or None, "logical_cpus": logical or None,
            "cpus_allowed": allowed}


def start_window(sim, sc, drop_seed: int):
    from difffr_amd.warmstart import start_drop
    start_drop(sim, sc, STATE_SLOT, drop_seed)


def train_network(args, sim, sc):
    """Untimed: capture (features, converged kappa) pairs from rollouts of
    `train_drops` seeded drops from the settled state and fit the warm-start
    network on the GPU, in `train_rounds` rounds (difffr_amd.warmstart).  Round
    1 captures rollouts warm-started the reference's way (halved previous
    kappa); each later round captures rollouts warm-started by the network so
    far and continues its training.  Every rollout's Jacobi solves iterate to
    the same tolerance, so every target is the stiffness the solver converged
    to (measured on the held-out drop over the 100 steps after it, 32-unit
    network: 2.17-2.25 pressure iterations after two rounds of 3000 or 6000
    Adam iterations, 2.12 after three of 6000)."""
    import torch
    from difffr_amd.warmstart import train_on_drops
    return train_on_drops(sim, sc, STATE_SLOT, drops=args.train_drops, steps=args.train_steps,
                          rounds=args.train_rounds, iters=args.train_iters, pairs=args.train_pairs,
                          lr=args.train_lr, seed=args.seed, sync=torch.cuda.synchronize, log=progress)


def timed_run(args, sim, sc, eval_seed: int, learned: bool, P=None, norm=None):
    """W warm-up steps of the window (recorded and reversed with every kernel
    class timed: names the dominant class), an untimed rollout that sizes the
    tape arena for K steps, then the timed K-step rollout + exact adjoint from
    the same state, bracketed by device synchronisation.  In the timed run only
    the dominant class carries events (HIP events on the library's stream)."""
    import torch
    from difffr_amd.simulation import AdjointRangeError
    k = args.checkpoint_every
    if learned:
        sim.mlp_set_params(P, norm)
    sim.mlp_enable(learned)
    seed = np.zeros((sim.n_bodies, 7))
    seed[-1] = np.random.default_rng(args.seed).normal(size=7)
    dominant, times = "adj_velocity_update", {}
    if args.warmup > 0:
        start_window(sim, sc, eval_seed)
        sim.reset_stats()
        sim.set_profiling(True)
        sim.adjoint_begin(k)
        sim.step(args.warmup)
        try:
            sim.adjoint_backward(seed)
        except AdjointRangeError:
            pass  # the reverse passes ran in full; only their result left the FP32 range
        sim.set_profiling(False)
        times = sim.kernel_times()
        dominant = max(times, key=lambda c: times[c][0] * times[c][1])
    if k == 0 and args.steps > args.warmup:
        start_window(sim, sc, eval_seed)  # tape arena for K steps (allocation stays outside the timed region)
        sim.adjoint_begin(0)
        sim.step(args.steps)
    start_window(sim, sc, eval_seed)
    sim.reset_stats()
    sim.set_profiling(True, classes=[dominant])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.adjoint_begin(k)
    sim.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    try:
        sim.adjoint_backward(seed)
        range_note = None
    except AdjointRangeError as e:
        range_note = str(e)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    start_grad = None if range_note else sim.adjoint_start_gradient()[-1]
    sim.set_profiling(False)
    st = sim.stats()
    live = sim.kernel_times()[dominant]
    nst = sim.neighbor_stats()
    nst.update(sim.halo_stats())
    steps_done = max(st["steps_since_reset"], 1)
    n = sc.n_fluid
    return {"elapsed": t2 - t0, "fwd": t1 - t0, "adj": t2 - t1, "value": n * args.steps / (t2 - t0),
            "forward_value": n * args.steps / (t1 - t0), "times": times, "nst": nst,
            "dominant": dominant, "live": live,
            "I_p": st["total_iterations"] / steps_done, "I_v": st["total_iterations_v"] / steps_done,
            "err": st["avg_density_err"],
            "start_grad_norm": None if start_grad is None else float(np.linalg.norm(start_grad)),
            "range_note": range_note}


def reference_mode_run(args, sim, sc, eval_seed: int):
    """DiffFR's own gradient (forward mode, gradient mode 1: pair Jacobians every
    Jacobi iteration and the per-step chain rule) over the same window, plain
    warm start: forward particle-steps/s on the GPU."""
    import torch
    sim.mlp_enable(False)
    start_window(sim, sc, eval_seed)
    sim.set_compute_gradient(True)
    sim.reset_stats()
    sim.set_profiling(True, classes=["gradient_pairs"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sim.set_profiling(False)
    sim.set_compute_gradient(False)
    st = sim.stats()
    g = sim.kernel_times()["gradient_pairs"]
    return {"forward_with_gradient_value": sc.n_fluid * args.steps / (t1 - t0),
            "ms_per_step": 1e3 * (t1 - t0) / args.steps,
            "mean_pressure_iterations": st["total_iterations"] / max(st["steps_since_reset"], 1),
            "gradient_pairs_avg_ms": g[0], "gradient_pairs_launches": g[1]}


def window_scene(sim, sc):
    """The current GPU state (fluid in particle-id order, bodies, stiffness
    fields, time) as a scene + state for the CPU oracle."""
    from difffr_amd import scenes
    x, v = sim.fluid()
    bodies = []
    rig = []
    for k, b in enumerate(sc.bodies):
        s = sim.rigid(k)
        rig.append(np.concatenate([s["x"], s["q"], s["v"], s["omega"]]))
        bodies.append(scenes.Body(b.x0, b.dynamic, s["x"].copy(), s["q"].copy(), b.density, name=b.name))
    state = {"kappa": sim.field("kappa").astype(np.float64), "kappa_v": sim.field("kappa_v").astype(np.float64),
             "time": np.array([sim.stats()["time"]]), "rigid": rig}
    return scenes.Scene(sc.name + "_window", dict(sc.params), x.astype(np.float64), v.astype(np.float64),
                        bodies, 0), state


def cpu_baseline(wsc, state, steps: int):
    """The CPU oracle (OpenMP, FP32 fluid) from the GPU window's start state:
    (1) the reference's path, forward + DiffFR's forward-mode gradient (gradient
    mode 1), and (2) forward + the exact FP64 adjoint; `steps` steps each after
    one warm-up step."""
    from oracle.pyoracle import OracleSim
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or (os.cpu_count() or 1))

    def make(params):
        o = OracleSim(wsc, "f32", params)
        for name in ("kappa", "kappa_v", "time"):
            o.set_state(name, state[name])
        for k, r in enumerate(state["rigid"]):
            o.set_state(f"rigid{k}", r)
        return o

    o = make({"compute_gradient": 1, "gradient_mode": 1})
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    ref = wsc.n_fluid * steps / (t1 - t0)
    del o
    o = make({"compute_gradient": 0})
    o.step(1)
    o.set_record(True)
    t2 = time.perf_counter()
    o.step(steps)
    t3 = time.perf_counter()
    seed = np.zeros((o.n_bodies, 7))
    seed[-1, :] = 1.0
    o.adjoint(seed)
    t4 = time.perf_counter()
    del o
    out = dict(value=ref, unit="fluid particle-steps/s", cores=threads, kind="port",
               sample=f"{steps} steps (after 1 warm-up step) of the GPU window's start state "
                      f"({wsc.n_fluid} fluid particles, settled tank + drop), CPU oracle FP32 forward + DiffFR "
                      f"forward-mode gradient (gradient mode 1), OpenMP",
               reference_path_s=t1 - t0,
               exact_adjoint_leg={"value": wsc.n_fluid * steps / (t4 - t2), "forward_s": t3 - t2,
                                  "adjoint_s": t4 - t3,
                                  "what": "CPU oracle FP32 forward + FP64 exact adjoint, same sample"})
    out.update(cpu_info())
    return out


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
    sc = build(seed=args.seed, steps=SETTLE_STEPS + args.steps)
    n = sc.n_fluid
    eval_seed = 10_000 + args.seed  # held out: training drops use seeds 1..train_drops
    # the exact adjoint replaces DiffFR's forward-mode gradient on the measured path
    sim = Simulation(sc, params_override={"compute_gradient": 0})
    sim.step(SETTLE_STEPS)
    sim.state_save(STATE_SLOT)
    progress(f"{sc.name}: {n} fluid particles settled ({SETTLE_STEPS} steps)")
    plain = timed_run(args, sim, sc, eval_seed, False)
    progress(f"plain warm start timed: {plain['value'] / 1e6:.1f}M particle-steps/s")
    learned_on = not args.no_learned
    if learned_on:
        P, norm, tinfo = train_network(args, sim, sc)
        progress(f"network trained ({len(tinfo['rounds'])} rounds, {tinfo['capture_s'] + tinfo['train_s']:.1f} s)")
        learned = timed_run(args, sim, sc, eval_seed, True, P, norm)
        progress(f"learned warm start timed: {learned['value'] / 1e6:.1f}M particle-steps/s")
        main_run = learned
    else:
        main_run = plain
    refmode = reference_mode_run(args, sim, sc, eval_seed)
    progress("reference-mode gradient timed")
    wsc = wstate = None
    if not args.no_cpu_baseline:
        start_window(sim, sc, eval_seed)
        wsc, wstate = window_scene(sim, sc)
    sim.adjoint_end()
    sim.close()

    times, nst = main_run["times"], main_run["nst"]
    dominant = main_run["dominant"]
    nbar = nst["mean_fluid"] + nst["mean_boundary"]
    own = own_bytes_per_particle()
    bytes_per_launch = n * (4 + 4 * nbar + own[dominant]) if dominant in own else None
    avg_ms = main_run["live"][0]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if bytes_per_launch else None
    traffic, traffic_state = load_traffic(dominant, args.workload)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                "traffic_measured_on": traffic_state,
                "kernel": dominant, "avg_launch_ms": avg_ms, "launches": main_run["live"][1],
                "algorithmic_bytes_per_launch": bytes_per_launch}
    summary = lambda r: {"fwd_adj_value": r["value"], "forward_value": r["forward_value"],
                         "ms_per_step": 1e3 * r["elapsed"] / args.steps,
                         "forward_ms_per_step": 1e3 * r["fwd"] / args.steps,
                         "adjoint_ms_per_step": 1e3 * r["adj"] / args.steps,
                         "mean_pressure_iterations": r["I_p"], "mean_divergence_iterations": r["I_v"],
                         "final_avg_density_err": r["err"],
                         "adjoint_start_state_grad_norm": r["start_grad_norm"],
                         "adjoint_range_error": r["range_note"]}
    line = {
        "metric": METRIC, "value": main_run["value"], "unit": "fluid particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * main_run["elapsed"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded jittered lattice settled for 300 steps, procedural rigid bodies dropped with a "
                "held-out seed; warm-start network trained on 64 other seeded drops in three rounds)",
        "config": {
            "workload": f"{sc.name}: {n} fluid particles + {sc.n_boundary} boundary samples, dt=1e-3, "
                        f"0.01% density / 0.1% divergence tolerance; {SETTLE_STEPS} settle steps, then the "
                        f"bodies dropped (seed {eval_seed}); timed window = the {args.steps} steps after the drop",
            "path": f"{'learned' if learned_on else 'plain'} kappa warm start; forward rollout + exact adjoint "
                    f"(every reverse pass of every step, seeded by dL/d(x, q) of the last body; "
                    f"adjoint_start_state_grad_norm = |dL/d(x, q, v, omega)| at the window start); "
                    + ("every step's tape kept in HBM (no recomputation)" if args.checkpoint_every == 0 else
                       f"checkpoint every {args.checkpoint_every} step(s), steps re-run with recording in the "
                       f"backward pass"),
            "mean_fluid_neighbors": nst["mean_fluid"], "mean_boundary_neighbors": nst["mean_boundary"],
            "largest_halo": nst["largest_halo"], "halo_lds_capacity": nst["capacity"],
            "plain": summary(plain),
            "reference_mode_gradient_gpu": refmode,
        },
        "roofline": roofline,
        "kernel_ms": {c: {"avg_ms": v[0], "launches": v[1]} for c, v in times.items()},
        "kernel_ms_source": "the W warm-up steps of the window + their adjoint, every class timed",
    }
    if learned_on:
        line["config"]["learned"] = summary(learned)
        line["config"]["learned_over_plain"] = learned["value"] / plain["value"]
        line["config"]["forward_learned_over_plain"] = learned["forward_value"] / plain["forward_value"]
        line["config"]["pressure_iterations_saved"] = plain["I_p"] - learned["I_p"]
        line["config"]["training"] = tinfo
    if wsc is not None:
        line["cpu_baseline"] = cpu_baseline(wsc, wstate, args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
