"""Training protocol of the learned stiffness warm start on seeded drops.

A settled tank is kept in a device state slot; each training rollout restores
it, drops every dynamic body with its own seed (drop_bodies) and records
(features, converged (kappa_v, kappa)) pairs while the plain solver runs to
its tolerance (dfr_mlp_capture).  The network is then fitted on the GPU
(dfr_mlp_train: bf16-MFMA forward/backward, FP32 Adam).  Later rounds capture
rollouts warm-started by the network of the previous round, so the targets
are the stiffness the solver converges to when it starts from the network's
own predictions.  Evaluation drops use seeds outside 1..drops.
"""
from __future__ import annotations

import time

import numpy as np


def drop_bodies(sim, sc, seed: int):
    """Places every dynamic body just above the fluid surface of `sc`: seeded
    position across the tank (bodies spread along x), rotation, velocity
    (v_x, v_z in [-2, 2], v_y in [-2, 0] m/s) and angular velocity in [-3, 3]^3."""
    rng = np.random.default_rng(seed)
    r = float(sc.params["radius"])
    top = float(sc.fluid_x[:, 1].max())
    walls = sc.bodies[0].x0
    lo, hi = walls.min(0), walls.max(0)
    dyn = [k for k, b in enumerate(sc.bodies) if b.dynamic]
    for j, k in enumerate(dyn):
        b = sc.bodies[k]
        rad = float(np.linalg.norm(b.x0, axis=1).max())
        margin = rad + 4 * r
        x = lo[0] + (hi[0] - lo[0]) * (j + 0.5 + rng.uniform(-0.3, 0.3)) / len(dyn)
        z = lo[2] + (hi[2] - lo[2]) * rng.uniform(0.3, 0.7)
        pos = np.array([np.clip(x, lo[0] + margin, hi[0] - margin), top + rad + 2 * r,
                        np.clip(z, lo[2] + margin, hi[2] - margin)])
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = np.array([rng.uniform(-2, 2), rng.uniform(-2, 0), rng.uniform(-2, 2)])
        w = rng.uniform(-3, 3, size=3)
        sim.set_rigid(k, pos, q, v, w)


def start_drop(sim, sc, slot: int, seed: int):
    """Restores the settled state kept in `slot` and drops the bodies with `seed`."""
    sim.state_load(slot)
    drop_bodies(sim, sc, seed)


def train_on_drops(sim, sc, slot: int, drops: int = 64, steps: int = 25, rounds: int = 2,
                   iters: int = 3000, pairs: int = 4_000_000, lr: float = 2e-3, batch: int = 16384,
                   seed: int = 0, sync=None, log=None):
    """Captures pairs from `drops` seeded drops (seeds 1..drops, `steps` steps
    each) from the state in `slot` and fits the network, `rounds` times (round
    1 warm-started the reference's way, each later round by the network so
    far; the parameters carry over between rounds).  `pairs` bounds the
    dataset: each step keeps a seeded fraction of its particles.  Leaves the
    network disabled and returns (params, norm, info); `sync` (e.g.
    torch.cuda.synchronize) makes the reported capture/train times exact;
    `log(message)` is called after each round."""
    sync = sync or (lambda: None)
    per_drop = max(sc.n_fluid * steps, 1)
    fraction = min(1.0, pairs / (drops * per_drop))
    rounds_info = []
    t_capture = t_train = 0.0
    sim.mlp_enable(False)
    for r in range(rounds):
        t0 = time.perf_counter()
        for d in range(drops):
            start_drop(sim, sc, slot, 1 + d)
            sim.mlp_capture(capacity=pairs, fraction=fraction, seed=r * 1000 + d)  # appends after drop 1
            sim.step(steps)
        sim.mlp_capture(0)
        n_pairs = sim.mlp_dataset()[0].shape[0]
        sync()
        t1 = time.perf_counter()
        losses = sim.mlp_train(iters, batch=batch, lr=lr, seed=seed + r)
        sync()
        t2 = time.perf_counter()
        t_capture += t1 - t0
        t_train += t2 - t1
        rounds_info.append({"capture_warm_start": "reference" if r == 0 else "network of the previous round",
                            "pairs": int(n_pairs), "loss_first": float(losses[:50].mean()),
                            "loss_last": float(losses[-50:].mean())})
        if log:
            log(f"training round {r + 1}: {n_pairs} pairs, loss {rounds_info[-1]['loss_first']:.3f} -> "
                f"{rounds_info[-1]['loss_last']:.3f}")
        sim.mlp_enable(True)
    sim.mlp_enable(False)
    P, norm = sim.mlp_get_params()
    info = {"training_drops": f"seeds 1..{drops}", "steps_per_drop": steps, "capture_fraction": fraction,
            "adam_iterations_per_round": iters, "learning_rate": lr, "batch": batch, "rounds": rounds_info,
            "capture_s": t_capture, "train_s": t_train}
    return P, norm, info
