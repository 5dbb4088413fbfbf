"""Exact adjoint on the GPU (dfr_adjoint_*), through the C ABI.

The oracle's adjoint is pinned by FP64 central differences on the CPU
(test_adjoint_oracle.py); here the GPU reverse pass is compared with it on the
same scene, the oracle replaying the GPU run's Jacobi iteration counts.
Tolerance: 1e-3 of the gradient's largest component (FP32 bulk against FP64).
"""
import numpy as np
import pytest

from difffr_amd import scenes

pytestmark = pytest.mark.gpu

STEPS = 15


def _gpu_rollout(sc, steps, k=1, params=None):
    from difffr_amd.simulation import Simulation
    sim = Simulation(sc, params_override=params or {"compute_gradient": 0})
    sim.adjoint_begin(k)
    sim.step(steps)
    return sim


def _oracle_adjoint(sc, steps, seed, iters, variant="f64"):
    from oracle import pyoracle as po
    s = po.OracleSim(sc, variant)
    s.set_fixed_iterations(*iters)
    s.set_record(True)
    s.step(steps)
    return s.adjoint(seed)


@pytest.mark.parametrize("body", ["sphere", "ellipsoid"])
def test_adjoint_matches_oracle(gpu_ready, body):
    sc = scenes.submerged_body(body=body, steps=STEPS)
    seed = np.zeros((2, 7))
    seed[1] = np.random.default_rng(3).normal(size=7)
    print(f"[adjoint vs oracle] body={body} steps={STEPS} n={sc.n_fluid}", flush=True)
    sim = _gpu_rollout(sc, STEPS)
    g = sim.adjoint_backward(seed)[1]
    iters = sim.adjoint_iterations()
    ref = _oracle_adjoint(sc, STEPS, seed, iters)[1]
    scale = np.abs(ref).max()
    print("gpu   ", g, "\noracle", ref, flush=True)
    assert np.all(np.isfinite(g))
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-3 * scale)


def test_adjoint_checkpoint_intervals_agree(gpu_ready):
    sc = scenes.submerged_body(body="ellipsoid", steps=12)
    seed = np.zeros((2, 7))
    seed[1] = np.random.default_rng(4).normal(size=7)
    out = []
    # k = 0: every step's tape kept in HBM, no recomputation; k >= 1: each segment re-run
    # once with its tapes kept; seg_mem = 0: states kept, every step re-run twice
    for k, seg_mem in ((0, -1), (1, -1), (5, -1), (12, -1), (5, 0), (12, 0)):
        print(f"[checkpoint interval] k={k} segment memory={seg_mem}", flush=True)
        from difffr_amd.simulation import Simulation
        sim = Simulation(sc, params_override={"compute_gradient": 0})
        sim.adjoint_segment_memory(seg_mem)
        sim.adjoint_begin(k)
        sim.step(12)
        out.append(sim.adjoint_backward(seed)[1])
        sim.close()
    scale = np.abs(out[0]).max()
    for g in out[1:]:
        np.testing.assert_allclose(g, out[0], rtol=0, atol=1e-5 * scale)


def test_adjoint_restores_state_and_repeats(gpu_ready):
    sc = scenes.submerged_body(steps=8)
    sim = _gpu_rollout(sc, 8)
    before = sim.rigid(1)
    x_before, _ = sim.fluid()
    seed = np.zeros((2, 7))
    seed[1, :3] = 1.0
    g1 = sim.adjoint_backward(seed)
    g2 = sim.adjoint_backward(seed)
    after = sim.rigid(1)
    x_after, _ = sim.fluid()
    for key in ("x", "q", "v", "omega"):
        np.testing.assert_array_equal(before[key], after[key])
    np.testing.assert_array_equal(x_before, x_after)
    np.testing.assert_allclose(g1, g2, rtol=0, atol=1e-9 * np.abs(g1).max())
    # stepping on after backward continues the rollout
    sim.step(2)
    assert sim.get_step_count() == 10


def test_adjoint_free_flight_closed_form(gpu_ready):
    sc = scenes.submerged_body(steps=10)
    sc.fluid_x = sc.fluid_x[:0]
    sc.fluid_v = sc.fluid_v[:0]
    sim = _gpu_rollout(sc, 10)
    dt = float(np.float32(sc.params["dt"]))  # the rigid update runs with the FP32 step size widened to FP64
    for k in range(3):
        seed = np.zeros((2, 7))
        seed[1, k] = 1.0
        g = sim.adjoint_backward(seed)[1]
        expect = np.zeros(6)
        expect[k] = 10 * dt
        np.testing.assert_allclose(g, expect, atol=1e-12)


def test_autograd_function(gpu_ready):
    import torch
    from difffr_amd.autograd import rigid_rollout
    from difffr_amd.simulation import Simulation
    sc = scenes.submerged_body(body="ellipsoid", steps=10)
    sim = Simulation(sc, params_override={"compute_gradient": 0})
    p = torch.tensor(np.concatenate([sc.bodies[1].v0, sc.bodies[1].w0]), dtype=torch.float64, requires_grad=True)
    target = torch.tensor([0.2, 0.15, 0.17], dtype=torch.float64)
    out = rigid_rollout(p, sim, 1, 10)
    loss = ((out[:3] - target) ** 2).sum() + (1.0 - out[3]) ** 2
    loss.backward()
    g = p.grad.detach().numpy().copy()
    # the same gradient straight from the C ABI
    seed = np.zeros((2, 7))
    o = out.detach().numpy()
    seed[1, :3] = 2.0 * (o[:3] - target.numpy())
    seed[1, 3] = -2.0 * (1.0 - o[3])
    ref = sim.adjoint_backward(seed)[1]
    np.testing.assert_allclose(g, ref, rtol=1e-7, atol=1e-15)
    # a small step against the gradient lowers the loss
    with torch.no_grad():
        q = p - 1e-2 * p.grad / p.grad.norm()
    out2 = rigid_rollout(q, sim, 1, 10)
    loss2 = ((out2[:3] - target) ** 2).sum() + (1.0 - out2[3]) ** 2
    assert loss2.item() < loss.item()


def _three_body_scene(steps):
    """A 26 x 10 x 26 tank (spacing 0.02) with a sphere, a box and an ellipsoid
    half submerged (the fluid inside each removed), seeded release velocities."""
    sc = scenes.tank((26, 10, 26), 0.02, seed=2, steps=steps, wall_height=0.5, name="three_bodies",
                     body_specs=[("sphere", 0.05, 500.0), ("box", (0.1, 0.05, 0.08), 500.0),
                                 ("ellipsoid", (0.07, 0.035, 0.05), 600.0)])
    r = sc.params["radius"]
    top = float(sc.fluid_x[:, 1].max())
    keep = np.ones(sc.n_fluid, bool)
    for b in sc.bodies[1:]:
        b.x_rb = b.x_rb.copy()
        b.x_rb[1] = top
        b.v0 = 0.5 * b.v0
        ext = np.abs(b.x0).max(axis=0) + 2 * r
        keep &= ~np.all(np.abs(sc.fluid_x - b.x_rb) < ext, axis=1)
    sc.fluid_x = sc.fluid_x[keep]
    sc.fluid_v = sc.fluid_v[keep]
    return sc


def test_three_bodies_adjoint_matches_oracle(gpu_ready):
    sc = _three_body_scene(12)
    nb = len(sc.bodies)
    seed = np.random.default_rng(8).normal(size=(nb, 7))
    seed[0] = 0.0
    print(f"[three bodies] nf={sc.n_fluid} samples={[b.x0.shape[0] for b in sc.bodies]}", flush=True)
    sim = _gpu_rollout(sc, 12)
    g = sim.adjoint_backward(seed)
    iters = sim.adjoint_iterations()
    ref = _oracle_adjoint(sc, 12, seed, iters)
    print("gpu\n", g[1:], "\noracle\n", ref[1:], flush=True)
    for k in range(1, nb):
        scale = np.abs(ref[k]).max()
        assert scale > 0
        np.testing.assert_allclose(g[k], ref[k], rtol=0, atol=1e-3 * scale)
    # every body's forward state after the rollout against the oracle's, same iteration counts
    from oracle import pyoracle as po
    o = po.OracleSim(sc, "f32_rigid64")
    o.set_fixed_iterations(*iters)
    o.step(12)
    for k in range(1, nb):
        a, b = sim.rigid(k), o.rigid(k)
        np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-6)
        np.testing.assert_allclose(a["q"], b["q"], rtol=0, atol=1e-5)
