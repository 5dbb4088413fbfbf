"""Start-of-trajectory windows of beginStep (TimeStepDiffDFSPH.cpp:353-430) on
the GPU against the oracle: the passive-body ramp with uniform_acc_time as the
reference's scenes use it (0.05 s and 0.25 s) and the release mode
(useReleaseRigidBodyMode).

Tolerances as in test_gpu_parity.py (rigid pose 1e-4 relative after the
rollout; reference-mode Jacobians against the FP32-vs-FP64 oracle spread) and
test_gpu_adjoint.py (exact adjoint within 1e-3 of the FP64 oracle's largest
component, iteration counts replayed).
"""
import numpy as np
import pytest

from difffr_amd import scenes
from difffr_amd.simulation import Simulation
from oracle.pyoracle import OracleSim

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# 0.05 s and 0.25 s less half a step: the window's edge falls between two step
# times, so FP32 (GPU, reference) and FP64 (oracle) time sums agree on it
@pytest.mark.parametrize("uat,release,steps", [(0.0495, 0, 100), (0.2495, 0, 300), (0.0495, 1, 100)])
def test_window_rollout_matches_oracle(gpu_ready, uat, release, steps):
    sc = scenes.sphere_in_box(n_per_axis=12, steps=steps)
    over = {"uniform_acc_time": uat, "release_mode": release}
    print(f"[window rollout] uat={uat} release={release} steps={steps} n={sc.n_fluid}", flush=True)
    g = Simulation(sc, params_override=over)
    o32 = OracleSim(sc, "f32_rigid64", params_override=over)
    o64 = OracleSim(sc, "f64", params_override=over)
    for s in range(steps):
        g.step(1)
        if s == 2:  # inside the window
            r = g.rigid(1)
            f = 0.0 if release else min(2 * sc.params["dt"] / uat, 1.0)
            np.testing.assert_allclose(r["v"], f * sc.bodies[1].v0, rtol=1e-6, atol=1e-9)
    o32.step(steps)
    o64.step(steps)
    rg, r32, r64 = g.rigid(1), o32.rigid(1), o64.rigid(1)
    e32 = max(rel(rg["x"], r32["x"]), rel(rg["q"], r32["q"]))
    floor = max(rel(r32["x"], r64["x"]), rel(r32["q"], r64["q"]))
    J, J32, J64 = g.jacobians_raw(1), o32.jacobians(1), o64.jacobians(1)
    print(f"  pose vs f32 oracle {e32:.2e} (oracle f32-vs-f64 {floor:.2e}); J rel vs f32 {rel(J, J32):.2e}, "
          f"vs f64 {rel(J, J64):.2e}, floor {rel(J32, J64):.2e}", flush=True)
    assert e32 < 1e-4
    assert rel(J, J64) < max(5 * rel(J32, J64), 1e-3)
    # the body moved away from its start after the window
    assert np.linalg.norm(rg["x"] - sc.bodies[1].x_rb) > 1e-3


@pytest.mark.parametrize("release", [0, 1])
def test_window_adjoint_matches_oracle(gpu_ready, release):
    sc = scenes.submerged_body(body="ellipsoid", steps=15)
    sc.params["uniform_acc_time"] = 0.0045  # window edge between steps: t = 5 ms in, 6 ms out
    sc.params["release_mode"] = release
    seed = np.zeros((2, 7))
    seed[1] = np.random.default_rng(8).normal(size=7)
    print(f"[window adjoint] release={release}", flush=True)
    sim = Simulation(sc, params_override={"compute_gradient": 0})
    sim.adjoint_begin(5)
    sim.step(15)
    g = sim.adjoint_backward(seed)[1]
    iters = sim.adjoint_iterations()
    sim.close()
    o = OracleSim(sc, "f64")
    o.set_fixed_iterations(*iters)
    o.set_record(True)
    o.step(15)
    ref = o.adjoint(seed)[1]
    print("gpu   ", g, "\noracle", ref, flush=True)
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-3 * np.abs(ref).max())
