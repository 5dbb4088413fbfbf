"""Pins the oracle's exact adjoint (oracle/dfsph_adjoint.hpp) on the CPU.

* central finite differences of the FP64 oracle with the Jacobi iteration
  counts held at the recorded values (the map is then smooth wherever no clamp
  or mask flips inside the difference interval; over these horizons none does);
* the fluid-free closed form dx_N/dv0 = N dt I (two release-window steps
  included: both override v with v0, so every step moves the body by dt v0);
* FP32 tapes against FP64 tapes (the GPU computes in FP32);
* with Standard or XSPH viscosity in the advection, central differences again.
"""
import numpy as np
import pytest

from difffr_amd import scenes
from oracle import pyoracle as po


def _run(scene, variant, v0, w0, steps, fix=None, record=False):
    s = po.OracleSim(scene, variant)
    s.set_init_velocity(1, v0, w0)
    if fix is not None:
        s.set_fixed_iterations(*fix)
    if record:
        s.set_record(True)
    s.step(steps)
    r = s.rigid(1)
    return s, np.concatenate([r["x"], r["q"]])


def _adjoint_and_fd(scene, steps, seed, eps=1e-6):
    b = scene.bodies[1]
    s, _ = _run(scene, "f64", b.v0, b.w0, steps, record=True)
    fix = s.recorded_iterations()
    sd = np.zeros((len(scene.bodies), 7))
    sd[1] = seed
    g = s.adjoint(sd)[1]
    p0 = np.concatenate([b.v0, b.w0])
    fd = np.zeros(6)
    for k in range(6):
        p = p0.copy(); p[k] += eps
        m = p0.copy(); m[k] -= eps
        _, a = _run(scene, "f64", p[:3], p[3:], steps, fix)
        _, c = _run(scene, "f64", m[:3], m[3:], steps, fix)
        fd[k] = seed @ (a - c) / (2 * eps)
    return g, fd


@pytest.mark.parametrize("body", ["sphere", "ellipsoid"])
def test_adjoint_matches_central_differences(body):
    sc = scenes.submerged_body(body=body, steps=15)
    seed = np.random.default_rng(1).normal(size=7)
    g, fd = _adjoint_and_fd(sc, 15, seed)
    scale = np.abs(fd).max()
    assert scale > 1e-4  # the coupling is active: the gradient is far from the fluid-free one
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-6 * scale)


@pytest.mark.parametrize("release", [0, 1])
def test_adjoint_with_ramp_and_release_windows(release):
    """uniform_acc_time = 4.5 ms (TimeStepDiffDFSPH::beginStep :353-430): the
    passive-body ramp drives the body with (t / uat) (v0, omega0) and holds it
    animated through uat + dt; the release mode holds body 1 at rest through
    uat + dt and then sets (v0, omega0) at the start of every step.  Adjoint
    against central differences of the whole rollout."""
    sc = scenes.submerged_body(body="ellipsoid", steps=15)
    sc.params["uniform_acc_time"] = 0.0045  # window edge between steps: t = 5 ms in, 6 ms out
    sc.params["release_mode"] = release
    seed = np.random.default_rng(6).normal(size=7)
    g, fd = _adjoint_and_fd(sc, 15, seed)
    scale = np.abs(fd).max()
    assert scale > 1e-4
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-6 * scale)
    # inside the window the body is animated: step 3 starts at t = 2 dt, so the
    # ramp gives 0.4 (v0, omega0) and the release mode holds it at rest
    s, _ = _run(sc, "f64", sc.bodies[1].v0, sc.bodies[1].w0, 3)
    r = s.rigid(1)
    f = 0.0 if release else 2 * sc.params["dt"] / 0.0045
    np.testing.assert_allclose(r["v"], f * sc.bodies[1].v0, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(r["omega"], f * sc.bodies[1].w0, rtol=1e-12, atol=1e-15)


def test_adjoint_free_flight_closed_form():
    sc = scenes.submerged_body(steps=12)
    sc.fluid_x = sc.fluid_x[:0]
    sc.fluid_v = sc.fluid_v[:0]
    sc.bodies = sc.bodies[1:]  # the body alone
    sc.bodies.insert(0, scenes.Body(np.zeros((0, 3)), False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1.0))
    n, dt = 12, sc.params["dt"]
    b = sc.bodies[1]
    s, _ = _run(sc, "f64", b.v0, b.w0, n, record=True)
    for k in range(3):
        sd = np.zeros((2, 7)); sd[1, k] = 1.0
        g = s.adjoint(sd)[1]
        expect = np.zeros(6); expect[k] = n * dt
        np.testing.assert_allclose(g, expect, atol=1e-12)


def test_adjoint_fp32_tapes_track_fp64():
    sc = scenes.submerged_body(body="ellipsoid", steps=20)
    b = sc.bodies[1]
    seed = np.zeros((2, 7)); seed[1] = np.random.default_rng(2).normal(size=7)
    s64, _ = _run(sc, "f64", b.v0, b.w0, 20, record=True)
    s32, _ = _run(sc, "f32", b.v0, b.w0, 20, fix=s64.recorded_iterations(), record=True)
    g64 = s64.adjoint(seed)[1]
    g32 = s32.adjoint(seed)[1]
    np.testing.assert_allclose(g32, g64, rtol=0, atol=1e-3 * np.abs(g64).max())


def test_adjoint_after_impact_is_the_derivative_of_a_rough_map():
    """A body driven into the fluid (small tank, v0 = (0.5, -1.5, 0.3)), 100
    steps: the gradient has grown about 30x over the 20 steps since contact.
    Differences with eps = 1e-10 (no eps guard or neighbour-count switch flips
    inside the interval) reproduce the adjoint; with eps >= 1e-8 the secants are
    dominated by those switches and bear no relation to it.  The adjoint is the
    derivative of the discrete step; its growth is a property of that step."""
    sc = scenes.small_tank(seed=0, steps=100)
    b = sc.bodies[1]
    b.v0 = np.array([0.5, -1.5, 0.3])
    seed = np.random.default_rng(11).normal(size=7)
    s, _ = _run(sc, "f64", b.v0, b.w0, 100, record=True)
    fix = s.recorded_iterations()
    sd = np.zeros((len(sc.bodies), 7))
    sd[1] = seed
    g = s.adjoint(sd)[1]
    del s
    assert np.abs(g).max() > 1.0  # free flight alone would give about 0.14
    p0 = np.concatenate([b.v0, b.w0])

    def secant(k, eps):
        p = p0.copy(); p[k] += eps
        m = p0.copy(); m[k] -= eps
        _, a = _run(sc, "f64", p[:3], p[3:], 100, fix)
        _, c = _run(sc, "f64", m[:3], m[3:], 100, fix)
        return seed @ (a - c) / (2 * eps)

    fine = secant(1, 1e-10)
    print(f"adjoint {g[1]:.6f}, eps=1e-10 secant {fine:.6f}", flush=True)
    assert abs(fine - g[1]) < 2e-3 * abs(g[1])
    coarse = secant(1, 1e-8)
    print(f"eps=1e-8 secant {coarse:.6f}", flush=True)
    assert abs(coarse - g[1]) > 0.5 * abs(g[1])


@pytest.mark.parametrize("visc,nu", [(1, 0.01), (2, 0.2)])
def test_adjoint_with_viscosity_matches_central_differences(visc, nu):
    """Standard viscosity at the reference fluid's default coefficient 0.01
    (Viscosity_Standard.cpp:262-268) and XSPH (Viscosity_XSPH.cpp:62-68) in the
    advection of every step: the adjoint carries the force's dependence on the
    fluid velocities, positions and densities.  Over these 10 steps the
    viscosity moves the gradient by several percent of its largest component,
    so an adjoint without its terms misses the differences by far more than
    the tolerance (checked below against the inviscid scene's adjoint)."""
    sc = scenes.submerged_body(body="ellipsoid", steps=10)
    sc.params.update(viscosity_method=visc, viscosity=nu)
    seed = np.random.default_rng(7).normal(size=7)
    g, fd = _adjoint_and_fd(sc, 10, seed)
    scale = np.abs(fd).max()
    assert scale > 1e-4
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-6 * scale)
    sc0 = scenes.submerged_body(body="ellipsoid", steps=10)
    s0, _ = _run(sc0, "f64", sc0.bodies[1].v0, sc0.bodies[1].w0, 10, record=True)
    sd = np.zeros((len(sc0.bodies), 7))
    sd[1] = seed
    assert np.abs(s0.adjoint(sd)[1] - fd).max() > 1e-2 * scale


def test_adjoint_refuses_surface_tension_and_cfl():
    for extra in (dict(surface_tension_method=2), dict(cfl_method=1)):
        sc = scenes.submerged_body(body="sphere", steps=2)
        sc.params.update(extra)
        s, _ = _run(sc, "f64", sc.bodies[1].v0, sc.bodies[1].w0, 2, record=True)
        with pytest.raises(RuntimeError):
            s.adjoint(np.zeros((len(sc.bodies), 7)))
