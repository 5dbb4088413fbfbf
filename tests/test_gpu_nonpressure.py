"""Non-pressure forces and the CFL step size on the GPU against the oracle
(Simulation::computeNonPressureForces, Simulation.cpp:618-631, and
updateTimeStepSizeCFL, :544-616): Akinci 2013 and He 2014 surface tension
(with He 2014's boundary reaction on the body), Standard and XSPH viscosity, CFL with the step-size bookkeeping of TimeStepDiffDFSPH::step
(the step's starting size advects velocities and positions and advances the
time; the new size runs the pressure solve, the rigid bodies and the next step).

Tolerances: one step's accelerations within 1e-5 of their largest component
(FP32 sums in a different order), step sizes within 1e-5 relative (the bound is
a maximum over FP32 speeds), and over a 40-step rollout (steps of up to 5 ms)
with the GPU's iteration counts replayed in the oracle the rigid pose and the
fluid within 1e-4 (relative, m RMS) or three times the spread between the
oracle's own FP32 and FP64 runs, whichever is larger.  The exact adjoint
covers the viscosity (compared with the oracle's adjoint) and refuses surface
tension and the CFL step size.
"""
import numpy as np
import pytest

from difffr_amd import scenes
from difffr_amd.simulation import DfrError, Simulation
from oracle.pyoracle import OracleSim

pytestmark = pytest.mark.gpu

FORCES = dict(surface_tension_method=2, surface_tension=0.2, surface_tension_boundary=0.05,
              viscosity_method=1, viscosity=0.01, cfl_method=1, cfl_factor=0.5,
              cfl_min_time_step_size=1e-4, cfl_max_time_step_size=0.005)


def _scene():
    sc = scenes.submerged_body(body="sphere", steps=40)
    sc.params.update(FORCES)
    return sc


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("st,visc", [(2, 1), (2, 2), (3, 1)])
def test_one_step_accelerations_and_step_size(gpu_ready, st, visc):
    """Surface tension 2 = Akinci 2013 (normals compared), 3 = He 2014 (colour
    field and |grad c|^2 compared; its boundary term pushes the body)."""
    sc = _scene()
    sc.params["viscosity_method"] = visc
    sc.params["surface_tension_method"] = st
    print(f"[non-pressure one step] surface tension {st}, viscosity method {visc}, n={sc.n_fluid}", flush=True)
    g = Simulation(sc)
    o = OracleSim(sc, "f32_rigid64")
    g.step(1)
    it = g.stats()
    o.set_fixed_iterations([it["iterations_v"]], [it["iterations"]])
    o.step(1)
    ag, ao = g.field("accel"), o.field("accel")
    e_a = np.abs(ag - ao).max() / np.abs(ao - np.asarray(sc.params["gravity"])).max()
    if st == 2:
        ng, no = g.field("normal"), o.field("normal")
        e_n = np.abs(ng - no).max() / np.abs(no).max()
    else:
        e_n = max(np.abs(g.field(f) - o.field(f)).max() / np.abs(o.field(f)).max() for f in ("color", "gradC2"))
    print(f"  accel err {e_a:.2e} of the force scale, normal / colour err {e_n:.2e}, dt GPU {it['dt']:.7g} "
          f"oracle {o.field('dt'):.7g}", flush=True)
    assert e_a < 1e-5 and e_n < 1e-5
    assert it["dt"] == pytest.approx(o.field("dt"), rel=1e-5)
    assert it["dt"] != pytest.approx(sc.params["dt"], rel=1e-3)  # the CFL rule moved it
    g.close()


@pytest.mark.parametrize("st", [2, 3])
def test_rollout_with_forces_and_cfl(gpu_ready, st):
    sc = _scene()
    sc.params["surface_tension_method"] = st
    print(f"[non-pressure rollout] surface tension {st}, {sc.steps} steps, n={sc.n_fluid}", flush=True)
    g = Simulation(sc)
    o = OracleSim(sc, "f32_rigid64")
    o64 = OracleSim(sc, "f64")
    dts_g, dts_o = [], []
    for s in range(sc.steps):
        g.step(1)
        st = g.stats()
        for oo in (o, o64):
            oo.set_fixed_iterations([st["iterations_v"]], [st["iterations"]])
            oo.step(1)
        dts_g.append(st["dt"])
        dts_o.append(o.field("dt"))
    rg, ro, r64 = g.rigid(1), o.rigid(1), o64.rigid(1)
    xg, _ = g.fluid()
    xo, _ = o.fluid()
    x64, _ = o64.fluid()
    rms = lambda a, b: float(np.sqrt(((a - b) ** 2).sum(1).mean()))
    pose = lambda a, b: max(rel(a["x"], b["x"]), rel(a["q"], b["q"]))
    e, floor = pose(rg, ro), pose(ro, r64)
    print(f"  dt range {min(dts_g):.3g}..{max(dts_g):.3g}, max dt rel err "
          f"{np.max(np.abs(np.array(dts_g) / np.array(dts_o) - 1)):.1e}; time GPU {g.stats()['time']:.6f} "
          f"oracle {o.field('time'):.6f}; pose rel vs FP32 oracle {e:.1e} (FP32 vs FP64 oracle {floor:.1e}); "
          f"fluid rms {rms(xg, xo):.1e} m (oracle FP32 vs FP64 {rms(xo, x64):.1e} m)", flush=True)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-5)
    assert g.stats()["time"] == pytest.approx(o.field("time"), rel=1e-5)
    # within the precision floor of the step (its FP32 and FP64 runs' spread)
    assert e < max(1e-4, 3 * floor)
    assert rms(xg, xo) < max(1e-4, 3 * rms(xo, x64))
    g.close()


def test_he2014_boundary_reaction_matches_oracle(gpu_ready):
    """He 2014's boundary term and its reaction on a dynamic ball (fluid term
    off, fluid below rest density, no divergence solve: the body's force and
    torque are the reactions alone) against the oracle, within 1e-4 relative
    (FP32 terms summed in FP64 in a different order), and equal to
    -sum_i m (a_i - g) of the GPU's own accelerations."""
    r = 0.025
    fluid = scenes.block_lattice((-0.3, -0.3, -0.3), (13, 13, 13), 3 * r)
    fluid = fluid[np.linalg.norm(fluid, axis=1) > 0.12 + 2 * r]
    p = scenes.default_params(r, 1e-3)
    # zero gravity: the FP32 accelerations then hold the boundary terms without cancellation against g
    p.update(compute_gradient=0, surface_tension_method=3, surface_tension=0.0, surface_tension_boundary=0.5,
             enable_divergence_solver=0, gravity=(0.0, 0.0, 0.0))
    body = scenes.Body(scenes.sphere_surface(0.12, 2 * r), True, np.array([0.01, -0.02, 0.015]),
                       np.array([1.0, 0, 0, 0]), 800.0)
    sc = scenes.Scene("ball", p, fluid, np.zeros_like(fluid), [body], 1)
    print(f"[He 2014 boundary reaction] n={sc.n_fluid}", flush=True)
    g = Simulation(sc)
    o = OracleSim(sc, "f32_rigid64")
    g.step(1)
    it = g.stats()
    o.set_fixed_iterations([it["iterations_v"]], [it["iterations"]])
    o.step(1)
    rg, ro = g.rigid(0), o.rigid(0)
    m = 0.8 * (2 * r) ** 3 * 1000.0
    own = -m * g.field("accel").astype(np.float64).sum(0)
    print(f"  force GPU {rg['force']} oracle {ro['force']}; torque GPU {rg['torque']} oracle {ro['torque']}",
          flush=True)
    assert np.abs(ro["force"]).max() > 0 and np.abs(ro["torque"]).max() > 0
    assert rel(rg["force"], ro["force"]) < 1e-4 and rel(rg["torque"], ro["torque"]) < 1e-4
    assert rel(rg["force"], own) < 1e-4
    g.close()


def test_exact_adjoint_refuses_surface_tension_and_cfl(gpu_ready):
    for extra in (dict(viscosity_method=0, cfl_method=0), dict(surface_tension_method=0, viscosity_method=0)):
        sc = _scene()
        sc.params.update(extra)
        sim = Simulation(sc, params_override={"compute_gradient": 0})
        with pytest.raises(DfrError, match="non-pressure"):
            sim.adjoint_begin(0)
        sim.close()


@pytest.mark.parametrize("visc,nu", [(1, 0.01), (2, 0.2)])
def test_exact_adjoint_with_viscosity_matches_oracle(gpu_ready, visc, nu):
    """The reverse of the advection with Standard (the reference fluid's default,
    nu = 0.01) or XSPH viscosity against the FP64 oracle's adjoint (pinned by
    central differences on the CPU), the oracle replaying the GPU's iteration
    counts; within 1e-3 of the largest component (FP32 against FP64).  The
    inviscid scene's adjoint misses by more than 1e-2, so the force's terms are
    what brings the two together.  Every step's tape kept (k = 0), segments
    re-run with their tapes kept (k = 5) and steps re-run one at a time
    (segment memory 0) agree within 1e-5."""
    steps = 12
    sc = scenes.submerged_body(body="ellipsoid", steps=steps)
    sc.params.update(viscosity_method=visc, viscosity=nu)
    seed = np.zeros((2, 7))
    seed[1] = np.random.default_rng(7).normal(size=7)
    out = []
    for k, seg_mem in ((0, -1), (5, -1), (5, 0)):
        print(f"[viscosity adjoint] method {visc} nu {nu} k={k} segment memory={seg_mem} n={sc.n_fluid}", flush=True)
        sim = Simulation(sc, params_override={"compute_gradient": 0})
        sim.adjoint_segment_memory(seg_mem)
        sim.adjoint_begin(k)
        sim.step(steps)
        out.append(sim.adjoint_backward(seed)[1])
        iters = sim.adjoint_iterations()
        sim.close()
    o = OracleSim(sc, "f64")
    o.set_fixed_iterations(*iters)
    o.set_record(True)
    o.step(steps)
    ref = o.adjoint(seed)[1]
    sc0 = scenes.submerged_body(body="ellipsoid", steps=steps)
    o0 = OracleSim(sc0, "f64")
    o0.set_record(True)
    o0.step(steps)
    inviscid = o0.adjoint(seed)[1]
    scale = np.abs(ref).max()
    print("gpu     ", out[0], "\noracle  ", ref, "\ninviscid", inviscid, flush=True)
    assert np.all(np.isfinite(out[0]))
    np.testing.assert_allclose(out[0], ref, rtol=0, atol=1e-3 * scale)
    assert np.abs(inviscid - ref).max() > 1e-2 * scale
    for g in out[1:]:
        np.testing.assert_allclose(g, out[0], rtol=0, atol=1e-5 * scale)
