"""Pins the oracle's coupling, solver and gradient functions (CPU only).

Each check compares an oracle function with something it was not written
from, so that a plausible mistake (a dropped term, a wrong sign, index or
lever arm, a transposed product) fails:

* reference-mode pair Jacobians (computeGradient, TimeStepDiffDFSPH.cpp:1281-1482):
  central differences of the per-pair force and torque map evaluated by the
  oracle's own forward functions, modes 0 and 1;
* Akinci coupling (velocity update + addForce, :1123-1196, :1847-1901,
  BoundaryModel.h:46-58): Newton's third law for force and angular momentum
  in FP64;
* Jacobi solves (:654-881, :964-1059, :1698-1788): the warm-start rules and the
  stiffness bookkeeping from the recorded fields, and the reported density /
  divergence errors against a brute-force numpy recomputation;
* Akinci boundary volume (BoundaryModel_Akinci2012.cpp:257-284 via
  Simulation::updateBoundaryVolume, Simulation.cpp:831-902): closed form on a
  flat wall lattice, and which samples enter the sum;
* forward-mode chain rule (BoundaryModel_Akinci2012.cpp:543-889) under
  prescribed force/torque laws: central differences of the rigid update map.
"""
import math

import numpy as np
import pytest

from difffr_amd import scenes
from oracle import pyoracle as po


# ------------------------------------------------------------ helpers ----
def cubic_w(r, h):
    """Textbook cubic spline W (k = 8 / (pi h^3))."""
    k = 8.0 / (math.pi * h ** 3)
    q = np.linalg.norm(np.atleast_2d(r), axis=-1) / h
    return np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1),
                    np.where(q <= 1.0, 2 * k * (1 - q) ** 3, 0.0))


def cubic_gradw(r, h):
    """Textbook gradient of the cubic spline (l = 48 / (pi h^3))."""
    l = 48.0 / (math.pi * h ** 3)
    r = np.atleast_2d(r)
    rl = np.linalg.norm(r, axis=-1)
    q = rl / h
    gq = r / np.maximum(rl, 1e-300)[:, None] / h
    c = np.where(q <= 0.5, l * q * (3 * q - 2), np.where(q <= 1.0, -l * (1 - q) ** 2, 0.0))
    c = np.where(rl > 1e-5, c, 0.0)
    return c[:, None] * gq


def rotation(q):
    """R(q) for q = (w, x, y, z), the polynomial form (Eigen's toRotationMatrix)."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rotate_h(q, p):
    """q p q* for q = (w, x, y, z) without normalisation (homogeneous of degree 2).
    Equal to R(q) p on unit quaternions; GradientUtils' d(R(q)p)/dq is the
    derivative of this form (its w column is 2 (w p + u x p))."""
    w, u = q[0], np.asarray(q[1:])
    return (w * w - u @ u) * p + 2 * (u @ p) * u + 2 * w * np.cross(u, p)


def central(fun, x0, eps):
    """Central differences of a vector function: columns = d fun / d x0[k]."""
    x0 = np.asarray(x0, np.float64)
    cols = []
    for k in range(x0.size):
        e = np.zeros_like(x0)
        e[k] = eps
        cols.append((np.asarray(fun(x0 + e)) - np.asarray(fun(x0 - e))) / (2 * eps))
    return np.stack(cols, axis=-1)


def close(a, b, rtol):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale


# ------------------------------------------------------ pair Jacobians ----
DT_EXACT = 1.0 / 1024.0  # 1/dt and 1/dt^2 are integers: the unsigned truncation (:1266) is exact


def _pair_scene(kind="submerged"):
    """submerged: a half-submerged ellipsoid after 6 steps (the solver has
    brought rho* to about 1); ball: a compressed lattice around a sphere with
    random velocities (rho* > 1 and d > 0 both common) after the release window."""
    if kind == "submerged":
        sc = scenes.submerged_body(counts=(12, 8, 12), body="ellipsoid", v0=(0.3, -0.8, 0.2), w0=(1.5, -2.0, 1.0),
                                   steps=8, dt=DT_EXACT)
        n = 6
    else:
        sc = _ball_around_body(DT_EXACT)
        n = 2
    o = po.OracleSim(sc, "f64", {"compute_gradient": 0})
    o.step(n)
    o.stage(0)  # lists, densities and factors at the current positions
    return sc, o


def _pairs(o, sc, mode, want=6):
    """Fluid/sample pairs whose force term is active (b_i > 0, alpha_i < 0) and
    well inside the support."""
    body = 1
    rig = o.rigid(body)
    xb = o.body_field(body, "x")
    vb = o.body_field(body, "v")
    x, v = o.fluid()
    h = 4 * sc.params["radius"]
    lists = o.neighbor_lists(2, body)
    out = []
    for j, nb in enumerate(lists):
        for i in nb:
            d = np.linalg.norm(x[i] - xb[j])
            if not (0.25 * h < d < 0.85 * h):
                continue
            r = o.pair_eval(body, j, int(i), mode, xb[j], vb[j], v[i], rig["x"])
            if r["b"] > 1e-4 and r["alpha"] < 0:
                out.append((j, int(i)))
                break
        if len(out) >= want:
            break
    return out


@pytest.mark.parametrize("kind,mode", [("submerged", 1), ("ball", 1), ("ball", 0)])
def test_pair_jacobians_match_central_differences(kind, mode):
    """Each per-sample Jacobian is the derivative of the pair's force (torque)
    f = m_f k_i g_ij, k_i = b_i alpha_i c, in the sense the reference uses:
    F_x, T_x: explicit in the body origin (v_i held);
    F_v, F_w, T_v, T_w: through v_j, with v_i responding implicitly,
      v_i = v_i0 + dt dvc(v_i, v_j) -> (I - dt dvc/dv_i)^-1 (:1401-1404);
    F_q, T_q: the explicit part through x_j(q) plus the implicit part through
      v_j(q) = omega x R(q) x0_j + v_rb (:1425-1480), R(q) p taken as q p q*.
    Mode 0 keeps the reference's extra 1/rho0 on the density term of d rho*/d x_j
    (:1539-1546): its map replaces the change of the pair's density term
    V_j W_ij by that change over rho0 (same value at the base point, that
    derivative)."""
    sc, o = _pair_scene(kind)
    body = 1
    rig = o.rigid(body)
    xrb, q, vrb, w = rig["x"], rig["q"], rig["v"], rig["omega"]
    x, v = o.fluid()
    x0 = sc.bodies[body].x0
    dt = DT_EXACT
    rho0 = sc.params["rho0"]
    pairs = _pairs(o, sc, mode)
    assert len(pairs) >= 4, "scene has too few active pairs"
    for j, i in pairs:
        xj_c = rotation(q) @ x0[j] + xrb
        vj_c = np.cross(w, xj_c - xrb) + vrb
        vi_c = v[i]

        wij0 = o.pair_eval(body, j, i, mode, xj_c, vj_c, vi_c, xrb)["wij"]

        def ev(xj, vj, vi, xo):
            r = o.pair_eval(body, j, i, mode, xj, vj, vi, xo)
            if mode == 0:  # the reference's density term of d rho*/d x_j carries 1/rho0
                s = (r["b"] - (r["wij"] - wij0) * (1 - 1 / rho0)) / r["b"]
                return r["dvc"] * s, r["f"] * s, r["tau"] * s
            return r["dvc"], r["f"], r["tau"]

        dvc_c = ev(xj_c, vj_c, vi_c, xrb)[0]
        vi0 = vi_c - dt * dvc_c
        # dvc is affine in v_i on the active side of the clamp: exact slope
        A = central(lambda u: ev(xj_c, vj_c, u, xrb)[0], vi_c, 1e-3)

        def implicit(xj, vj, xo):
            d = ev(xj, vj, vi_c, xo)[0]
            vi = np.linalg.solve(np.eye(3) - dt * A, vi0 + dt * d - dt * A @ vi_c)
            _, f, t = ev(xj, vj, vi, xo)
            return np.concatenate([f, t])

        def explicit(xj, vj, xo):
            _, f, t = ev(xj, vj, vi_c, xo)
            return np.concatenate([f, t])

        eps = 1e-6
        Dx = central(lambda p: explicit(xj_c + (p - xrb), vj_c, p), xrb, eps)
        # affine in v_j and omega on the active side: larger steps, less rounding
        Dv = central(lambda p: implicit(xj_c, vj_c + (p - vrb), xrb), vrb, 1e-3)
        Dw = central(lambda p: implicit(xj_c, np.cross(p, xj_c - xrb) + vrb, xrb), w, 1e-3)
        Dq = (central(lambda p: explicit(rotate_h(p, x0[j]) + xrb, vj_c, xrb), q, eps)
              + central(lambda p: implicit(xj_c, np.cross(w, rotate_h(p, x0[j])) + vrb, xrb), q, eps))
        J = o.pair_jacobians(body, j, i, mode)
        fd = dict(Fx=Dx[:3], Tx=Dx[3:], Fv=Dv[:3], Tv=Dv[3:], Fw=Dw[:3], Tw=Dw[3:], Fq=Dq[:3], Tq=Dq[3:])
        for name, ref in fd.items():
            assert close(J[name], ref, 1e-6), (mode, j, i, name, J[name], ref)


def test_pair_jacobian_finite_difference_check_catches_a_dropped_term():
    """The check above is sharp: dropping the implicit v_i response (the inverse
    factor of F_v) is visible in the same comparison."""
    sc, o = _pair_scene()
    body, mode, dt = 1, 1, DT_EXACT
    rig = o.rigid(body)
    xrb, q, vrb, w = rig["x"], rig["q"], rig["v"], rig["omega"]
    x, v = o.fluid()
    x0 = sc.bodies[body].x0
    worst = 0.0
    for j, i in _pairs(o, sc, mode):
        xj_c = rotation(q) @ x0[j] + xrb
        vj_c = np.cross(w, xj_c - xrb) + vrb
        explicit = lambda p: o.pair_eval(body, j, i, mode, xj_c, vj_c + (p - vrb), v[i], xrb)["f"]
        Fv_explicit = central(explicit, vrb, 1e-6)
        J = o.pair_jacobians(body, j, i, mode)
        worst = max(worst, np.abs(J["Fv"] - Fv_explicit).max() / np.abs(J["Fv"]).max())
    assert worst > 1e-4, worst


# -------------------------------------------------- Newton's third law ----
def _ball_around_body(dt=1e-3, squeeze=0.92):
    """Compressed fluid lattice around a dynamic sphere, no wall in reach (the
    only static samples sit far away), random fluid velocities."""
    r = 0.01
    sp = 2 * r * squeeze
    fl = scenes.block_lattice((0.0, 0.0, 0.0), (18, 18, 18), sp)
    c = fl.mean(0)
    rad = 0.05
    fl = fl[np.linalg.norm(fl - c, axis=1) > rad + r]
    rng = np.random.default_rng(3)
    vel = rng.normal(scale=0.3, size=fl.shape)
    far = scenes.block_lattice((5.0, 5.0, 5.0), (3, 1, 3), 2 * r)
    bodies = [scenes.Body(far, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0, name="far"),
              scenes.Body(scenes.sphere_surface(rad, 2 * r), True, c, np.array([0.9, 0.1, -0.3, 0.2]) / np.linalg.norm([0.9, 0.1, -0.3, 0.2]),
                          500.0, np.array([0.2, -0.1, 0.3]), np.array([2.0, -1.0, 0.5]), name="sphere")]
    sc = scenes.Scene("ball", scenes.default_params(r, dt), fl, vel, bodies, 4)
    return sc


@pytest.mark.parametrize("which", [0, 1])
def test_velocity_update_conserves_momentum_with_the_body(which):
    """One Jacobi velocity update (0: pressure form, 1: divergence form): the
    fluid-fluid terms are antisymmetric in (i, j) and parallel to x_i - x_j, so
    the fluid's momentum and its angular momentum about the body origin change
    by exactly minus the impulse and angular impulse the coupling put on the
    body (FP64)."""
    sc = _ball_around_body()
    o = po.OracleSim(sc, "f64", {"compute_gradient": 0})
    o.step(2)  # release window: the body moves with (v0, omega0)
    x, _ = o.fluid()
    dv, ft = o.velocity_update_probe(which)
    m = 0.8 * (2 * sc.params["radius"]) ** 3 * sc.params["rho0"]
    dt = sc.params["dt"]
    F, T = ft[1, :3], ft[1, 3:]
    xrb = o.rigid(1)["x"]
    assert np.abs(F).max() > 1e-3, "coupling inactive in the test scene"
    P = m * dv.sum(0) / dt
    L = m * np.cross(x - xrb, dv).sum(0) / dt
    scale = m * np.abs(dv).sum() / dt
    assert np.abs(P + F).max() <= 1e-10 * scale, (P, F)
    assert np.abs(L + T).max() <= 1e-10 * scale * 0.1, (L, T)
    # the static samples far away receive nothing
    assert np.all(ft[0] == 0)


# -------------------------------------------------- boundary volume -------
def _wall(n, sp, y=0.0):
    a = (np.arange(n) - n // 2) * sp
    X, Z = np.meshgrid(a, a, indexing="ij")
    return np.stack([X.ravel(), np.full(X.size, y), Z.ravel()], axis=1)


def test_boundary_volume_flat_wall_closed_form_and_sample_sets():
    """V_b = 1 / (W0 + sum_k W(x_b - x_k)).  On a square lattice of spacing 2r
    (h = 4r) an interior sample sees 4 samples at 2r and 4 at 2 sqrt(2) r
    (those at 4r = h contribute W = 0).  Static samples of every static body
    enter the sum; a dynamic body's samples see only their own body."""
    r = 0.025
    sp, h = 2 * r, 4 * r
    k = 8.0 / (math.pi * h ** 3)
    w = lambda q: k * (6 * q ** 3 - 6 * q ** 2 + 1) if q <= 0.5 else 2 * k * (1 - q) ** 3
    iso = 1.0 / (k + 4 * w(0.5) + 4 * w(math.sqrt(2) / 2))
    n = 21
    wall = _wall(n, sp)
    c = (n * n) // 2
    far_fluid = np.array([[10.0, 10.0, 10.0]])
    ident = np.array([1.0, 0, 0, 0])

    def vol(bodies):
        sc = scenes.Scene("walls", scenes.default_params(r), far_fluid, np.zeros_like(far_fluid), bodies, 1)
        o = po.OracleSim(sc, "f64")
        return [o.body_field(b, "volume") for b in range(len(bodies))]

    V = vol([scenes.Body(wall, False, np.zeros(3), ident, 1000.0)])
    assert V[0][c] == pytest.approx(iso, rel=1e-13)
    # a second static wall 2r above: its samples enter (1 at 2r, 4 at 2 sqrt2 r, 4 at sqrt12 r)
    V = vol([scenes.Body(wall, False, np.zeros(3), ident, 1000.0),
             scenes.Body(wall, False, np.array([0.0, sp, 0.0]), ident, 1000.0)])
    both = 1.0 / (k + 4 * w(0.5) + 4 * w(math.sqrt(2) / 2) + w(0.5) + 4 * w(math.sqrt(2) / 2)
                  + 4 * w(math.sqrt(12) / 4))
    assert V[0][c] == pytest.approx(both, rel=1e-13)
    assert V[1][c] == pytest.approx(both, rel=1e-13)
    # a dynamic wall 2r above the static one: each keeps its own set only
    V = vol([scenes.Body(wall, False, np.zeros(3), ident, 1000.0),
             scenes.Body(wall, True, np.array([0.0, sp, 0.0]), ident, 500.0)])
    assert V[0][c] == pytest.approx(iso, rel=1e-13)
    assert V[1][c] == pytest.approx(iso, rel=1e-13)


# ------------------------------------------------------ Jacobi solves -----
def test_warm_starts_and_stiffness_bookkeeping():
    """Recorded fields of consecutive steps: the warm starts follow the
    reference's rules from the previous step's stored stiffness
    (warmstartPressureSolve :964-1059, warmstartDivergenceSolve :1698-1788),
    and the stored stiffness of a step is its warm start plus every
    iteration's increment, in stored units (kappa dt^2, kappa_v dt; :753-766,
    :866-879)."""
    sc = _ball_around_body(squeeze=0.8)
    o = po.OracleSim(sc, "f64", {"compute_gradient": 0})
    o.step(3)
    # stored stiffness beyond the clamps (-0.00025, -0.5) for some particles
    rng = np.random.default_rng(5)
    o.set_state("kappa", np.where(rng.random(o.n) < 0.4, -0.0004, o.field("kappa")))
    o.set_state("kappa_v", np.where(rng.random(o.n) < 0.05, -0.7, o.field("kappa_v")))
    o.set_record(True)
    o.step(4)
    iv, ip = o.recorded_iterations()
    dt = sc.params["dt"]
    for s in range(4):
        kPrev, kvPrev = o.tape_field(s, "kPrev"), o.tape_field(s, "kvPrev")
        rsW, kW = o.tape_field(s, "rsW"), o.tape_field(s, "kW")
        d0, kvW = o.tape_field(s, "d0"), o.tape_field(s, "kvW")
        assert (rsW > 1).sum() > 10 and (d0 > 0).sum() > 10
        if s == 0:
            assert (kPrev[rsW > 1] < -0.00025).sum() > 3 and (kvPrev[d0 > 0] < -0.5).sum() > 3
        np.testing.assert_allclose(kW, np.where(rsW > 1, 0.5 * np.maximum(kPrev, -0.00025) / dt ** 2, 0.0),
                                   rtol=1e-14, atol=0)
        np.testing.assert_allclose(kvW, np.where(d0 > 0, 0.5 * np.maximum(kvPrev, -0.5) / dt, 0.0),
                                   rtol=1e-14, atol=0)
        if s < 3:
            kp = kW + sum(o.tape_field(s, "kP", it) for it in range(ip[s]))
            kv = kvW + sum(o.tape_field(s, "kD", it) for it in range(iv[s]))
            np.testing.assert_allclose(o.tape_field(s + 1, "kPrev"), kp * dt ** 2, rtol=1e-12,
                                       atol=1e-15 * np.abs(kp).max() * dt ** 2)
            np.testing.assert_allclose(o.tape_field(s + 1, "kvPrev"), kv * dt, rtol=1e-12,
                                       atol=1e-15 * np.abs(kv).max() * dt)


def test_solver_errors_against_brute_force_recomputation():
    """A tank with static walls only: the reported average density error
    (mean rho0 (rho* - 1) after the pressure solve) and divergence error
    (mean rho0 d after the divergence solve, d = max(drho/dt, 0), zero below 20
    neighbours) equal an O(N^2) numpy recomputation from the recorded state,
    and both are within their tolerances (:654-768, :770-881, :1926-2041)."""
    sc = scenes.tank((10, 8, 10), 0.02, body_kind="none", steps=20, wall_height=0.4)
    o = po.OracleSim(sc, "f64", {"compute_gradient": 0})
    o.step(12)
    o.set_record(True)
    o.step(1)
    p = sc.params
    r, dt, rho0 = p["radius"], p["dt"], p["rho0"]
    h = 4 * r
    Vf = 0.8 * (2 * r) ** 3
    x = o.tape_field(0, "x").reshape(-1, 3)
    vdiv = o.tape_field(0, "vDiv").reshape(-1, 3)
    _, vend = o.fluid()
    xb = o.body_field(0, "x")
    Vb = o.body_field(0, "volume")
    d_ff = x[:, None, :] - x[None, :, :]
    d_fb = x[:, None, :] - xb[None, :, :]
    nb_ff = (np.sum(d_ff * d_ff, -1) < h * h) & ~np.eye(len(x), dtype=bool)
    nb_fb = np.sum(d_fb * d_fb, -1) < h * h
    W_ff = np.where(nb_ff, cubic_w(d_ff.reshape(-1, 3), h).reshape(nb_ff.shape), 0.0)
    W_fb = np.where(nb_fb, cubic_w(d_fb.reshape(-1, 3), h).reshape(nb_fb.shape), 0.0)
    rho = rho0 * (Vf * 8 / (math.pi * h ** 3) + Vf * W_ff.sum(1) + (W_fb * Vb[None, :]).sum(1))
    np.testing.assert_allclose(o.tape_field(0, "rho"), rho, rtol=1e-12)
    G_ff = np.where(nb_ff[..., None], cubic_gradw(d_ff.reshape(-1, 3), h).reshape(d_ff.shape), 0.0)
    G_fb = np.where(nb_fb[..., None], cubic_gradw(d_fb.reshape(-1, 3), h).reshape(d_fb.shape), 0.0)

    def delta(vel):  # walls do not move
        dvv = vel[:, None, :] - vel[None, :, :]
        return Vf * np.sum(dvv * G_ff, -1).sum(1) + np.sum(vel[:, None, :] * G_fb * Vb[None, :, None], -1).sum(1)

    st = o.stats()
    rs = np.maximum(rho / rho0 + dt * delta(vend), 1.0)
    err_p = np.mean(rho0 * rs - rho0)
    assert st["avg_density_err"] == pytest.approx(err_p, rel=1e-9, abs=1e-12)
    assert st["iterations"] < p["max_iterations"] and err_p <= p["max_error"] * 0.01 * rho0
    nn = nb_ff.sum(1) + nb_fb.sum(1)
    d = np.where(nn < 20, 0.0, np.maximum(delta(vdiv), 0.0))
    err_v = np.mean(rho0 * d)
    assert st["avg_div_err"] == pytest.approx(err_v, rel=1e-9, abs=1e-12)
    assert st["iterations_v"] < p["max_iterations_v"] and err_v <= (1 / dt) * p["max_error_v"] * 0.01 * rho0
    assert (nn < 20).any() and (d > 0).any()


# ------------------------------------------------------ chain rule --------
def _rigid_only_scene(dt):
    """A lone dynamic ellipsoid (anisotropic inertia, rotated), far from the
    only fluid particle and static samples."""
    r = 0.01
    far = scenes.block_lattice((5.0, 5.0, 5.0), (2, 1, 2), 2 * r)
    q0 = np.array([0.8, 0.3, -0.4, 0.33])
    q0 /= np.linalg.norm(q0)
    ell = scenes.Body(scenes.ellipsoid_surface((0.12, 0.05, 0.08), 2 * r), True, np.zeros(3), q0, 600.0,
                      np.array([0.4, -0.2, 0.7]), np.zeros(3), name="ellipsoid")
    fl = np.array([[-5.0, -5.0, -5.0]])
    return scenes.Scene("rigid_only", scenes.default_params(r, dt), fl, np.zeros_like(fl),
                        [scenes.Body(far, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0), ell], 1)


CHAIN_CASES = {
    # mode: which state columns (x, v, omega, q) the laws may use
    1: (False, True, True, False),
    0: (True, True, True, True),
    2: None,  # RigidGradOnly: constant force, no torque
}


@pytest.mark.parametrize("mode", [1, 0, 2, "manager"])
def test_chain_rule_under_force_laws_matches_central_differences(mode):
    """perform_chain_rule (modes 1, 0, 2) and the RigidBodyGradientManager chain
    (RigidBodyGradientManager.cpp:96-423) with the gyroscopic term, the
    quaternion update and normalisation, against central differences of the
    rigid update map in (v0, omega0) over 40 steps, the force/torque given by
    affine laws in the body state (Sim::ForceModel).  The map uses the frozen
    inverse inertia that carries the reference's '+' on the d(I^-1) term."""
    dt = 0.01
    sc = _rigid_only_scene(dt)
    manager = mode == "manager"
    if manager:
        mode = 1  # the manager uses the v and omega sums only, whatever the mode
    rng = np.random.default_rng(7 + mode + (10 if manager else 0))
    A = rng.normal(scale=0.3, size=(3, 13))
    B = rng.normal(scale=0.01, size=(3, 13))
    F0 = rng.normal(scale=1.0, size=3)
    T0 = rng.normal(scale=0.02, size=3)
    if mode == 2:
        A[:] = 0.0; B[:] = 0.0; T0[:] = 0.0
    else:
        use = CHAIN_CASES[mode]
        for c, on in zip(((0, 3), (3, 6), (6, 9), (9, 13)), use):
            if not on:
                A[:, c[0]:c[1]] = 0.0; B[:, c[0]:c[1]] = 0.0
    steps = 40
    o = po.OracleSim(sc, "f64", {"compute_gradient": 1, "gradient_mode": mode,
                                 "use_gradient_manager": int(manager)})
    v0 = sc.bodies[1].v0.copy()
    w0 = np.zeros(3)
    o.set_force_model(1, F0, T0, A, B, start=2, record_frozen=True)
    o.model_step(steps)
    J = o.jacobians(1)
    if manager:  # the body's own Jacobians stay at their reset value
        J0 = np.zeros(78)
        J0[[18, 22, 26, 45, 49, 53]] = 1.0
        np.testing.assert_array_equal(J, J0)
        J = o.gradient_manager(1)[0]
    base = o.rigid(1)
    assert np.abs(base["omega"]).max() > 0.1 or mode == 2

    def run(p):
        o.set_init_velocity(1, p[:3], p[3:])
        o.reset_bodies()
        o.set_force_model(1, F0, T0, A, B, start=2, use_frozen=True)
        o.model_step(steps)
        s = o.rigid(1)
        return np.concatenate([s["x"], s["v"], s["omega"], s["q"]])

    p0 = np.concatenate([v0, w0])
    np.testing.assert_allclose(run(p0)[:3], base["x"], rtol=0, atol=1e-14)
    D = central(run, p0, 1e-6)
    Jxv, Jxw, Jvv, Jvw, Jwv, Jww = (J[9 * k:9 * k + 9].reshape(3, 3) for k in range(6))
    Jqv, Jqw = J[54:66].reshape(4, 3), J[66:78].reshape(4, 3)
    pairs = dict(Jxv=(Jxv, D[0:3, 0:3]), Jxw=(Jxw, D[0:3, 3:6]), Jvv=(Jvv, D[3:6, 0:3]), Jvw=(Jvw, D[3:6, 3:6]),
                 Jwv=(Jwv, D[6:9, 0:3]), Jww=(Jww, D[6:9, 3:6]), Jqv=(Jqv, D[9:13, 0:3]), Jqw=(Jqw, D[9:13, 3:6]))
    for name, (a, b) in pairs.items():
        if np.abs(b).max() < 1e-12:
            assert np.abs(a).max() < 1e-10, (name, a)
            continue
        assert close(a, b, 1e-6), (mode, name, a, b)
