"""Pinning of the oracle's rigid-rigid contact (the penalty form of
RigidContactSolver, RigidContactSolver.cpp:140-170, :307-366, :419-555; the
body updates of Dynamic3dRigidBody.h:113-142, 228-233), CPU only:

* rest values: inside a flat sample lattice (every neighbour with a complete
  neighbourhood) the rest density vol0 (W(0) + sum W) is exactly gamma;
* two mirror-image boxes pressed into each other (no gravity, no friction)
  are pushed apart with equal and opposite momentum and no spin;
* bodies out of each other's reach feel gravity only; the contact response is
  linear in the penalty stiffness;
* two small spheres in contact with friction and spin: the samples' rest
  values, the contact force and torque sums and the bodies' velocities after
  the step against a direct evaluation of the definitions (including the
  per-call gyroscopic term of addTorque and the later updateBoundaryForces
  calls).
"""
import numpy as np
import pytest

from difffr_amd import scenes
from oracle import pyoracle as po
from oracle.pyoracle import OracleSim

R_P = 0.015  # particle radius


def _params(**kw):
    p = scenes.default_params(R_P, 1e-3)
    p.update(gravity=(0.0, 0.0, 0.0), compute_gradient=0, use_rigid_contact=1, rigid_contact_beta=2000.0)
    p.update(kw)
    return p


def _scene(p, bodies):
    empty = np.zeros((0, 3))
    return scenes.Scene("contact", p, empty, empty.copy(), bodies, 1)


def _body(x0, xr, dynamic=True, v0=(0.0, 0.0, 0.0), w0=(0.0, 0.0, 0.0), density=800.0):
    return scenes.Body(np.asarray(x0, np.float64), dynamic, np.asarray(xr, np.float64), np.array([1.0, 0, 0, 0]),
                       density, np.asarray(v0, np.float64), np.asarray(w0, np.float64))


def test_rest_density_is_gamma_inside_a_lattice():
    sp = 2 * R_P
    plate = scenes.block_lattice((0.0, 0.0, 0.0), (15, 1, 15), sp)
    far = scenes.box_surface((5.0, 5.0, 5.0), (5.2, 5.2, 5.2), sp)
    for gamma in (0.7, 0.5):
        o = OracleSim(_scene(_params(rigid_contact_gamma=gamma), [_body(plate, [0, 0, 0], False),
                                                                   _body(far, [0, 0, 0], False)]), "f64")
        d0 = o.body_field(0, "contact_density0")
        idx = np.rint((plate - plate.min(0)) / sp).astype(int)
        deep = np.all((idx[:, [0, 2]] >= 4) & (idx[:, [0, 2]] <= 10), axis=1)
        assert deep.sum() == 49
        np.testing.assert_allclose(d0[deep], gamma, rtol=1e-12)
        assert np.abs(d0[~deep] - gamma).max() > 1e-3


def _boxes(gap, **kw):
    sp = 2 * R_P
    box = scenes.box_surface((-0.06, -0.06, -0.06), (0.06, 0.06, 0.06), sp)
    box -= box.mean(0)
    c = 0.5 * (box[:, 0].max() - box[:, 0].min()) + 0.5 * gap
    return _scene(_params(**kw), [_body(box, [-c, 0, 0]), _body(box * np.array([-1.0, 1.0, 1.0]), [c, 0, 0])])


def test_mirror_boxes_push_apart_with_equal_momentum():
    sc = _boxes(gap=0.4 * R_P)  # faces 0.4 r apart: well inside the kernel radius
    o = OracleSim(sc, "f64")
    m = [o.mass_props(k)["mass"] for k in range(2)]
    o.step(1)
    r0, r1 = o.rigid(0), o.rigid(1)
    assert r0["v"][0] < -1e-3 and r1["v"][0] > 1e-3
    p = m[0] * r0["v"] + m[1] * r1["v"]
    assert np.abs(p).max() <= 1e-12 * m[0] * np.abs(r0["v"]).max()
    np.testing.assert_allclose(r0["v"][1:], 0.0, atol=1e-12 * abs(r0["v"][0]))
    np.testing.assert_allclose(r0["omega"], 0.0, atol=1e-9)
    f0 = o.body_field(0, "contact_force")
    np.testing.assert_allclose(f0, -o.body_field(1, "contact_force"), rtol=0, atol=1e-12 * np.abs(f0).max())


def test_out_of_reach_gravity_only_and_linear_in_stiffness():
    sc = _boxes(gap=3 * 4 * R_P, gravity=(0.0, -9.81, 0.0))
    o = OracleSim(sc, "f64")
    o.step(1)
    for k in range(2):
        np.testing.assert_allclose(o.rigid(k)["v"], [0.0, -9.81e-3, 0.0], rtol=1e-14)
        assert np.all(o.body_field(k, "contact_force") == 0.0)
    dv = []
    for beta in (1000.0, 2000.0):
        o = OracleSim(_boxes(gap=0.4 * R_P, rigid_contact_beta=beta, gyroscopic=0), "f64")
        o.step(1)
        dv.append(np.concatenate([o.rigid(0)["v"], o.rigid(1)["v"]]))
    np.testing.assert_allclose(dv[1], 2 * dv[0], rtol=1e-12, atol=1e-15)


def _W(hc, r):
    W, _, _ = po.kernel_eval(hc, np.atleast_2d(r))
    return W


@pytest.mark.parametrize("gyro", [1, 0])
def test_two_spheres_against_direct_evaluation(gyro):
    """RigidContactSolver's penalty step evaluated sample by sample from its
    definitions (own-body rest values, contact density, n, v_rel divided by the
    own-body weight sum, Coulomb friction along normalized v_rel, one
    addForce/addTorque per contact sample, then one updateBoundaryForces call
    per body).  The contact kernel's radius is the search radius (the default
    factor 4): with a smaller one, a sample whose other-body neighbours all lie
    between the two radii has density / density0 = 1 up to rounding, and
    whether it calls addTorque (adding one more gyroscopic term) depends on
    the summation order; the test asserts no contact sample is that close."""
    sp = 2 * R_P
    ball = scenes.sphere_surface(0.05, sp)
    ball = ball - ball.mean(0)
    beta, gamma, mu, fac = 3000.0, 0.7, 0.4, 4.0
    g = np.array([0.0, -9.81, 0.0])
    dt = 1e-3
    p = _params(rigid_contact_beta=beta, rigid_contact_gamma=gamma, rigid_contact_friction=mu,
                rigid_contact_support_radius_factor=fac, gyroscopic=gyro, gravity=tuple(g))
    xa, xb = np.array([0.0, 0.0, 0.0]), np.array([0.093, 0.012, -0.008])
    va, wa = np.array([0.3, -0.1, 0.2]), np.array([2.0, -1.0, 3.0])
    vb, wb = np.array([-0.2, 0.1, 0.0]), np.array([-1.0, 0.5, 0.7])
    sc = _scene(p, [_body(ball, xa, v0=va, w0=wa), _body(ball, xb, v0=vb, w0=wb, density=600.0)])
    o = OracleSim(sc, "f64")
    X = [o.body_field(k, "x") for k in range(2)]
    XR = [o.rigid(k)["x"] for k in range(2)]
    MP = [o.mass_props(k) for k in range(2)]
    o.step(1)

    h, hc = 4 * R_P, fac * R_P
    W0 = _W(hc, np.zeros(3))[0]
    pts = np.concatenate(X)
    body = np.repeat([0, 1], [len(X[0]), len(X[1])])
    n = len(pts)
    d = pts[:, None, :] - pts[None, :, :]
    nb = (np.einsum("ijk,ijk->ij", d, d) < h * h) & ~np.eye(n, dtype=bool)
    Wm = _W(hc, d.reshape(-1, 3)).reshape(n, n) * nb
    own = body[:, None] == body[None, :]
    vol0 = gamma / (W0 + (Wm * own).sum(1))
    dens0 = vol0 * W0 + (Wm * own) @ vol0
    np.testing.assert_allclose(np.concatenate([o.body_field(k, "contact_vol0") for k in range(2)]), vol0, rtol=1e-12)
    np.testing.assert_allclose(np.concatenate([o.body_field(k, "contact_density0") for k in range(2)]), dens0,
                               rtol=1e-12)
    V = [va + g * dt, vb + g * dt]
    Wb = [wa, wb]
    vel = np.array([V[body[i]] + np.cross(Wb[body[i]], pts[i] - XR[body[i]]) for i in range(n)])
    v, w = [V[0].copy(), V[1].copy()], [wa.copy(), wb.copy()]
    F = [np.zeros(3), np.zeros(3)]
    T = [np.zeros(3), np.zeros(3)]
    contacts = 0
    for i in range(n):
        k = body[i]
        other = nb[i] & ~own[i]
        if not other.any():
            continue
        ratio = (vol0[i] * W0 + Wm[i] @ vol0) / dens0[i]
        assert abs(ratio - 1) > 1e-9
        if not ratio > 1:
            continue
        contacts += 1
        sw = (Wm[i] * own[i]).sum()
        nrm = pts[i] - (Wm[i] * own[i]) @ pts / sw
        vrel = vel[i] - (Wm[i] * other) @ vel / sw
        fn = -beta * (ratio - 1) * nrm
        f = fn - mu * np.linalg.norm(fn) * vrel / np.linalg.norm(vrel)
        tau = np.cross(pts[i] - XR[k], f)
        F[k] += f
        T[k] += tau
        v[k] += f / MP[k]["mass"] * dt
        rhs = tau + (np.cross(MP[k]["I"] @ w[k], w[k]) if gyro else 0.0)
        w[k] += MP[k]["Iinv"] @ rhs * dt
    assert contacts >= 4
    for _ in range(2):  # updateBoundaryForces, once per body: no fluid force, the gyroscopic term again
        for k in range(2):
            if gyro:
                w[k] += MP[k]["Iinv"] @ np.cross(MP[k]["I"] @ w[k], w[k]) * dt
    for k in range(2):
        np.testing.assert_allclose(o.body_field(k, "contact_force"), F[k], rtol=1e-10)
        np.testing.assert_allclose(o.body_field(k, "contact_torque"), T[k], rtol=1e-10)
        r = o.rigid(k)
        np.testing.assert_allclose(r["v"], v[k], rtol=1e-12)
        np.testing.assert_allclose(r["omega"], w[k], rtol=1e-12)


def _split_net(n):
    Fv, Fx, Fw, Tw, Tx, Tv = [n[9 * k:9 * k + 9].reshape(3, 3) for k in range(6)]
    return dict(Fv=Fv, Fx=Fx, Fw=Fw, Tw=Tw, Tx=Tx, Tv=Tv, Fq=n[54:66].reshape(3, 4), Tq=n[66:78].reshape(3, 4))


def _two_balls(mu, **kw):
    ball = scenes.sphere_surface(0.05, 2 * R_P)
    ball = ball - ball.mean(0)
    p = _params(rigid_contact_beta=3000.0, rigid_contact_friction=mu, gravity=(0.0, -9.81, 0.0),
                compute_gradient=1, use_gradient_manager=1, **kw)
    return _scene(p, [_body(ball, [0.0, 0.0, 0.0], v0=(0.3, -0.1, 0.2), w0=(2.0, -1.0, 3.0)),
                      _body(ball, [0.093, 0.012, -0.008], v0=(-0.2, 0.1, 0.0), w0=(-1.0, 0.5, 0.7), density=600.0)])


def _fd(probe, R, Q, which, out, ncol, eps=1e-7):
    cols = []
    for c in range(ncol):
        e = np.zeros(ncol)
        e[c] = eps
        a = probe.contact_probe(R, Q, **{which: e})[out]
        b = probe.contact_probe(R, Q, **{which: -e})[out]
        cols.append((a - b) / (2 * eps))
    return np.stack(cols, 1)


def test_contact_gradient_terms_against_differences():
    """The sums the contact solver hands the manager (solveRigidContactPenalty,
    RigidContactSolver.cpp:471-553) where they are derivatives of the contact
    step: with no friction, d/dx and d/dq of sum f_n through the density (its
    other-body part; n held; df_n/drho taken as -beta n as the reference does)
    and of the torques, for the body itself and for the other body (the q
    derivative of the homogeneous rotation form the reference differentiates);
    with friction, d/dv and d/domega of the friction sum (|f_n| and the weights
    held) for the body itself and d/domega for the other body.  The other
    body's d/dv is the reference's approximation (-the body's own term) and is
    not a derivative."""
    for mu in (0.0, 0.4):
        sc = _two_balls(mu)
        probe = OracleSim(sc, "f64")
        o = OracleSim(sc, "f64")
        o.step(1)
        for R, Q in ((0, 0), (0, 1), (1, 0), (1, 1)):
            N = _split_net(o.gradient_manager_pair(R, Q)[1])
            check = []
            if mu == 0.0:
                check += [("Fx", _fd(probe, R, Q, "dx", 0, 3)), ("Tx", _fd(probe, R, Q, "dx", 1, 3)),
                          ("Fq", _fd(probe, R, Q, "dq", 0, 4))]
            else:
                check += [("Fw", _fd(probe, R, Q, "dw", 2, 3))]
                if R == Q:
                    check += [("Fv", _fd(probe, R, Q, "dv", 2, 3))]
            for name, fd in check:
                scale = np.abs(fd).max()
                assert scale > 1e-3, (R, Q, name)
                np.testing.assert_allclose(N[name], fd, rtol=0, atol=1e-6 * scale, err_msg=f"{name} ({R},{Q}) mu {mu}")


def test_multi_body_manager_without_contact_matches_the_single_body_chain():
    """With the contact solver on but the bodies apart, the manager's table
    across bodies (the chains after the fluid and after the rigid coupling,
    RigidBodyGradientManager.cpp:455-468) keeps the cross entries at zero and
    its diagonal equals the single-body chain pinned in
    test_oracle_pinning_coupling.py (gyroscopic term off: with it the contact
    path's extra updateBoundaryForces calls move omega)."""
    sc = scenes.submerged_body(body="ellipsoid", steps=8)
    far = scenes.sphere_surface(0.04, 2 * sc.params["radius"])
    sc.bodies.append(scenes.Body(far, True, np.array([0.17, 0.6, 0.17]), np.array([1.0, 0, 0, 0]), 500.0))
    sc.params.update(compute_gradient=1, use_gradient_manager=1, gyroscopic=0)
    runs = []
    for contact in (0, 1):
        p = dict(sc.params, use_rigid_contact=contact)
        o = OracleSim(scenes.Scene(sc.name, p, sc.fluid_x, sc.fluid_v, sc.bodies, sc.steps), "f64")
        o.step(8)
        runs.append(o)
    for k in (1, 2):
        j1, n1 = runs[0].gradient_manager(k)
        j2, n2 = runs[1].gradient_manager_pair(k, k)
        scale = np.abs(j1).max()
        np.testing.assert_allclose(j2, j1, rtol=0, atol=1e-9 * scale)
    np.testing.assert_array_equal(runs[1].gradient_manager_pair(1, 2)[0], 0.0)
    np.testing.assert_array_equal(runs[1].gradient_manager_pair(2, 1)[0], 0.0)
    assert np.abs(runs[1].gradient_manager_pair(1, 1)[0][18:27] - np.eye(3).ravel()).max() > 1e-3  # v/v0 moved


def test_cross_body_gradient_after_a_hit():
    """A ball driven into a resting one: after the contact the passive ball's
    position depends on the active ball's initial velocity through the manager's
    cross entries (the quantity diff-rigid-contact-multi.py reads,
    get_grad_x_to_v0(passive, active)); along the line of the hit it is
    positive (a faster hit pushes the passive ball further)."""
    ball = scenes.sphere_surface(0.05, 2 * R_P)
    ball = ball - ball.mean(0)
    p = _params(rigid_contact_beta=3000.0, compute_gradient=1, use_gradient_manager=1)
    sc = _scene(p, [_body(ball, [0.0, 0.0, 0.0], v0=(1.0, 0.0, 0.0)), _body(ball, [0.105, 0.0, 0.0])])
    o = OracleSim(sc, "f64")
    o.step(40)
    J = o.gradient_manager_pair(1, 0)[0]
    dx_dv0 = J[0:9].reshape(3, 3)  # grad_x_to_v0 (layout: xv, xw, vv, vw, wv, ww, qv, qw)
    assert o.rigid(1)["v"][0] > 0.1  # the passive ball was hit
    assert dx_dv0[0, 0] > 1e-3
