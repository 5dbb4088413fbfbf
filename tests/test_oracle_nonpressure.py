"""Pinning of the oracle's non-pressure forces and CFL step size (CPU only).

* Akinci 2013 cohesion / adhesion kernels: bit-exact against the reference's
  own SPHKernels.h compiled from the reference tree (oracle/_ref), and their
  zeros and extremum;
* a free cube of fluid at rest in zero gravity (8 x 8 x 8 lattice plus a far
  static plate so the scene has a boundary): surface tension (Akinci 2013)
  conserves momentum (the pair terms are antisymmetric), vanishes at the
  cube's centre by symmetry and pulls the face centres inward; normals vanish
  at the centre and point inward on the faces;
* viscosity (Standard and XSPH): a uniform translation gives exactly zero,
  the force is linear in the coefficient, XSPH's is proportional to 1/dt and
  a shear flow slows the fastest layer; Standard viscosity conserves momentum
  between equal-density particles;
* CFL: the step size equals cflFactor 0.4 (2r) / sqrt(max(0.1, |v + a dt|^2))
  clamped to [min, max], the velocities and positions advance and the time
  moves with the step's starting size, the next step starts with the new one.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from difffr_amd import scenes
from oracle import pyoracle as po
from oracle.pyoracle import OracleSim

REF_LIB = os.path.join(os.path.dirname(po.__file__), "_ref", "libref.so")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_cohesion_adhesion_bit_exact_against_reference_build():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    L = C.CDLL(REF_LIB)
    L.ref_cohesion_adhesion_eval.argtypes = [C.c_double, C.POINTER(C.c_double), C.c_int,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rng = np.random.default_rng(3)
    for h in (0.1, 0.06, 0.0425):
        r = rng.uniform(-1.1 * h, 1.1 * h, size=(20000, 3))
        r[:3] = 0.0
        r[3:13, 0] = np.linspace(0, h, 10)  # r = h/2 and r = h exactly
        n = r.shape[0]
        Cr, Ar = np.empty(n), np.empty(n)
        L.ref_cohesion_adhesion_eval(h, _dp(r), n, _dp(Cr), _dp(Ar))
        Co, Ao = po.cohesion_adhesion_eval(h, r, double=False)
        assert np.array_equal(Cr, Co)
        assert np.array_equal(Ar, Ao)


def test_cohesion_adhesion_shapes():
    h = 0.1
    r = np.zeros((5, 3))
    r[:, 0] = [0.5 * h, 0.75 * h, h, 0.7 * h, 0.8 * h]
    C_, A_ = po.cohesion_adhesion_eval(h, r)
    assert A_[0] == 0.0 and A_[2] == pytest.approx(0.0, abs=1e-12) and C_[2] == pytest.approx(0.0, abs=1e-12)
    assert A_[1] > A_[3] and A_[1] > A_[4]  # the adhesion kernel peaks at 3h/4
    # cohesion: C(0) is the constant term alone, -h^6/64 as the reference
    # evaluates it (SPHKernels.h:473: not multiplied by the kernel's factor,
    # unlike Akinci et al.'s formula), and C > 0 on the rest of the support
    C0, _ = po.cohesion_adhesion_eval(h, np.zeros((1, 3)))
    assert C0[0] == pytest.approx(-h ** 6 / 64, rel=1e-12)
    rr = np.zeros((200, 3))
    rr[:, 0] = np.linspace(0.01 * h, 0.995 * h, 200)
    Cs, _ = po.cohesion_adhesion_eval(h, rr)
    assert Cs.min() > 0


def _cube(params, n=8, radius=0.025, v=None):
    sp = 2 * radius
    x = scenes.block_lattice((0.0, 0.0, 0.0), (n, n, n), sp)
    plate = scenes.box_surface((5.0, 5.0, 5.0), (5.4, 5.05, 5.4), sp)
    p = scenes.default_params(radius, 1e-3)
    p.update(gravity=(0.0, 0.0, 0.0), compute_gradient=0)
    p.update(params)
    vv = np.zeros_like(x) if v is None else v(x)
    return scenes.Scene("cube", p, x, vv, [scenes.Body(plate, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0)], 1)


def _accel(sc, variant="f64"):
    o = OracleSim(sc, variant)
    o.step(1)
    return o, o.field("accel")


def _centre_and_faces(x):
    c = x.mean(0)
    centre = int(np.argmin(np.linalg.norm(x - c, axis=1)))
    faces = []
    for k in range(3):
        for sgn in (-1, 1):
            target = c.copy()
            target[k] = x[:, k].max() if sgn > 0 else x[:, k].min()
            d = np.linalg.norm(x - target, axis=1)
            faces.append((int(np.argmin(d)), k, sgn))
    return centre, faces


def test_surface_tension_invariants():
    sc = _cube(dict(surface_tension_method=2, surface_tension=0.2, surface_tension_boundary=0.0))
    o, a = _accel(sc)
    n = o.field("normal")
    x = sc.fluid_x
    centre, faces = _centre_and_faces(x)
    scale = np.abs(a).max()
    assert scale > 0
    # momentum: sum of m a over the fluid (equal masses) vanishes
    assert np.abs(a.sum(0)).max() <= 1e-10 * np.abs(a).sum()
    assert np.abs(a[centre]).max() <= 1e-9 * scale
    assert np.abs(n[centre]).max() <= 1e-9 * np.abs(n).max()
    for i, k, sgn in faces:
        assert sgn * a[i, k] < -0.05 * scale, (i, k, sgn, a[i])  # pulled inward
        assert sgn * n[i, k] < 0, (i, k, sgn, n[i])              # normal points inward


def test_viscosity_invariants():
    shear = lambda x: np.stack([2.0 * x[:, 1], np.zeros(len(x)), np.zeros(len(x))], 1)
    for method in (1, 2):
        # a uniform translation: exactly no viscous acceleration
        sc = _cube(dict(viscosity_method=method, viscosity=0.01), v=lambda x: np.tile([0.3, -0.2, 0.1], (len(x), 1)))
        _, a = _accel(sc)
        assert np.array_equal(a, np.zeros_like(a))
        # linear in the coefficient
        a1 = _accel(_cube(dict(viscosity_method=method, viscosity=0.01), v=shear))[1]
        a2 = _accel(_cube(dict(viscosity_method=method, viscosity=0.02), v=shear))[1]
        np.testing.assert_allclose(a2, 2 * a1, rtol=1e-12, atol=1e-15)
        # the fastest (top) layer is slowed, the slowest (bottom) layer dragged along
        x = _cube({}).fluid_x
        top = x[:, 1] == x[:, 1].max()
        bot = x[:, 1] == x[:, 1].min()
        assert np.all(a1[top, 0] < 0) and np.all(a1[bot, 0] > 0)
        if method == 1:
            assert np.abs(a1.sum(0)).max() <= 1e-10 * np.abs(a1).sum()  # equal densities: momentum kept
    # XSPH scales with 1/dt (Viscosity_XSPH.cpp:47)
    sc1 = _cube(dict(viscosity_method=2, viscosity=0.01), v=shear)
    sc2 = _cube(dict(viscosity_method=2, viscosity=0.01, dt=2e-3), v=shear)
    np.testing.assert_allclose(_accel(sc2)[1], 0.5 * _accel(sc1)[1], rtol=1e-12, atol=1e-15)


def test_cfl_step_size_and_time_bookkeeping():
    r = 0.025
    base = dict(cfl_method=1, cfl_factor=0.5, cfl_min_time_step_size=1e-4, cfl_max_time_step_size=1.0)
    # at rest with gravity: the bound is the 0.1 floor of the squared speed
    sc = _cube(dict(base, gravity=(0.0, -9.81, 0.0)))
    o = OracleSim(sc, "f64")
    o.step(1)
    h1 = 0.5 * 0.4 * (2 * r) / math.sqrt(0.1)
    assert o.field("dt") == pytest.approx(h1, rel=1e-12)
    assert o.field("time") == pytest.approx(1e-3, rel=1e-12)  # the first step ran with the initial size
    x1, v1 = o.fluid()
    o.step(1)
    assert o.field("time") == pytest.approx(1e-3 + h1, rel=1e-12)
    # one fast particle sets the bound (|v + a dt| with the step's starting size)
    fast = lambda x: np.where(np.arange(len(x))[:, None] == 7, np.array([[6.0, 0.0, 8.0]]), 0.0)
    sc = _cube(dict(base, gravity=(0.0, -9.81, 0.0)), v=fast)
    o = OracleSim(sc, "f64")
    o.step(1)
    # after the divergence solve the particle still moves at about 10 m/s; the
    # bound uses its velocity then, so the step is within a few percent of
    # 0.5 * 0.4 * 2r / 10
    assert o.field("dt") == pytest.approx(0.5 * 0.4 * 2 * r / 10.0, rel=0.1)
    # the clamps
    sc = _cube(dict(base, cfl_max_time_step_size=0.004, gravity=(0.0, -9.81, 0.0)))
    o = OracleSim(sc, "f64")
    o.step(1)
    assert o.field("dt") == pytest.approx(0.004, rel=1e-12)
    sc = _cube(dict(base, cfl_min_time_step_size=0.01, gravity=(0.0, -9.81, 0.0)), v=fast)
    o = OracleSim(sc, "f64")
    o.step(1)
    assert o.field("dt") == pytest.approx(0.01, rel=1e-12)


def test_surface_tension_two_particles_closed_form():
    """Two fluid particles at distance d < h, alone (a far plate gives the scene
    its boundary): rho = rho0 V (W(0) + W(d)) for both, n_1 = h (m / rho) gradW(x1 - x2)
    = -n_2, and the acceleration of particle 1 along e = (x1 - x2)/d is
    K (-k m C(d) - k 2 n_1.e) with K = 2 rho0 / (2 rho) (Akinci et al. 2013,
    eq. 1-4 as SurfaceTension_Akinci2013.cpp:78-110 evaluates them): the
    cohesion term attracts, the curvature term of this pair pushes apart."""
    r, k = 0.025, 0.2
    h = 4 * r
    for d in (0.3 * h, 0.7 * h):
        x = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
        plate = scenes.box_surface((5.0, 5.0, 5.0), (5.4, 5.05, 5.4), 2 * r)
        p = scenes.default_params(r, 1e-3)
        p.update(gravity=(0.0, 0.0, 0.0), compute_gradient=0, surface_tension_method=2, surface_tension=k,
                 surface_tension_boundary=0.0, enable_divergence_solver=0)
        sc = scenes.Scene("pair", p, x, np.zeros_like(x),
                          [scenes.Body(plate, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0)], 1)
        o = OracleSim(sc, "f64")
        o.step(1)
        a = o.field("accel")
        V = 0.8 * (2 * r) ** 3
        m = V * 1000.0
        W, gW, _ = po.kernel_eval(h, np.array([[0.0, 0.0, 0.0], [-d, 0.0, 0.0]]))
        rho = 1000.0 * V * (W[0] + W[1])
        n1 = h * (m / rho) * gW[1]          # gradW(x1 - x2), x1 - x2 = (-d, 0, 0)
        e = np.array([-1.0, 0.0, 0.0])      # (x1 - x2) / d
        Cd, _ = po.cohesion_adhesion_eval(h, np.array([[d, 0.0, 0.0]]))
        K = 2 * 1000.0 / (2 * rho)
        expect = K * (-k * m * Cd[0] - k * 2 * (n1 @ e))
        assert a[0] @ e == pytest.approx(expect, rel=1e-9)
        assert np.abs(a[0] + a[1]).max() <= 1e-12 * np.abs(a[0]).max()  # equal and opposite


def _he_cube(n, k=0.2, kb=0.0):
    return _cube(dict(surface_tension_method=3, surface_tension=k, surface_tension_boundary=kb), n=n)


def test_he2014_colour_field_and_invariants():
    """He et al. 2014 (SurfaceTension_He2014.cpp:29-160) on a free 13 x 13 x 13
    lattice cube in zero gravity: where a particle's neighbours all have complete
    neighbourhoods (at least two kernel radii inside) the colour field is exactly
    the Shepard sum 1 (rho_i = rho_j = rho0 V (W(0) + sum W)); dropping its W(0)
    term or a density factor breaks this.  |grad c|^2 vanishes at the centre by
    symmetry and peaks at the surface; the fluid pair force conserves momentum
    and pulls the face centres inward."""
    sc = _he_cube(13)
    o, a = _accel(sc)
    c, g2 = o.field("color"), o.field("gradC2")
    x = sc.fluid_x
    sp = 2 * 0.025
    idx = np.rint((x - x.min(0)) / sp).astype(int)
    deep = np.all((idx >= 4) & (idx <= 8), axis=1)
    assert deep.sum() == 125
    np.testing.assert_allclose(c[deep], 1.0, rtol=0, atol=1e-12)
    assert np.abs(c[~deep] - 1.0).max() > 1e-2  # near the surface the sum is not a partition of unity
    centre, faces = _centre_and_faces(x)
    assert g2[centre] <= 1e-20 * g2.max()
    assert np.argmax(g2) in np.flatnonzero(~deep)
    scale = np.abs(a).max()
    assert scale > 0
    assert np.abs(a.sum(0)).max() <= 1e-10 * np.abs(a).sum()
    assert np.abs(a[centre]).max() <= 1e-9 * scale
    for i, kk, sgn in faces:
        assert sgn * a[i, kk] < -0.05 * scale, (i, kk, sgn, a[i])
    # linear in the coefficient
    _, a2 = _accel(_he_cube(13, k=0.4))
    np.testing.assert_allclose(a2, 2 * a, rtol=1e-12, atol=1e-15)
    # a jittered cube has no symmetry left: momentum is kept because each pair's
    # terms are equal and opposite (equal masses)
    sc = _he_cube(9)
    sc.fluid_x = sc.fluid_x + np.random.default_rng(5).uniform(-0.004, 0.004, sc.fluid_x.shape)
    _, a = _accel(sc)
    assert np.abs(a.sum(0)).max() <= 1e-10 * np.abs(a).sum()


def test_he2014_three_particles_brute_force():
    """Three fluid particles at unequal distances (colours and densities all
    differ): density, colour, grad c, |grad c|^2 and the acceleration evaluated
    by direct summation of He et al. 2014's definitions (colour
    c_i = sum_j (m / rho_j) W_ij including j = i, grad c_i = (1 / c_i) sum_j
    (m / rho_j) c_j gradW_ij; the force (k / 4 rho_i) sum_j V_j (|grad c_i|^2 +
    |grad c_j|^2) gradW_ij) from the kernel's own evaluation."""
    r, k = 0.025, 0.3
    h = 4 * r
    x = np.array([[0.0, 0.0, 0.0], [0.35 * h, 0.1 * h, 0.0], [0.2 * h, 0.75 * h, 0.15 * h]])
    plate = scenes.box_surface((5.0, 5.0, 5.0), (5.4, 5.05, 5.4), 2 * r)
    p = scenes.default_params(r, 1e-3)
    p.update(gravity=(0.0, 0.0, 0.0), compute_gradient=0, surface_tension_method=3, surface_tension=k,
             surface_tension_boundary=0.0, enable_divergence_solver=0)
    sc = scenes.Scene("three", p, x, np.zeros_like(x),
                      [scenes.Body(plate, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0)], 1)
    o = OracleSim(sc, "f64")
    o.step(1)
    m = 0.8 * (2 * r) ** 3 * 1000.0
    Wm = np.zeros((3, 3))
    G = np.zeros((3, 3, 3))
    for i in range(3):
        W, gW, _ = po.kernel_eval(h, x[i] - x)
        Wm[i], G[i] = W, gW
    rho = m * Wm.sum(1)
    col = (Wm * (m / rho)[None, :]).sum(1)
    grad = np.einsum("j,ijk->ik", m / rho * col, G) / col[:, None]
    g2 = (grad ** 2).sum(1)
    acc = np.einsum("ij,ijk->ik", (k / (4 * rho))[:, None] * (m / rho)[None, :] * (g2[:, None] + g2[None, :]), G)
    np.testing.assert_allclose(o.field("density"), rho, rtol=1e-12)
    assert np.ptp(col) > 1e-2 and np.ptp(rho) > 1.0
    np.testing.assert_allclose(o.field("color"), col, rtol=1e-12)
    np.testing.assert_allclose(o.field("gradC2"), g2, rtol=1e-10)
    np.testing.assert_allclose(o.field("accel"), acc, rtol=1e-10, atol=1e-12 * np.abs(acc).max())


def test_he2014_two_particles_closed_form():
    """Two fluid particles at distance d < h (a far plate gives the scene its
    boundary): both colours are 1 (the Shepard sum of a symmetric pair),
    grad c_1 = (m / rho) gradW(x1 - x2), gradC2 = (m / rho)^2 |gradW|^2 for both,
    and a_1 = (k / 4 rho) (m / rho) 2 gradC2 gradW(x1 - x2) = -a_2."""
    r, k = 0.025, 0.3
    h = 4 * r
    for d in (0.3 * h, 0.7 * h):
        x = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
        plate = scenes.box_surface((5.0, 5.0, 5.0), (5.4, 5.05, 5.4), 2 * r)
        p = scenes.default_params(r, 1e-3)
        p.update(gravity=(0.0, 0.0, 0.0), compute_gradient=0, surface_tension_method=3, surface_tension=k,
                 surface_tension_boundary=0.0, enable_divergence_solver=0)
        sc = scenes.Scene("pair", p, x, np.zeros_like(x),
                          [scenes.Body(plate, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0)], 1)
        o = OracleSim(sc, "f64")
        o.step(1)
        V = 0.8 * (2 * r) ** 3
        m = V * 1000.0
        W, gW, _ = po.kernel_eval(h, np.array([[0.0, 0.0, 0.0], [-d, 0.0, 0.0]]))
        rho = 1000.0 * V * (W[0] + W[1])
        np.testing.assert_allclose(o.field("color"), 1.0, rtol=1e-12)
        g2 = (m / rho) ** 2 * (gW[1] @ gW[1])
        np.testing.assert_allclose(o.field("gradC2"), g2, rtol=1e-10)
        expect = (k / (4 * rho)) * (m / rho) * 2 * g2 * gW[1]
        a = o.field("accel")
        np.testing.assert_allclose(a[0], expect, rtol=1e-10, atol=1e-14 * np.abs(expect).max())
        np.testing.assert_allclose(a[1], -expect, rtol=1e-10, atol=1e-14 * np.abs(expect).max())


def test_he2014_boundary_reaction_on_the_body():
    """The boundary term of He 2014 pushes the fluid and, by
    BoundaryModel::addForce, the dynamic body with -m a: with the fluid term off
    (k = 0), no viscosity, the divergence solve off and a fluid sparse enough to
    stay below rest density (no pressure forces), the body's force after one step
    is exactly -sum_i m (a_i - g)."""
    r = 0.025
    sp = 3 * r  # 1.5 diameters: rho < rho0 everywhere
    fluid = scenes.block_lattice((-0.3, -0.3, -0.3), (13, 13, 13), sp)
    ball = scenes.sphere_surface(0.12, 2 * r)
    fluid = fluid[np.linalg.norm(fluid, axis=1) > 0.12 + 2 * r]
    p = scenes.default_params(r, 1e-3)
    p.update(compute_gradient=0, surface_tension_method=3, surface_tension=0.0, surface_tension_boundary=0.5,
             enable_divergence_solver=0)
    body = scenes.Body(ball, True, np.zeros(3), np.array([1.0, 0, 0, 0]), 800.0)
    sc = scenes.Scene("ball", p, fluid, np.zeros_like(fluid), [body], 1)
    o = OracleSim(sc, "f64")
    o.step(1)
    assert o.field("density").max() < 1000.0
    m = 0.8 * (2 * r) ** 3 * 1000.0
    a = o.field("accel") - np.asarray(p["gravity"])
    F = o.rigid(0)["force"]
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(F, -m * a.sum(0), rtol=1e-10, atol=1e-12 * np.abs(F).max())
