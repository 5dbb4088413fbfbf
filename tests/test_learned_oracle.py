"""Pins the oracle's learned warm start (oracle/warmnet.hpp, the learned
branches of oracle/dfsph_oracle.hpp and oracle/dfsph_adjoint.hpp) on the CPU.

* the C++ network against the numpy network (oracle/mlp_oracle.py), whose
  float64 gradients are pinned by central differences in test_mlp_oracle.py;
* a constant network: the warm starts become kappa_v = c0 / dt where d > 0 and
  kappa = c1 / dt^2 where rho* > 1, in closed form against the recorded masks;
* the recorded features against the recorded densities, neighbour counts and
  the previous step's stiffness fields;
* the exact adjoint through the network and its features against FP64 central
  differences (exact float64 network, Jacobi iteration counts held fixed).
"""
import numpy as np

from difffr_amd import scenes
from oracle import mlp_oracle as mo
from oracle import pyoracle as po

STEPS = 12


def _net(seed, norm):
    rng = np.random.default_rng(seed)
    P = np.zeros(mo.NP, np.float32)
    o = 0
    for m, n in mo.SIZES:
        P[o:o + m * n] = rng.uniform(-1, 1, m * n) * np.sqrt(6.0 / n)
        o += m * n
        P[o:o + m] = rng.normal(scale=0.1, size=m)
        o += m
    return P, np.asarray(norm, np.float32)


def _fitted_norm(sc, steps):
    """Feature normalisation from the features a first learned run records,
    output normalisation of the size of the stored stiffness on this scene."""
    probe = np.concatenate([np.zeros(6), np.ones(6), [-1.4e-3, -1e-8], [1.3e-3, 1.4e-7]])
    P, nm = _net(11, probe)
    s = po.OracleSim(sc, "f64")
    s.set_warm_net(P, nm, bf16=False)
    s.set_record(True)
    s.step(steps)
    f = np.concatenate([s.tape_field(k, "feat").reshape(-1, 6) for k in range(steps)])
    sd = f.std(axis=0)
    sd[sd < 1e-30] = 1.0
    return np.concatenate([f.mean(axis=0), sd, probe[12:]]).astype(np.float32)


def test_cpp_network_matches_numpy_network():
    rng = np.random.default_rng(0)
    norm = np.concatenate([rng.normal(size=6), rng.uniform(0.5, 2.0, 6), rng.normal(size=2), rng.uniform(0.5, 2.0, 2)])
    P, nm = _net(1, norm)
    feat = nm[:6] + rng.normal(size=(200, 6)) * nm[6:12]
    ybar = rng.normal(size=(200, 2))
    y64 = po.net_predict(P, nm, feat, bf16=False)
    np.testing.assert_allclose(y64, mo.predict(P, nm, feat, exact=True), rtol=1e-10, atol=1e-12)
    g64 = po.net_input_grad(P, nm, feat, ybar, bf16=False)
    np.testing.assert_allclose(g64, mo.input_grad(P, nm, feat, ybar, exact=True), rtol=1e-9, atol=1e-12)
    # rounded: the same bf16 decisions up to flips from float32 summation order
    y16 = po.net_predict(P, nm, feat)
    ref = mo.predict(P, nm, feat)
    assert np.abs(y16 - ref).max() <= 2e-2 * np.abs(ref).max()
    assert np.median(np.abs(y16 - ref)) <= 1e-5 * np.abs(ref).max()
    g16 = po.net_input_grad(P, nm, feat, ybar)
    gref = mo.input_grad(P, nm, feat, ybar)
    assert np.abs(g16 - gref).max() <= 2e-2 * np.abs(gref).max()


def test_constant_network_warm_starts_closed_form():
    sc = scenes.submerged_body(steps=6)
    c0, c1 = float(np.float32(-2.0e-3)), float(np.float32(-3.0e-7))  # stored units: kappa_v dt, kappa dt^2
    P = np.zeros(mo.NP, np.float32)
    P[-2:] = (c0, c1)  # output biases; every weight zero
    nm = np.concatenate([np.zeros(6), np.ones(6), np.zeros(2), np.ones(2)]).astype(np.float32)
    s = po.OracleSim(sc, "f64")
    s.set_warm_net(P, nm, bf16=False)
    s.set_record(True)
    dt = sc.params["dt"]
    prev_k = np.zeros(s.n)
    prev_kv = np.zeros(s.n)
    active = [0, 0]
    for k in range(6):
        s.step(1)
        d0, kvW = s.tape_field(k, "d0"), s.tape_field(k, "kvW")
        rsW, kW = s.tape_field(k, "rsW"), s.tape_field(k, "kW")
        active[0] += int((d0 > 0).sum())
        active[1] += int((rsW > 1).sum())
        np.testing.assert_allclose(kvW, np.where(d0 > 0, c0 / dt, 0.0), rtol=1e-14)
        np.testing.assert_allclose(kW, np.where(rsW > 1, c1 / dt ** 2, 0.0), rtol=1e-14)
        # features: rho/rho0 - 1, neighbour count, previous stiffness fields
        f = s.tape_field(k, "feat").reshape(-1, 6)
        np.testing.assert_allclose(f[:, 0], s.tape_field(k, "rho") / sc.params["rho0"] - 1.0, atol=1e-14)
        np.testing.assert_array_equal(f[:, 2], s.tape_field(k, "nn"))
        np.testing.assert_array_equal(f[:, 4], prev_k)
        np.testing.assert_array_equal(f[:, 5], prev_kv)
        # boundary density: nonzero exactly where a particle has boundary neighbours
        assert np.all(f[:, 3] >= 0) and (f[:, 3] > 0).any()
        # the divergence feature is dt times the unclamped rate, whose clamp is d0
        nn = f[:, 2]
        np.testing.assert_allclose(np.where(nn >= 20, np.maximum(f[:, 1] / dt, 0), 0), d0, rtol=1e-12, atol=1e-14)
        prev_k = s.field("kappa")
        prev_kv = s.field("kappa_v")
    assert active[0] > 0 and active[1] > 0  # both masks were exercised


def _run(sc, net, v0, w0, steps, fix=None, record=False):
    s = po.OracleSim(sc, "f64")
    s.set_warm_net(*net, bf16=False)
    s.set_init_velocity(1, v0, w0)
    if fix is not None:
        s.set_fixed_iterations(*fix)
    if record:
        s.set_record(True)
    s.step(steps)
    r = s.rigid(1)
    return s, np.concatenate([r["x"], r["q"]])


def test_learned_adjoint_matches_central_differences():
    sc = scenes.submerged_body(body="ellipsoid", steps=STEPS)
    net = _net(5, _fitted_norm(sc, STEPS))
    b = sc.bodies[1]
    s, _ = _run(sc, net, b.v0, b.w0, STEPS, record=True)
    fix = s.recorded_iterations()
    seed = np.random.default_rng(2).normal(size=7)
    sd = np.zeros((2, 7)); sd[1] = seed
    g = s.adjoint(sd)[1]
    p0 = np.concatenate([b.v0, b.w0])
    fd = np.zeros(6)
    eps = 1e-6
    for k in range(6):
        p = p0.copy(); p[k] += eps
        m = p0.copy(); m[k] -= eps
        _, a = _run(sc, net, p[:3], p[3:], STEPS, fix)
        _, c = _run(sc, net, m[:3], m[3:], STEPS, fix)
        fd[k] = seed @ (a - c) / (2 * eps)
    print("adjoint", g, "\nfd     ", fd)
    scale = np.abs(fd).max()
    assert scale > 1e-4
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-6 * scale)
    # the network changes the trajectory: the plain warm start's gradient differs
    s2 = po.OracleSim(sc, "f64")
    s2.set_record(True)
    s2.step(STEPS)
    g_plain = s2.adjoint(sd)[1]
    assert np.abs(g_plain - g).max() > 1e-4 * scale
