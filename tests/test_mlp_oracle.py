"""Pins the numpy restatement of the warm-start network (oracle/mlp_oracle.py):
float64 gradients against central differences, and the bfloat16 rounding
against values whose rounding is known."""
import numpy as np

from oracle import mlp_oracle as mo


def _net(seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(scale=0.3, size=mo.NP)
    norm = np.concatenate([rng.normal(size=6), rng.uniform(0.5, 2.0, 6), rng.normal(size=2), rng.uniform(0.5, 2.0, 2)])
    return P, norm, rng


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e-5], np.float32)
    r = mo.bf16(x)
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    np.testing.assert_array_equal(r[:4], np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0], np.float32))
    assert r[4] == np.float32(-2.5)
    assert abs(r[5] - 3.0e-5) / 3.0e-5 < 2.0 ** -8


def test_loss_gradient_matches_central_differences():
    P, norm, rng = _net(0)
    feat = rng.normal(size=(7, 6))
    targ = rng.normal(size=(7, 2))
    _, g = mo.loss_grad(P, norm, feat, targ, exact=True)
    n = feat.shape[0]
    idx = np.concatenate([rng.choice(np.arange(0, 192), 6), rng.choice(np.arange(192, 224), 3),
                          rng.choice(np.arange(224, 1248), 6), rng.choice(np.arange(2336, 2402), 6)])
    for k in idx:
        e = 1e-6
        Pp = P.copy(); Pp[k] += e
        Pm = P.copy(); Pm[k] -= e
        lp, _ = mo.loss_grad(Pp, norm, feat, targ, exact=True)
        lm, _ = mo.loss_grad(Pm, norm, feat, targ, exact=True)
        fd = (lp - lm) / (2 * e) / n
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_input_gradient_matches_central_differences():
    P, norm, rng = _net(1)
    feat = rng.normal(size=(3, 6))
    ybar = rng.normal(size=(3, 2))
    fb = mo.input_grad(P, norm, feat, ybar, exact=True)
    for p in range(3):
        for k in range(6):
            e = 1e-6
            fp = feat.copy(); fp[p, k] += e
            fm = feat.copy(); fm[p, k] -= e
            d = (mo.predict(P, norm, fp, exact=True) - mo.predict(P, norm, fm, exact=True)) / (2 * e)
            fd = float(np.sum(d * ybar))
            assert abs(fd - fb[p, k]) <= 1e-6 * max(1.0, abs(fd))


def test_rounded_network_tracks_exact_one():
    P, norm, rng = _net(2)
    feat = rng.normal(size=(50, 6))
    y32 = mo.predict(P, norm, feat)
    y64 = mo.predict(P, norm, feat, exact=True)
    assert np.abs(y32 - y64).max() <= 0.05 * np.abs(y64).max()
