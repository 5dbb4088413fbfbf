"""dfr_mlp_loss_grad_noalloc (fixed device scratch, no allocation) against
dfr_mlp_loss_grad (the same kernels, library-allocated buffers), through the C
ABI: two calls in a row, the second with fewer rows, so the second call runs
on the array the first one left behind.  Gradients agree to the order of their
float atomic sums (relative L2 1e-5), losses to 1e-6 relative."""
import ctypes as C

import numpy as np
import pytest

from difffr_amd import native, scenes
from difffr_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _call(fn, sim, feat, targ):
    g = np.zeros(native.MLP_NP, np.float32)
    loss = np.zeros(1)
    rc = fn(sim.ctx, feat.shape[0], _fp(feat), _fp(targ), _fp(g), _dp(loss))
    return rc, g, float(loss[0])


def test_no_alloc_loss_grad_matches_the_allocating_entry_point(gpu_ready):
    L = native.load()
    sc = scenes.submerged_body(steps=2)
    sim = Simulation(sc)
    sim.mlp_init_params(seed=3)
    rng = np.random.default_rng(11)
    for n in (40000, 3000):
        print(f"[no-alloc] rows={n}", flush=True)
        feat = rng.normal(size=(n, 6)).astype(np.float32)
        targ = rng.normal(size=(n, 2)).astype(np.float32)
        rc0, g0, l0 = _call(L.dfr_mlp_loss_grad, sim, feat, targ)
        rc1, g1, l1 = _call(L.dfr_mlp_loss_grad_noalloc, sim, feat, targ)
        assert rc0 == 0 and rc1 == 0
        assert np.linalg.norm(g1 - g0) <= 1e-5 * np.linalg.norm(g0)
        assert abs(l1 - l0) <= 1e-6 * abs(l0)
        assert np.abs(g0).max() > 0
    big = 65536 + 1
    feat = np.zeros((big, 6), np.float32)
    targ = np.zeros((big, 2), np.float32)
    rc, _, _ = _call(L.dfr_mlp_loss_grad_noalloc, sim, feat, targ)
    assert rc == -3  # DFR_ERR_CAPACITY
    assert "exceed the fixed scratch" in L.dfr_last_error(sim.ctx).decode()
    sim.close()
