"""Failure paths of the C ABI on the GPU.

* A neighbour list that outgrows its capacity fails dfr_step with
  DFR_ERR_CAPACITY (-3).
* A step that fails while a rollout is being recorded with every tape kept
  ends the recording; after dfr_reset a new checkpointed rollout and its
  backward pass give the gradient of a fresh context.
* A checkpointed backward pass whose segment tapes run out of device memory
  part-way through a segment (forced with the tape arena's byte limit) falls
  back to re-running every step twice and gives the kept-tape gradient.
"""
import ctypes as C

import numpy as np
import pytest

from difffr_amd import native, scenes
from difffr_amd.simulation import DfrError, Simulation

pytestmark = pytest.mark.gpu

R = 0.01


def _overcompressed():
    # lattice spacing 1.15 r: about 170 fluid neighbours per particle, beyond the 128 slots
    fluid = scenes.block_lattice((0.0, 0.0, 0.0), (12, 8, 12), 1.15 * R)
    inner = 1.15 * R * (np.array([12, 8, 12], np.float64) + 1)
    walls = scenes.box_surface((0.0, 0.0, 0.0), (inner[0], 2 * inner[1], inner[2]), 2 * R, open_top=True)
    bodies = [scenes.Body(walls, False, np.zeros(3), np.array([1.0, 0, 0, 0]), 1000.0, name="tank")]
    return scenes.Scene("overcompressed", scenes.default_params(R), fluid, np.zeros_like(fluid), bodies, 4)


def test_overflow_returns_capacity_code(gpu_ready):
    sc = _overcompressed()
    print(f"[capacity code] nf={sc.n_fluid}", flush=True)
    g = Simulation(sc)
    rc = native.load().dfr_step(g.ctx, 1)
    msg = native.load().dfr_last_error(g.ctx).decode()
    print("  rc", rc, msg, flush=True)
    assert rc == -3 and "capacity exceeded" in msg
    g.close()


def _seed():
    seed = np.zeros((2, 7))
    seed[1] = np.random.default_rng(11).normal(size=7)
    return seed


def test_failed_recorded_step_then_new_rollout(gpu_ready):
    sc = scenes.submerged_body(steps=8)
    params = {"compute_gradient": 0}
    fresh = Simulation(sc, params_override=params)
    fresh.adjoint_begin(1)
    fresh.step(6)
    want = fresh.adjoint_backward(_seed())[1]
    fresh.close()

    sim = Simulation(sc, params_override=params)
    sim.adjoint_arena_limit(1 << 20)  # far below one step's tape
    sim.adjoint_begin(0)
    print("[failed recording] kept-tape rollout under a 1 MiB tape limit", flush=True)
    with pytest.raises(DfrError, match="DFR_ERR_HIP"):
        sim.step(3)
    with pytest.raises(DfrError, match="DFR_ERR_STATE"):
        sim.adjoint_backward(_seed())  # the failed step ended the recording
    sim.adjoint_arena_limit(0)
    sim.reset()
    sim.adjoint_begin(1)
    sim.step(6)
    got = sim.adjoint_backward(_seed())[1]
    print("  fresh", want, "\n  after", got, flush=True)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * np.abs(want).max())
    sim.close()


def test_tape_memory_runs_out_mid_segment(gpu_ready):
    sc = scenes.submerged_body(body="ellipsoid", steps=8)
    params = {"compute_gradient": 0}
    ref = Simulation(sc, params_override=params)
    ref.adjoint_begin(0)
    ref.step(8)
    per_step = ref.adjoint_tape_bytes() / 8
    want = ref.adjoint_backward(_seed())[1]
    ref.close()
    print(f"[tape memory] {per_step / 2**20:.1f} MiB per step; limit 2.5 steps, segments of 4", flush=True)
    sim = Simulation(sc, params_override=params)
    sim.adjoint_arena_limit(int(2.5 * per_step))
    sim.adjoint_begin(4)
    sim.step(8)
    got = sim.adjoint_backward(_seed())[1]
    print("  kept", want, "\n  fallback", got, flush=True)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
    # the rollout's final state is back and a second backward pass agrees
    again = sim.adjoint_backward(_seed())[1]
    np.testing.assert_allclose(again, got, rtol=0, atol=1e-9 * np.abs(got).max())
    sim.close()
