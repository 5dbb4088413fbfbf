"""Host-side wrapper of one libdfr_hip.so context.

`Simulation` owns a dfr_ctx built from a `difffr_amd.scenes.Scene`; its
`boundary_model(i)` objects expose the getters DiffFR's drivers read from
BoundaryModel_Akinci2012 (pySPlisHSPlasH/BoundaryModelModule.cpp:43-91), and the
simulation exposes the TimeStepDiffDFSPH-level ones
(pySPlisHSPlasH/DiffDFSPHModule.cpp:44-102).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DfrError(RuntimeError):
    pass


class AdjointRangeError(DfrError):
    """The exact gradient left the FP32 range; `grad` holds what the reverse pass produced."""

    def __init__(self, msg: str, grad: np.ndarray):
        super().__init__(msg)
        self.grad = grad


def make_params(p: dict) -> native.Params:
    L = native.load()
    q = native.Params()
    L.dfr_default_params(C.byref(q), float(p["radius"]), float(p["dt"]))
    for k in ("rho0", "max_error", "max_error_v", "uniform_acc_time", "target_time", "surface_tension",
              "surface_tension_boundary", "viscosity", "viscosity_boundary", "cfl_factor",
              "cfl_min_time_step_size", "cfl_max_time_step_size", "rigid_contact_beta", "rigid_contact_gamma",
              "rigid_contact_support_radius_factor", "rigid_contact_friction"):
        if k in p:
            setattr(q, k, float(p[k]))
    for k in ("min_iterations", "max_iterations", "max_iterations_v", "use_pressure_warmstart",
              "use_divergence_warmstart", "enable_divergence_solver", "gradient_mode", "optimize_rotation",
              "gyroscopic", "compute_gradient", "max_neighbors", "max_boundary_neighbors", "release_mode",
              "use_gradient_manager", "surface_tension_method", "viscosity_method", "cfl_method",
              "use_rigid_contact"):
        if k in p:
            setattr(q, k, int(p[k]))
    if "gravity" in p:
        for k in range(3):
            q.gravity[k] = float(p["gravity"][k])
    return q


class BoundaryModel:
    """Getters of one body (BoundaryModel_Akinci2012.h:169-323)."""

    def __init__(self, sim: "Simulation", index: int):
        self.sim = sim
        self.index = index

    def _rigid(self):
        return self.sim.rigid(self.index)

    def get_position_rb(self):
        return self._rigid()["x"]

    def get_quaternion_rb_vec4(self):
        return self._rigid()["q"]

    def get_velocity_rb(self):
        return self._rigid()["v"]

    def get_angular_velocity_rb(self):
        return self._rigid()["omega"]

    def getForce(self):
        return self._rigid()["force"]

    def getTorque(self):
        return self._rigid()["torque"]

    def _J(self):
        return self.sim.jacobians(self.index)

    def get_grad_x_to_v0(self):
        return self._J()["Jxv"]

    def get_grad_x_to_omega0(self):
        return self._J()["Jxw"]

    def get_grad_v_to_v0(self):
        return self._J()["Jvv"]

    def get_grad_v_to_omega0(self):
        return self._J()["Jvw"]

    def get_grad_omega_to_v0(self):
        return self._J()["Jwv"]

    def get_grad_omega_to_omega0(self):
        return self._J()["Jww"]

    def get_grad_quaternion_to_v0(self):
        return self._J()["Jqv"]

    def get_grad_quaternion_to_omega0(self):
        return self._J()["Jqw"]


def unpack_jacobians(o: np.ndarray) -> dict:
    names = ("Jxv", "Jxw", "Jvv", "Jvw", "Jwv", "Jww")
    d = {n: o[9 * k:9 * k + 9].reshape(3, 3).copy() for k, n in enumerate(names)}
    d["Jqv"] = o[54:66].reshape(4, 3).copy()
    d["Jqw"] = o[66:78].reshape(4, 3).copy()
    return d


def unpack_net(o: np.ndarray) -> dict:
    names = ("Fv", "Fx", "Fw", "Tw", "Tx", "Tv")
    d = {n: o[9 * k:9 * k + 9].reshape(3, 3).copy() for k, n in enumerate(names)}
    d["Fq"] = o[54:66].reshape(3, 4).copy()
    d["Tq"] = o[66:78].reshape(3, 4).copy()
    return d


class Simulation:
    """One MI355X simulation context built from a scene."""

    def __init__(self, scene, device: int = 0, params_override: dict | None = None):
        L = native.load()
        p = dict(scene.params)
        if params_override:
            p.update(params_override)
        self.params = p
        self._p = make_params(p)
        self.ctx = L.dfr_create(C.byref(self._p), int(device))
        if not self.ctx:
            raise DfrError("dfr_create failed (no HIP device?)")
        self.n = scene.n_fluid
        x = np.ascontiguousarray(scene.fluid_x, np.float32)
        v = np.ascontiguousarray(scene.fluid_v, np.float32)
        self._check(L.dfr_set_fluid(self.ctx, self.n, _fp(x), _fp(v)))
        self.n_bodies = 0
        self.body_sizes = []
        self._dyn = [bool(b.dynamic) for b in scene.bodies]
        for b in scene.bodies:
            x0 = np.ascontiguousarray(b.x0, np.float32)
            xr = np.ascontiguousarray(b.x_rb, np.float64)
            q = np.ascontiguousarray(b.q, np.float64)
            bid = L.dfr_add_boundary(self.ctx, x0.shape[0], _fp(x0), int(b.dynamic), _dp(xr), _dp(q),
                                     float(b.density))
            self._check(bid)
            v0 = np.ascontiguousarray(b.v0, np.float64)
            w0 = np.ascontiguousarray(b.w0, np.float64)
            self._check(L.dfr_set_init_velocity(self.ctx, bid, _dp(v0), _dp(w0)))
            self.n_bodies += 1
            self.body_sizes.append(x0.shape[0])
        self._check(L.dfr_init(self.ctx))

    def _check(self, r):
        if r < 0:
            msg = native.load().dfr_last_error(self.ctx)
            raise DfrError(f"{native.ERRORS.get(r, r)}: {msg.decode() if msg else ''}")
        return r

    def close(self):
        if getattr(self, "ctx", None):
            native.load().dfr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping -----------------------------------------------------------
    def step(self, n: int = 1):
        self._check(native.load().dfr_step(self.ctx, int(n)))

    def reset(self):
        self._check(native.load().dfr_reset(self.ctx))

    def reset_gradient(self):
        self._check(native.load().dfr_reset_gradient(self.ctx))

    def set_compute_gradient(self, on: bool):
        """DiffFR's forward-mode gradient accumulation on/off for the next steps."""
        self._check(native.load().dfr_set_compute_gradient(self.ctx, int(bool(on))))
        self.params["compute_gradient"] = int(bool(on))

    # -- state --------------------------------------------------------------
    def fluid(self):
        x = np.empty((self.n, 3), np.float32)
        v = np.empty((self.n, 3), np.float32)
        self._check(native.load().dfr_get_fluid(self.ctx, _fp(x), _fp(v)))
        return x, v

    def field(self, name: str) -> np.ndarray:
        """Per-particle fields in particle-id order: density, factor, density_adv,
        kappa, kappa_v; (n, 3): sum_grad_p_k, and with non-pressure forces on, the
        last step's accelerations (accel = g + forces) and Akinci 2013 normals;
        with He 2014 surface tension its colour field "color" and "gradC2"."""
        k = 3 if name in ("sum_grad_p_k", "accel", "normal") else 1
        out = np.empty(self.n * k, np.float32)
        self._check(native.load().dfr_get_field(self.ctx, name.encode(), _fp(out)))
        return out.reshape(self.n, 3) if k == 3 else out

    def set_field(self, name: str, values: np.ndarray):
        """"kappa", "kappa_v" ([n]) or "x", "v" ([n][3]), particle-id order."""
        a = np.ascontiguousarray(values, np.float32)
        k = 3 if name in ("x", "v") else 1
        if a.size != k * self.n:
            raise ValueError(f"{name}: expected {k * self.n} values, got {a.size}")
        self._check(native.load().dfr_set_field(self.ctx, name.encode(), _fp(a)))

    def set_time(self, time: float, step_count: int):
        self._check(native.load().dfr_set_time(self.ctx, float(time), int(step_count)))

    def body_field(self, body: int, name: str) -> np.ndarray:
        n = self.body_sizes[body]
        k = 1 if name == "volume" else 3
        out = np.empty(n * k, np.float32)
        self._check(native.load().dfr_get_boundary_field(self.ctx, body, name.encode(), _fp(out)))
        return out if k == 1 else out.reshape(n, 3)

    def rigid(self, body: int) -> dict:
        o = np.empty(19)
        self._check(native.load().dfr_get_rigid(self.ctx, body, _dp(o)))
        return dict(x=o[0:3], q=o[3:7], v=o[7:10], omega=o[10:13], force=o[13:16], torque=o[16:19])

    def set_rigid(self, body: int, x, q, v, omega):
        a = np.ascontiguousarray(np.concatenate([x, q, v, omega]), np.float64)
        self._check(native.load().dfr_set_rigid(self.ctx, body, _dp(a)))

    def jacobians_raw(self, body: int) -> np.ndarray:
        o = np.empty(78)
        self._check(native.load().dfr_get_jacobians(self.ctx, body, _dp(o)))
        return o

    def jacobians(self, body: int) -> dict:
        return unpack_jacobians(self.jacobians_raw(body))

    def gradient_manager_raw(self, R: int, RR: int):
        """RigidBodyGradientManager entries (R, RR): (Jacobians[78], per-step sums[78])."""
        j = np.empty(78)
        n = np.empty(78)
        self._check(native.load().dfr_get_gradient_manager(self.ctx, R, RR, _dp(j), _dp(n)))
        return j, n

    def net_gradients_raw(self, body: int) -> np.ndarray:
        o = np.empty(78)
        self._check(native.load().dfr_get_net_gradients(self.ctx, body, _dp(o)))
        return o

    def stats(self) -> dict:
        s = native.Stats()
        self._check(native.load().dfr_get_stats(self.ctx, C.byref(s)))
        return {k: getattr(s, k) for k, _ in native.Stats._fields_}

    def reset_stats(self):
        self._check(native.load().dfr_reset_stats(self.ctx))

    def neighbor_stats(self) -> dict:
        o = np.empty(3)
        self._check(native.load().dfr_neighbor_stats(self.ctx, _dp(o)))
        return dict(mean_fluid=o[0], mean_boundary=o[1], mean_fluid_per_dynamic_sample=o[2])

    def set_halo_capacity(self, slots: int = -1):
        """LDS halo capacity of the gather passes (-1: follow the largest halo;
        0: every halo read from global memory)."""
        self._check(native.load().dfr_set_halo_capacity(self.ctx, int(slots)))

    def halo_stats(self) -> dict:
        o = np.zeros(3, np.int64)
        self._check(native.load().dfr_halo_stats(self.ctx, o.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(largest_halo=int(o[0]), capacity=int(o[1]), max_capacity=int(o[2]))

    def neighbor_lists(self, kind: int):
        """kind 0: fluid->fluid (ids), 1: fluid->boundary (global sample index), 2: dynamic sample->fluid."""
        L = native.load()
        rows = self.n if kind in (0, 1) else sum(
            s for s, dyn in zip(self.body_sizes, self._dyn) if dyn)
        cap = max(rows, 1) * int(self._p.max_neighbors if kind != 2 else self._p.max_boundary_neighbors)
        counts = np.empty(max(rows, 1), np.uint32)
        lists = np.empty(max(cap, 1), np.uint32)
        r = L.dfr_get_neighbors(self.ctx, kind, counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                lists.ctypes.data_as(C.POINTER(C.c_uint32)), cap)
        self._check(int(r))
        off = np.concatenate([[0], np.cumsum(counts[:r].astype(np.int64))])
        return [lists[off[k]:off[k + 1]].copy() for k in range(int(r))]

    def set_profiling(self, on: bool, classes=None):
        """Kernel-class timing with HIP events; classes: names to time (None: all)."""
        if classes is None:
            self._check(native.load().dfr_set_profiling(self.ctx, int(on)))
            return
        names = list(self.kernel_times())
        mask = 0
        for cl in classes:
            mask |= 1 << names.index(cl)
        self._check(native.load().dfr_set_profiling_classes(self.ctx, mask if on else 0))

    def kernel_times(self) -> dict:
        n = 32
        names = (C.c_char * 32 * n)()
        avg = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        k = native.load().dfr_kernel_times(self.ctx, n, C.cast(names, C.c_void_p), _dp(avg),
                                           cnt.ctypes.data_as(C.POINTER(C.c_int64)))
        return {names[i].value.decode(): (float(avg[i]), int(cnt[i])) for i in range(k)}

    # -- exact adjoint --------------------------------------------------------
    def set_init_velocity(self, body: int, v0, w0):
        v = np.ascontiguousarray(v0, np.float64)
        w = np.ascontiguousarray(w0, np.float64)
        self._check(native.load().dfr_set_init_velocity(self.ctx, int(body), _dp(v), _dp(w)))

    def adjoint_begin(self, checkpoint_every: int = 1):
        """Start recording a rollout for adjoint_backward (checkpoint every k steps)."""
        self._check(native.load().dfr_adjoint_begin(self.ctx, int(checkpoint_every)))

    def adjoint_backward(self, seed: np.ndarray) -> np.ndarray:
        """seed (n_bodies, 7): dL/d(x_rb, q (w,x,y,z)) after the last recorded step.
        Returns (n_bodies, 6): dL/d(init_v_rb, init_omega_rb)."""
        sd = np.ascontiguousarray(seed, np.float64).reshape(self.n_bodies, 7)
        out = np.zeros((self.n_bodies, 6), np.float64)
        L = native.load()
        r = L.dfr_adjoint_backward(self.ctx, _dp(sd), _dp(out))
        if r == -5:
            msg = L.dfr_last_error(self.ctx)
            raise AdjointRangeError(f"DFR_ERR_RANGE: {msg.decode() if msg else ''}", out)
        self._check(r)
        return out

    def adjoint_start_gradient(self) -> np.ndarray:
        """(n_bodies, 13): dL/d(x_rb, q (w,x,y,z), v, omega) at the start of the
        recorded window, from the last adjoint_backward."""
        out = np.zeros((self.n_bodies, 13), np.float64)
        self._check(native.load().dfr_adjoint_start_gradient(self.ctx, _dp(out)))
        return out

    def state_save(self, slot: int = 0):
        """Keep the current state in device memory under `slot`."""
        self._check(native.load().dfr_state_save(self.ctx, int(slot)))

    def state_load(self, slot: int = 0):
        """Put back the state kept under `slot`."""
        self._check(native.load().dfr_state_load(self.ctx, int(slot)))

    def state_free(self, slot: int = 0):
        """Free the device memory of the state kept under `slot`."""
        self._check(native.load().dfr_state_free(self.ctx, int(slot)))

    def adjoint_arena_limit(self, nbytes: int = 0):
        """Test hook: bytes the adjoint tape arena may hand out (0: no limit);
        reaching it acts as running out of device memory."""
        self._check(native.load().dfr_adjoint_arena_limit(self.ctx, int(nbytes)))

    def adjoint_tape_bytes(self) -> int:
        """Device bytes the kept tapes of the current rollout occupy."""
        return int(self._check(native.load().dfr_adjoint_tape_bytes(self.ctx)))

    def adjoint_segment_memory(self, nbytes: int = -1):
        """Bound on one checkpoint segment's kept tapes in the backward pass
        (-1: free device memory; 0: re-run every step twice instead)."""
        self._check(native.load().dfr_adjoint_segment_memory(self.ctx, int(nbytes)))

    def adjoint_end(self):
        self._check(native.load().dfr_adjoint_end(self.ctx))

    def adjoint_iterations(self):
        L = native.load()
        n = int(L.dfr_adjoint_iterations(self.ctx, None, None, 0))
        iv = np.zeros(max(n, 1), np.int32)
        ip = np.zeros(max(n, 1), np.int32)
        L.dfr_adjoint_iterations(self.ctx, iv.ctypes.data_as(C.POINTER(C.c_int32)),
                                 ip.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return iv[:n].copy(), ip[:n].copy()

    def set_fixed_iterations(self, iv, ip):
        """Run the next len(iv) steps with exactly iv[s] / ip[s] iterations."""
        a = np.ascontiguousarray(iv, np.int32)
        b = np.ascontiguousarray(ip, np.int32)
        self._check(native.load().dfr_set_fixed_iterations(self.ctx, int(a.shape[0]),
                                                            a.ctypes.data_as(C.POINTER(C.c_int32)),
                                                            b.ctypes.data_as(C.POINTER(C.c_int32))))

    # -- learned warm start ----------------------------------------------------
    def mlp_capture(self, capacity: int, fraction: float = 1.0, seed: int = 0):
        """Record (features, converged (kappa_v, kappa)) pairs during the next steps."""
        self._check(native.load().dfr_mlp_capture(self.ctx, int(capacity), float(fraction), int(seed)))

    def mlp_dataset(self):
        L = native.load()
        n = int(self._check(L.dfr_mlp_dataset_size(self.ctx)))
        f = np.zeros((max(n, 1), 6), np.float32)
        t = np.zeros((max(n, 1), 2), np.float32)
        self._check(L.dfr_mlp_dataset(self.ctx, _fp(f), _fp(t), n))
        return f[:n], t[:n]

    def mlp_init_params(self, seed: int = 0):
        self._check(native.load().dfr_mlp_init_params(self.ctx, int(seed)))

    def mlp_set_params(self, params: np.ndarray, norm: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        q = np.ascontiguousarray(norm, np.float32)
        assert p.shape == (native.MLP_NP,) and q.shape == (16,)
        self._check(native.load().dfr_mlp_set_params(self.ctx, _fp(p), _fp(q)))

    def mlp_get_params(self):
        p = np.zeros(native.MLP_NP, np.float32)
        q = np.zeros(16, np.float32)
        self._check(native.load().dfr_mlp_get_params(self.ctx, _fp(p), _fp(q)))
        return p, q

    def mlp_train(self, iters: int, batch: int = 16384, lr: float = 2e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(int(iters))
        self._check(native.load().dfr_mlp_train(self.ctx, int(iters), int(batch), float(lr), int(seed), _dp(losses)))
        return losses

    def mlp_enable(self, on: bool = True):
        self._check(native.load().dfr_mlp_enable(self.ctx, int(on)))

    def mlp_predict(self, feat: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, 6)
        out = np.zeros((f.shape[0], 2), np.float32)
        self._check(native.load().dfr_mlp_predict(self.ctx, f.shape[0], _fp(f), _fp(out)))
        return out

    def mlp_loss_grad(self, feat: np.ndarray, targ: np.ndarray):
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(targ, np.float32).reshape(-1, 2)
        g = np.zeros(native.MLP_NP, np.float32)
        loss = np.zeros(1)
        self._check(native.load().dfr_mlp_loss_grad(self.ctx, f.shape[0], _fp(f), _fp(t), _fp(g), _dp(loss)))
        return float(loss[0]), g

    def mlp_input_grad(self, feat: np.ndarray, ybar: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, 6)
        y = np.ascontiguousarray(ybar, np.float32).reshape(-1, 2)
        out = np.zeros((f.shape[0], 6), np.float32)
        self._check(native.load().dfr_mlp_input_grad(self.ctx, f.shape[0], _fp(f), _fp(y), _fp(out)))
        return out

    # -- TimeStepDiffDFSPH-level getters -------------------------------------
    def get_boundary_model(self, i: int) -> BoundaryModel:
        return BoundaryModel(self, i)

    def is_trajectory_finish_callback(self) -> bool:
        return bool(self.stats()["trajectory_finished"])

    def get_step_count(self) -> int:
        return int(self.stats()["step_count"])
