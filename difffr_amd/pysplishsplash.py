"""The `pysplishsplash` module surface that DiffFR's optimisation drivers use,
over one libdfr_hip.so context.

The reference registers these classes in pySPlisHSPlasH/main.cpp:47-70; the
drivers (e.g. experiments/rigid_body_trajectory_optimization/python/
gradient-based-optimize.py:90-358,423-525) use the names below with the same
arity and meaning, so such a driver runs with `import difffr_amd.pysplishsplash
as sph` in place of `import pysplishsplash as sph`, given a scene this library
can build:

    base = sph.Exec.SimulatorBase()
    base.init(sceneFile=scene, useGui=False, stopAt=...)   # a scenes.Scene
    base.initSimulation()
    sim = sph.Simulation.getCurrent()
    timestep = sim.getTimeStep()                           # TimeStepDiffDFSPH
    sim.setGradientMode(1)
    base.setTimeStepCB(callback)                           # after every step
    base.runSimulation()
    base.cleanup()

What differs, and why:
* `init(sceneFile=...)` takes a scene JSON file (read by `difffr_amd.sceneio`:
  its fluid blocks, OBJ bodies sampled on their surface, and the step's
  settings; CFL time stepping, surface tension, viscosity and emitters are
  not run and are listed in `Scene.unsupported`), a `difffr_amd.scenes.Scene`,
  or the name of a scene builder in `difffr_amd.scenes`.  State files the
  reference wrote (`stateFile`, `loadFluidPos[AndVel]`,
  `loadStateWithRigidExisted(path)`) are read; `--param` overrides are not
  (NotImplementedError), nor is state exported (`setValueBool(STATE_EXPORT,
  True)` raises).
* There is no GUI: `GUI.Simulator_GUI_imgui(base)` constructs an inert object
  and `runSimulation` runs the headless loop of SimulatorBase::runSimulation
  (SimulatorBase.cpp:649-683).  With `useGui=False` a negative stopAt is an
  error there and here; with `useGui=True` the loop runs until the callback
  ends it (the drivers call `quit()`), or until stopAt when it is positive.
* The RigidBodyGradientManager (scene parameter `use_gradient_manager`, the
  scene key useRigidGradientManager) holds the single-body entries (R, R) of
  its chain; with the rigid contact solver (useRigidContactSolver) it holds
  the entries between bodies too, which the contact solver's gradient terms
  fill (at most 16 bodies).  The exact adjoint does not cover the contact
  solver.
* A reference C++ `assert` or crash on a bad index is an IndexError here.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
from typing import Callable, Optional

import numpy as np

from . import native
from . import sceneio as _sceneio
from . import scenes as _scenes
from .simulation import BoundaryModel as _CoreBoundaryModel
from .simulation import DfrError
from .simulation import Simulation as _CoreSimulation

__all__ = ["Exec", "GUI", "Simulation", "TimeManager", "TimeStepDiffDFSPH", "BoundaryModelAkinci2012",
           "RigidBodyObject", "RigidBodyGradientManager", "Utilities", "DfrError"]

_current: dict = {"base": None}


def _core() -> _CoreSimulation:
    base = _current["base"]
    if base is None or base._core is None:
        raise DfrError("no current simulation: call SimulatorBase.initSimulation() first")
    return base._core


def _vec(v, n: int) -> np.ndarray:
    a = np.asarray(v, np.float64).reshape(-1)
    if a.shape[0] != n:
        raise ValueError(f"expected {n} values, got {a.shape[0]}")
    return a.copy()


def _quat_matrix(q) -> np.ndarray:
    """Rotation matrix of the unit quaternion q = (w, x, y, z)."""
    w, x, y, z = (float(c) for c in q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def _quat_mul(a, b) -> np.ndarray:
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz,
                     aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw])


def target_quaternion(angle) -> np.ndarray:
    """(w, x, y, z) of Rx(a0) * Ry(a1) * Rz(a2), the composition of three axis
    rotations (SimulationDataDiffDFSPH.h:169-184)."""
    a = np.asarray(angle, np.float64)
    qx = np.array([math.cos(a[0] / 2), math.sin(a[0] / 2), 0.0, 0.0])
    qy = np.array([math.cos(a[1] / 2), 0.0, math.sin(a[1] / 2), 0.0])
    qz = np.array([math.cos(a[2] / 2), 0.0, 0.0, math.sin(a[2] / 2)])
    return _quat_mul(_quat_mul(qx, qy), qz)


# ---------------------------------------------------------------------------
# RigidBodyObject (pySPlisHSPlasH/RigidBodyModule.cpp:19-45)
# ---------------------------------------------------------------------------
class RigidBodyObject:
    """The rigid body behind a boundary model.  Rotations cross this interface
    as Eigen coefficient vectors (x, y, z, w), as the reference's
    getRotation/setRotation lambdas do (RigidBodyModule.cpp:28-36)."""

    def __init__(self, core: _CoreSimulation, index: int):
        self._core = core
        self._index = index

    def _body(self):
        m = np.zeros(1)
        d = C.c_int(0)
        a = C.c_int(0)
        self._core._check(native.load().dfr_get_body(self._core.ctx, self._index, m.ctypes.data_as(
            C.POINTER(C.c_double)), C.byref(d), C.byref(a)))
        return float(m[0]), bool(d.value), bool(a.value)

    def _state(self):
        return self._core.rigid(self._index)

    def _set(self, **kw):
        s = self._state()
        s.update(kw)
        self._core.set_rigid(self._index, s["x"], s["q"], s["v"], s["omega"])

    def isDynamic(self) -> bool:
        return self._body()[1]

    def isAnimated(self) -> bool:
        return self._body()[2]

    def getMass(self) -> float:
        return self._body()[0]

    def getPosition(self) -> np.ndarray:
        return self._state()["x"].copy()

    def setPosition(self, x):
        self._set(x=_vec(x, 3))

    def getWorldSpacePosition(self) -> np.ndarray:
        return self.getPosition()

    def getVelocity(self) -> np.ndarray:
        return self._state()["v"].copy()

    def setVelocity(self, v):
        self._set(v=_vec(v, 3))

    def getAngularVelocity(self) -> np.ndarray:
        return self._state()["omega"].copy()

    def setAngularVelocity(self, w):
        self._set(omega=_vec(w, 3))

    def getRotation(self) -> np.ndarray:
        q = self._state()["q"]
        return np.array([q[1], q[2], q[3], q[0]])

    def setRotation(self, q_xyzw):
        c = _vec(q_xyzw, 4)
        self._set(q=np.array([c[3], c[0], c[1], c[2]]))

    def getWorldSpaceRotation(self) -> np.ndarray:
        return _quat_matrix(self._state()["q"])


# ---------------------------------------------------------------------------
# BoundaryModelAkinci2012 (pySPlisHSPlasH/BoundaryModelModule.cpp:18-91)
# ---------------------------------------------------------------------------
class BoundaryModelAkinci2012(_CoreBoundaryModel):
    """Per-body getters: the rigid state and forward-mode gradients of the base
    class plus the sample accessors and velocity setters."""

    def numberOfParticles(self) -> int:
        return int(self.sim.body_sizes[self.index])

    def _sample(self, i: int, name: str):
        n = self.numberOfParticles()
        if not 0 <= i < n:
            raise IndexError(f"boundary sample {i} out of range [0, {n})")
        return self.sim.body_field(self.index, name)[i]

    def getPosition(self, i: int) -> np.ndarray:
        return self._sample(i, "x").astype(np.float64)

    def getVelocity(self, i: int) -> np.ndarray:
        return self._sample(i, "v").astype(np.float64)

    def getVolume(self, i: int) -> float:
        return float(self._sample(i, "volume"))

    def getRigidBodyObject(self) -> RigidBodyObject:
        return RigidBodyObject(self.sim, self.index)

    def getPointVelocity(self, x) -> np.ndarray:
        """v_rb + omega x (x - x_rb) (BoundaryModel::getPointVelocity)."""
        s = self._rigid()
        return s["v"] + np.cross(s["omega"], _vec(x, 3) - s["x"])

    def set_velocity_rb(self, v):
        self.getRigidBodyObject().setVelocity(v)

    def set_angular_velocity_rb(self, w):
        self.getRigidBodyObject().setAngularVelocity(w)


# ---------------------------------------------------------------------------
# TimeStepDiffDFSPH (pySPlisHSPlasH/DiffDFSPHModule.cpp:44-102)
# ---------------------------------------------------------------------------
class TimeStepDiffDFSPH:
    """The differentiable time step's Python surface.  Loss values, the learning
    rate and the log message are host-side bookkeeping as in
    SimulationDataDiffDFSPH (learning rate 1 at start, SimulationDataDiffDFSPH.cpp:25);
    the initial rigid velocities and the gradients live in the library."""

    # parameter ids of the reference's ParameterObject (values of this module)
    ENABLE_COUNT_NEIGHBOR_DOF = 1001

    def __init__(self, base: "Exec.SimulatorBase"):
        self._base = base
        core = base._core
        self._loss = 0.0
        self._loss_x = 0.0
        self._loss_rotation = 0.0
        self._lr = 1.0
        self._log = ""
        self._finished = False
        self._new_trajectory = False
        self._count_dof = False
        sc = base._scene
        self._init_v = [np.asarray(b.v0, np.float64).copy() for b in sc.bodies]
        self._init_w = [np.asarray(b.w0, np.float64).copy() for b in sc.bodies]
        self._target_x = [np.asarray(b.target_x, np.float64).copy() for b in sc.bodies]
        self._target_angle = [np.asarray(b.target_angle, np.float64).copy() for b in sc.bodies]
        self._core = core

    def _body(self, i: int) -> int:
        if not 0 <= int(i) < self._core.n_bodies:
            raise IndexError(f"boundary model {i} out of range [0, {self._core.n_bodies})")
        return int(i)

    def get_boundary_model(self, i: int) -> BoundaryModelAkinci2012:
        return BoundaryModelAkinci2012(self._core, self._body(i))

    # losses and learning rate (TimeStepDiffDFSPH.cpp:2157-2195)
    def get_loss(self) -> float:
        return self._loss

    def set_loss(self, v: float):
        self._loss = float(v)

    def get_loss_x(self) -> float:
        return self._loss_x

    def set_loss_x(self, v: float):
        self._loss_x = float(v)

    def get_loss_rotation(self) -> float:
        return self._loss_rotation

    def set_loss_rotation(self, v: float):
        self._loss_rotation = float(v)

    def get_lr(self) -> float:
        return self._lr

    def set_lr(self, v: float):
        self._lr = float(v)

    # initial rigid velocities (TimeStepDiffDFSPH.cpp:2087-2125); beginStep
    # imposes them on the first steps (:379-429)
    def set_init_v_rb(self, i: int, v):
        k = self._body(i)
        self._init_v[k] = _vec(v, 3)
        self._core.set_init_velocity(k, self._init_v[k], self._init_w[k])

    def set_init_omega_rb(self, i: int, w):
        k = self._body(i)
        self._init_w[k] = _vec(w, 3)
        self._core.set_init_velocity(k, self._init_v[k], self._init_w[k])

    def get_init_v_rb(self, i: int) -> np.ndarray:
        return self._init_v[self._body(i)].copy()

    def get_init_omega_rb(self, i: int) -> np.ndarray:
        return self._init_w[self._body(i)].copy()

    # targets (TimeStepDiffDFSPH.cpp:2137-2155)
    def get_target_x(self, i: int) -> np.ndarray:
        return self._target_x[self._body(i)].copy()

    def set_target_x(self, i: int, x):
        self._target_x[self._body(i)] = _vec(x, 3)

    def get_target_angle_in_radian(self, i: int) -> np.ndarray:
        return self._target_angle[self._body(i)].copy()

    def get_target_quaternion_vec4(self, i: int) -> np.ndarray:
        return target_quaternion(self._target_angle[self._body(i)])

    # trajectory bookkeeping (endStep, TimeStepDiffDFSPH.cpp:432-490; :2197-2215)
    def is_trajectory_finish_callback(self) -> bool:
        return self._finished

    def clear_all_callbacks(self):
        self._finished = False

    def is_in_new_trajectory(self) -> bool:
        return self._new_trajectory

    def set_custom_log_message(self, s: str):
        self._log += str(s)  # appends, as the reference does (:2217-2220)

    def get_custom_log_message(self) -> str:
        return self._log

    def add_log(self, s: str):
        print(s, flush=True)

    def get_step_count(self) -> int:
        return self._core.get_step_count()

    def reset_gradient(self):
        self._core.reset_gradient()

    def get_num_1ring_fluid_particle(self) -> int:
        """(fluid particle, dynamic body) pairs with at least one neighbour sample
        of the body in the last step's lists, when neighbour-DOF counting is on
        (TimeStepDiffDFSPH.cpp:283-330, :542-543); 0 otherwise."""
        if not self._count_dof:
            return 0
        return int(self._core._check(native.load().dfr_count_1ring(self._core.ctx)))

    def setValueBool(self, pid: int, v: bool):
        if pid != self.ENABLE_COUNT_NEIGHBOR_DOF:
            raise KeyError(f"unknown parameter id {pid}")
        self._count_dof = bool(v)

    def getValueBool(self, pid: int) -> bool:
        if pid != self.ENABLE_COUNT_NEIGHBOR_DOF:
            raise KeyError(f"unknown parameter id {pid}")
        return self._count_dof

    def _after_step(self):
        st = self._core.stats()
        self._finished = bool(st["trajectory_finished"])
        if self._finished:
            self._log = ""  # endStep clears the custom message (:464)


# ---------------------------------------------------------------------------
# Simulation and TimeManager (SimulationModule.cpp:143-210, TimeModule.cpp:17-26)
# ---------------------------------------------------------------------------
class Simulation:
    """The current simulation: object getters and the gradient mode."""

    def __init__(self, base: "Exec.SimulatorBase"):
        self._base = base

    @staticmethod
    def getCurrent() -> "Simulation":
        base = _current["base"]
        if base is None or base._core is None:
            raise DfrError("no current simulation: call SimulatorBase.initSimulation() first")
        return base._simulation

    def getTimeStep(self) -> TimeStepDiffDFSPH:
        return self._base._timestep

    def getBoundaryModel(self, i: int) -> BoundaryModelAkinci2012:
        return self._base._timestep.get_boundary_model(i)

    def numberOfBoundaryModels(self) -> int:
        return self._base._core.n_bodies

    def numberOfFluidModels(self) -> int:
        return 1 if self._base._core.n > 0 else 0

    def setGradientMode(self, mode: int):
        if int(mode) not in (0, 1, 2):
            raise ValueError("gradient mode is 0 (complete), 1 (incomplete) or 2 (rigid only)")
        core = self._base._core
        core._check(native.load().dfr_set_gradient_mode(core.ctx, int(mode)))
        self._base._gradient_mode = int(mode)

    def getGradientMode(self) -> int:
        return self._base._gradient_mode

    def useRigidGradientManager(self) -> bool:
        return bool(int(self._base._core.params.get("use_gradient_manager", 0)))


class RigidBodyGradientManager:
    """SPH.RigidBodyGradientManager (SimulationModule.cpp:406-413): the chain
    of RigidBodyGradientManager.cpp:96-423 that replaces perform_chain_rule
    when the scene sets useRigidGradientManager.  Getters return copies, as
    pybind11 does for Eigen values."""

    def __init__(self, base: "Exec.SimulatorBase"):
        self._base = base

    def _entries(self, R: int, RR: int):
        core = self._base._require()
        n = core.n_bodies
        if not (0 <= int(R) < n and 0 <= int(RR) < n):
            raise IndexError(f"body indices ({R}, {RR}) out of range [0, {n})")
        return core.gradient_manager_raw(int(R), int(RR))

    def reset(self):
        """Jacobians back to (I, 0) and the handed sums to zero (also resets
        the boundary models' own Jacobians)."""
        core = self._base._require()
        core._check(native.load().dfr_reset_gradient(core.ctx))

    def get_grad_net_force_to_vn(self, R: int, RR: int) -> np.ndarray:
        return self._entries(R, RR)[1][0:9].reshape(3, 3).copy()

    def get_grad_net_torque_to_omega_n(self, R: int, RR: int) -> np.ndarray:
        return self._entries(R, RR)[1][27:36].reshape(3, 3).copy()

    def get_grad_x_to_v0(self, R: int, RR: int) -> np.ndarray:
        return self._entries(R, RR)[0][0:9].reshape(3, 3).copy()

    def get_grad_x_to_omega0(self, R: int, RR: int) -> np.ndarray:
        return self._entries(R, RR)[0][9:18].reshape(3, 3).copy()

    def get_grad_q_to_v0(self, R: int, RR: int) -> np.ndarray:
        return self._entries(R, RR)[0][54:66].reshape(4, 3).copy()

    def get_grad_q_to_omega0(self, R: int, RR: int) -> np.ndarray:
        return self._entries(R, RR)[0][66:78].reshape(4, 3).copy()


class TimeManager:
    @staticmethod
    def getCurrent() -> "TimeManager":
        _core()
        return TimeManager()

    def getTime(self) -> float:
        return float(_core().stats()["time"])

    def getTimeStepSize(self) -> float:
        """The step size the next step starts with (the CFL rule changes it)."""
        return float(_core().stats()["dt"])


# ---------------------------------------------------------------------------
# Exec.SimulatorBase (SimulationModule.cpp:217-362, SimulatorBase.cpp:649-1200)
# ---------------------------------------------------------------------------
class _SimulatorBase:
    # parameter ids (ParameterObjectModule.cpp:17-29 and the static ids of
    # SimulatorBase); values of this module
    STOP_AT = 1
    STATE_EXPORT = 2
    STATE_EXPORT_FPS = 3
    PAUSE = 4

    def __init__(self):
        self._scene = None
        self._core: Optional[_CoreSimulation] = None
        self._timestep: Optional[TimeStepDiffDFSPH] = None
        self._simulation: Optional[Simulation] = None
        self._cb: Optional[Callable[[], None]] = None
        self._reset_cb: Optional[Callable[[], None]] = None
        self._stop_at = -1.0
        self._use_gui = True
        self._output = ""
        self._gui = None
        self._gradient_mode = 1
        self._values = {self.STATE_EXPORT: False, self.STATE_EXPORT_FPS: 25.0, self.PAUSE: False}
        self._states: dict = {}  # saveState names -> device state slots, least recently used first

    def init(self, sceneFile="", programName="pySPlisHSPlasH", useCache=True, stateFile="",
             loadFluidPos=False, loadFluidPosAndVel=False, outputDir="", initialPause=True, useGui=True,
             stopAt=-1.0, param=""):
        """SimulatorBase::init through py_init_simulator (SimulationModule.cpp:32-68)."""
        if isinstance(sceneFile, _scenes.Scene):
            scene = sceneFile
        elif isinstance(sceneFile, str) and sceneFile.endswith(".json"):
            if not os.path.isfile(sceneFile):
                raise FileNotFoundError(sceneFile)
            scene = _sceneio.load_scene(sceneFile)
        elif isinstance(sceneFile, str) and sceneFile and not os.path.splitext(sceneFile)[1] \
                and callable(getattr(_scenes, sceneFile, None)):
            scene = getattr(_scenes, sceneFile)()
        else:
            raise NotImplementedError(
                "expected a scene JSON file, a difffr_amd.scenes.Scene or the name of a builder in "
                f"difffr_amd.scenes (got {sceneFile!r})")
        if (loadFluidPos or loadFluidPosAndVel) and not stateFile:
            raise ValueError("loadFluidPos / loadFluidPosAndVel need a stateFile")
        if stateFile and not os.path.isfile(stateFile if stateFile.endswith(".bin") else stateFile + ".bin"):
            raise FileNotFoundError(stateFile)
        self._state_file = stateFile
        self._load_fluid = "posvel" if loadFluidPosAndVel else ("pos" if loadFluidPos else "")
        if param:
            raise NotImplementedError("--param overrides are not supported; set scene.params instead")
        self._scene = scene
        self._stop_at = float(stopAt)
        self._use_gui = bool(useGui)
        self._gradient_mode = int(scene.params.get("gradient_mode", 1))
        self._output = outputDir or os.path.join(os.getcwd(), "output", scene.name)

    def initSimulation(self):
        if self._scene is None:
            raise DfrError("SimulatorBase.init must run before initSimulation")
        if self._core is not None:
            self._core.close()
        self._states = {}
        self._core = _CoreSimulation(self._scene)
        self._timestep = TimeStepDiffDFSPH(self)
        self._simulation = Simulation(self)
        self._manager = RigidBodyGradientManager(self)
        _current["base"] = self
        if getattr(self, "_state_file", ""):
            # SimulatorBase::initSimulation: loadFluidParticlePositions[AndVelocities] or loadState
            if self._load_fluid:
                self._apply_state_file(self._state_file, fluid_only=True,
                                       velocities=self._load_fluid == "posvel")
            else:
                self._apply_state_file(self._state_file)

    def _apply_state_file(self, path: str, fluid_only: bool = False, velocities: bool = True):
        """Fluid fields from the state's BGEO file (SimulatorBase::readFluidParticlesState,
        SimulatorBase.cpp:2545-2600: matched by particle id) and, unless
        fluid_only, the dynamic bodies' x, q, v, omega and the time from the
        .bin (SimulatorBase::loadStateWithRigidExisted, :2124-2290)."""
        core = self._require()
        bin_path = path if path.endswith(".bin") else path + ".bin"
        dyn = [k for k, b in enumerate(self._scene.bodies) if b.dynamic]
        st = _sceneio.read_state(bin_path, 0 if fluid_only else len(dyn))
        if not st["fluids"]:
            raise FileNotFoundError(f"no fluid BGEO file beside {bin_path}")
        fl = next(iter(st["fluids"].values()))
        n = fl["position"].shape[0]
        if n != core.n:
            raise ValueError(f"{bin_path}: {n} fluid particles, the scene has {core.n}")
        ids = fl["id"].astype(np.int64) if "id" in fl else np.arange(n)
        if not np.array_equal(np.sort(ids), np.arange(n)):
            raise ValueError(f"{bin_path}: particle ids are not 0..{n - 1}")

        def by_id(a):
            out = np.empty_like(a)
            out[ids] = a
            return out

        core.set_field("x", by_id(fl["position"]))
        core.set_field("v", by_id(fl["velocity"]) if velocities and "velocity" in fl
                       else np.zeros((n, 3), np.float32))
        if fluid_only:
            return
        for name in ("kappa", "kappa_v"):
            if name in fl:
                core.set_field(name, by_id(fl[name]))
        for k, r in zip(dyn, st["rigid"]):
            core.set_rigid(k, r["x"], r["q"], r["v"], r["omega"])
        if st["time"] is not None:
            m = re.fullmatch(r"state_(\d+)", os.path.basename(bin_path[:-4]))
            core.set_time(st["time"], int(m.group(1)) if m else core.get_step_count())

    def initSimulationWithDeferredInit(self):
        self.initSimulation()

    def _require(self) -> _CoreSimulation:
        if self._core is None:
            raise DfrError("no simulation: call initSimulation() first")
        return self._core

    def setGui(self, gui):
        self._gui = gui

    def getOutputPath(self) -> str:
        return self._output

    def setTimeStepCB(self, fn: Optional[Callable[[], None]]):
        self._cb = fn

    def setResetCB(self, fn: Optional[Callable[[], None]]):
        self._reset_cb = fn

    def setStateExportPath(self, path: str):
        self._values["state_export_path"] = str(path)

    def setValueBool(self, pid: int, v: bool):
        if pid == self.STATE_EXPORT and v:
            raise NotImplementedError("state export is not supported by this library")
        if pid not in (self.STATE_EXPORT, self.PAUSE):
            raise KeyError(f"unknown bool parameter id {pid}")
        self._values[pid] = bool(v)

    def setValueFloat(self, pid: int, v: float):
        if pid == self.STOP_AT:
            self._stop_at = float(v)
        elif pid == self.STATE_EXPORT_FPS:
            self._values[pid] = float(v)
        else:
            raise KeyError(f"unknown float parameter id {pid}")

    def setValueInt(self, pid: int, v: int):
        raise KeyError(f"unknown int parameter id {pid}")

    def getValueFloat(self, pid: int) -> float:
        if pid == self.STOP_AT:
            return self._stop_at
        return float(self._values[pid])

    def getValueBool(self, pid: int) -> bool:
        return bool(self._values[pid])

    # states kept in device memory at once; saving more evicts the least recently used
    MAX_KEPT_STATES = 16

    def saveState(self, stateFileDir: str = "") -> str:
        """Keeps the current state (SimulatorBase::saveState,
        SimulatorBase.cpp:1899-1960) in device memory and returns the name the
        reference gives its file, <dir>/state_<step> (without ".bin",
        SimulatorBase.cpp:1905-1920); no file is written.
        loadStateWithRigidExisted accepts that name, with or without ".bin".
        At most MAX_KEPT_STATES states are kept: saving another one frees the
        least recently saved or loaded one (a per-step loop such as
        experiments/others/python/finite-difference.py:55-110 reloads only the
        state of its current step)."""
        core = self._require()
        step = core.get_step_count()
        d = stateFileDir or os.path.join(self._output, "state")
        name = os.path.normpath(os.path.join(d, f"state_{step}"))
        if name in self._states:
            slot = self._states.pop(name)
        elif len(self._states) < self.MAX_KEPT_STATES:
            slot = len(self._states)
        else:
            old = next(iter(self._states))
            slot = self._states.pop(old)
            core.state_free(slot)
        core.state_save(slot)
        self._states[name] = slot
        return name

    def loadStateWithRigidExisted(self, stateFile: str):
        """Restores a state kept by saveState (SimulatorBase::loadStateWithRigidExisted,
        SimulatorBase.cpp:2124-2290): fluid positions, velocities and stiffness
        fields, boundary samples, rigid x, q, v, omega, time and step count.
        Unlike the reference, which leaves them as the last step made them, the
        forward-mode rigid Jacobians are restored too, so a perturbed step
        that is undone leaves no trace in the gradient chain.  A state file
        the reference wrote (".bin" + "_particle_<id>.bgeo") is read instead
        when no kept state has that name: fluid fields, the dynamic bodies'
        x, q, v, omega and, when its name carries it, the time."""
        core = self._require()
        name = stateFile[:-4] if stateFile.endswith(".bin") else stateFile
        name = os.path.normpath(name)
        if name not in self._states:
            if os.path.isfile(name + ".bin"):  # a state file the reference wrote
                self._apply_state_file(name + ".bin")
                return
            raise FileNotFoundError(
                f"{stateFile!r} is neither a state file nor one of the last {self.MAX_KEPT_STATES} "
                "states kept by saveState")
        slot = self._states.pop(name)
        self._states[name] = slot  # most recently used
        core.state_load(slot)

    def getRigidBodyGradientManager(self) -> RigidBodyGradientManager:
        self._require()
        return self._manager

    def getBoundarySimulator(self) -> "_BoundarySimulator":
        """SimulatorBase::getBoundarySimulator (SimulationModule.cpp:383)."""
        self._require()
        return _BoundarySimulator(self._scene)

    # -- stepping -------------------------------------------------------------
    def _one_step(self):
        core = self._require()
        core.step(1)
        self._timestep._after_step()
        if self._cb is not None:
            self._cb()

    def timeStepNoGUI(self) -> bool:
        """One step then the callback, unless past stopAt (SimulatorBase.cpp:1142-1200)."""
        core = self._require()
        if self._stop_at > 0.0 and self._stop_at < float(core.stats()["time"]):
            return False
        self._one_step()
        return True

    def singleTimeStep(self):
        self._one_step()

    def timeStep(self):
        if not self._values[self.PAUSE]:
            self._one_step()

    def forwardFixedSteps(self, n: int):
        for _ in range(int(n)):
            self._one_step()

    def runFixedTimeSteps(self, n: int):
        self.forwardFixedSteps(n)

    def runSimulation(self):
        """The headless loop of SimulatorBase::runSimulation (SimulatorBase.cpp:649-683)."""
        self._require()
        if not self._use_gui and self._stop_at < 0.0:
            raise DfrError("StopAt parameter must be set when starting without GUI.")
        while self.timeStepNoGUI():
            pass

    def runNewTrajectory(self):
        """Reset, then step until the trajectory ends (SimulatorBase.cpp:1093-1111)."""
        ts = self._timestep
        ts._new_trajectory = True
        try:
            self.reset()
            while True:
                self.singleTimeStep()
                if bool(self._core.stats()["trajectory_finished"]):
                    break
        finally:
            ts._new_trajectory = False

    def reset(self):
        """Fluid, bodies, time and gradients back to their initial state
        (SimulatorBase::reset, SimulatorBase.cpp:887-951)."""
        self._require().reset()
        if self._reset_cb is not None:
            self._reset_cb()

    def cleanup(self):
        if self._core is not None:
            self._core.close()
        self._core = None
        if _current["base"] is self:
            _current["base"] = None


class _RigidContactSolver:
    """RigidContactSolver of a scene with useRigidContactSolver: the penalty
    form runs inside every step; reset() (RigidContactSolver.cpp:1080-1087)
    clears per-sample state that the step recomputes anyway."""

    def reset(self):
        return None


class _BoundarySimulator:
    def __init__(self, scene):
        self._contact = bool(scene.params.get("use_rigid_contact", 0))

    def getRigidContactSolver(self) -> Optional[_RigidContactSolver]:
        """The contact solver, or None when the scene does not use it (the
        reference returns a null pointer, RigidBody3dBoundarySimulator.cpp:370-380)."""
        return _RigidContactSolver() if self._contact else None


class Exec:
    SimulatorBase = _SimulatorBase


class _GUI:
    def __init__(self, base):
        self.base = base


class GUI:
    """No GUI: the constructor keeps the simulator and nothing else."""
    Simulator_GUI_imgui = _GUI


class _Timing:
    @staticmethod
    def printAverageTimes():
        """Mean kernel-class times (ms) when kernel timing is on."""
        t = _core().kernel_times()
        for name, (avg, cnt) in t.items():
            if cnt:
                print(f"{name}: {avg:.4f} ms ({cnt} calls)", flush=True)

    @staticmethod
    def printTimeSums():
        t = _core().kernel_times()
        for name, (avg, cnt) in t.items():
            if cnt:
                print(f"{name}: {avg * cnt:.3f} ms total", flush=True)


class Utilities:
    Timing = _Timing
