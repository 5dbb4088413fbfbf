"""The pysplishsplash mirror on the CPU: its surface carries the names DiffFR's
drivers call (pySPlisHSPlasH/DiffDFSPHModule.cpp:44-102, BoundaryModelModule.cpp:
18-91, RigidBodyModule.cpp:19-45, SimulationModule.cpp:143-362), the target
quaternion follows Rx(a0) Ry(a1) Rz(a2) (SimulationDataDiffDFSPH.h:169-184),
and the unsupported inputs fail loudly.  Stepping is covered by the GPU test
of the same module."""
import math

import numpy as np
import pytest

from difffr_amd import pysplishsplash as sph
from difffr_amd.simulation import DfrError


def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_target_quaternion_is_the_xyz_rotation_product():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.uniform(-math.pi, math.pi, 3)
        q = sph.target_quaternion(a)
        assert abs(np.linalg.norm(q) - 1.0) < 1e-14
        R = _rx(a[0]) @ _ry(a[1]) @ _rz(a[2])
        np.testing.assert_allclose(sph._quat_matrix(q), R, atol=1e-14)
        # the transposed order would differ for generic angles
        assert np.abs(sph._quat_matrix(q) - _rz(a[2]) @ _ry(a[1]) @ _rx(a[0])).max() > 1e-6


def test_target_quaternion_special_cases():
    np.testing.assert_allclose(sph.target_quaternion([0, 0, 0]), [1, 0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(sph.target_quaternion([math.pi, 0, 0]), [0, 1, 0, 0], atol=1e-15)
    np.testing.assert_allclose(sph.target_quaternion([0, math.pi, 0]), [0, 0, 1, 0], atol=1e-15)
    h = math.sqrt(0.5)
    np.testing.assert_allclose(sph.target_quaternion([0, 0, math.pi / 2]), [h, 0, 0, h], atol=1e-15)


SURFACE = {
    sph.Exec.SimulatorBase: [
        "init", "initSimulation", "initSimulationWithDeferredInit", "runSimulation", "runNewTrajectory",
        "singleTimeStep", "timeStep", "timeStepNoGUI", "forwardFixedSteps", "runFixedTimeSteps", "reset",
        "cleanup", "setTimeStepCB", "setResetCB", "setGui", "getOutputPath", "getRigidBodyGradientManager",
        "setStateExportPath", "setValueBool", "setValueFloat", "setValueInt", "saveState",
        "loadStateWithRigidExisted", "STATE_EXPORT",
        "STATE_EXPORT_FPS", "STOP_AT"],
    sph.Simulation: ["getCurrent", "getTimeStep", "getBoundaryModel", "numberOfBoundaryModels",
                     "setGradientMode", "getGradientMode", "useRigidGradientManager"],
    sph.TimeStepDiffDFSPH: [
        "get_boundary_model", "get_loss", "set_loss", "get_loss_x", "set_loss_x", "get_loss_rotation",
        "set_loss_rotation", "get_lr", "set_lr", "set_init_v_rb", "set_init_omega_rb", "get_init_v_rb",
        "get_init_omega_rb", "get_target_x", "set_target_x", "get_target_angle_in_radian",
        "get_target_quaternion_vec4", "is_trajectory_finish_callback", "clear_all_callbacks",
        "is_in_new_trajectory", "set_custom_log_message", "get_custom_log_message", "get_step_count",
        "add_log", "reset_gradient", "get_num_1ring_fluid_particle"],
    sph.BoundaryModelAkinci2012: [
        "numberOfParticles", "getPosition", "getVelocity", "getVolume", "getRigidBodyObject", "getForce",
        "getTorque", "getPointVelocity", "get_position_rb", "get_quaternion_rb_vec4", "get_velocity_rb",
        "get_angular_velocity_rb", "set_velocity_rb", "set_angular_velocity_rb", "get_grad_x_to_v0",
        "get_grad_x_to_omega0", "get_grad_quaternion_to_v0", "get_grad_quaternion_to_omega0",
        "get_grad_v_to_v0", "get_grad_v_to_omega0", "get_grad_omega_to_v0", "get_grad_omega_to_omega0"],
    sph.RigidBodyObject: [
        "isDynamic", "isAnimated", "getMass", "getPosition", "setPosition", "getWorldSpacePosition",
        "getVelocity", "setVelocity", "getRotation", "setRotation", "getWorldSpaceRotation",
        "getAngularVelocity", "setAngularVelocity"],
    sph.TimeManager: ["getCurrent", "getTime", "getTimeStepSize"],
    sph.RigidBodyGradientManager: [
        "reset", "get_grad_net_force_to_vn", "get_grad_net_torque_to_omega_n", "get_grad_x_to_omega0",
        "get_grad_x_to_v0", "get_grad_q_to_v0", "get_grad_q_to_omega0"],
}


@pytest.mark.parametrize("cls", list(SURFACE), ids=lambda c: c.__name__)
def test_surface_names(cls):
    missing = [n for n in SURFACE[cls] if not hasattr(cls, n)]
    assert not missing, missing


def test_gui_is_construct_only():
    base = sph.Exec.SimulatorBase()
    gui = sph.GUI.Simulator_GUI_imgui(base)
    base.setGui(gui)
    assert gui.base is base


def test_unsupported_inputs_fail_loudly(tmp_path):
    base = sph.Exec.SimulatorBase()
    with pytest.raises(FileNotFoundError):
        base.init(sceneFile=str(tmp_path / "scene.json"), useGui=False, stopAt=1.0)
    with pytest.raises(NotImplementedError, match="scene JSON file"):
        base.init(sceneFile=str(tmp_path / "scene.xml"), useGui=False, stopAt=1.0)
    with pytest.raises(FileNotFoundError):
        base.init(sceneFile="sphere_in_box", stateFile=str(tmp_path / "state.bin"))
    with pytest.raises(ValueError):
        base.init(sceneFile="sphere_in_box", loadFluidPos=True)
    with pytest.raises(NotImplementedError, match="state export"):
        base.setValueBool(base.STATE_EXPORT, True)
    base.setValueBool(base.STATE_EXPORT, False)
    base.setValueFloat(base.STOP_AT, 0.25)
    assert base.getValueFloat(base.STOP_AT) == 0.25
    with pytest.raises(DfrError):  # before initSimulation
        base.getRigidBodyGradientManager()


def test_no_simulation_before_init_simulation():
    base = sph.Exec.SimulatorBase()
    with pytest.raises(DfrError):
        base.runSimulation()
    with pytest.raises(DfrError):
        base.singleTimeStep()
    if sph._current["base"] is None:
        with pytest.raises(DfrError):
            sph.Simulation.getCurrent()


def test_rigid_contact_solver_handle():
    """getBoundarySimulator().getRigidContactSolver(): None without
    useRigidContactSolver (the reference's null pointer), else an object whose
    reset() the drivers may call."""
    from difffr_amd import scenes
    sc = scenes.submerged_body(steps=1)
    assert sph._BoundarySimulator(sc).getRigidContactSolver() is None
    sc.params["use_rigid_contact"] = 1
    solver = sph._BoundarySimulator(sc).getRigidContactSolver()
    assert solver is not None and solver.reset() is None
