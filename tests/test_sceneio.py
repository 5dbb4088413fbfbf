"""Scene and state files of the reference (difffr_amd.sceneio), on the CPU.

Fixtures: tests/golden/reference_files (the experiments' scene JSONs, the
bottle-flip experiment's meshes and saved state; copied by
tests/golden/copy_reference_files.py).  Pins:
* scene keys and defaults against the files' own values (SceneLoader.cpp:40-360),
  the contact solver's keys included;
* the fluid-block layout (SimulatorBase::createFluidBlocks) against the particle
  counts of states the reference saved for the same scenes (billiards: 87,374,
  from state_17_particle_Fluid.bgeo's header; bottle flip: 5,904) and the
  lattice geometry of the three dense modes;
* regular triangle sampling against its closed-form counts on a square and
  the spacing it promises; Poisson disk sampling against its minimum distance
  and coverage;
* the BGEO reader against an independent writer of the format and the
  fixture's own header;
* the state reader's rigid records against the bottle-flip scene's body.
"""
import gzip
import math
import os
import struct

import numpy as np
import pytest

from difffr_amd import sceneio

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = os.path.join(HERE, "golden", "reference_files")
SCENES = os.path.join(FILES, "scenes")
BOTTLE = os.path.join(FILES, "bottle_flip_robot")


def _scene(name):
    return sceneio.read_scene_file(os.path.join(SCENES, name))


def test_scene_keys_and_defaults():
    sf = _scene("diff-stone-skipping.json")
    c = sf.config
    assert c["particleRadius"] == 0.015 and c["targetTime"] == 0.18 and c["uniformAccelerateRBTime"] == 0.05
    assert c["maxError"] == 0.05 and c["maxErrorV"] == 0.1 and c["gradientMode"] == 1
    assert c["timeStepSize"] == 0.001  # not in the file: the loader's default (SceneLoader.cpp:45)
    wall, stone = sf.rigid_bodies
    assert wall.is_wall and not wall.dynamic and wall.density == 1000.0
    np.testing.assert_array_equal(wall.scale, [3.5, 4, 2])
    assert stone.dynamic and stone.density == 2000
    np.testing.assert_array_equal(stone.init_velocity, [25, -16, 0])
    np.testing.assert_array_equal(stone.init_angular_velocity, [-3, 20, 6])
    np.testing.assert_array_equal(stone.target_x, [1.7, 1.6, 0.0])
    np.testing.assert_array_equal(stone.quaternion, [1, 0, 0, 0])  # angle 0
    (blk,) = sf.fluid_blocks
    np.testing.assert_array_equal(blk.box_min, [-1.5, 0.0, -0.75])
    np.testing.assert_array_equal(blk.box_max, [1.5, 1.5, 0.75])
    assert sf.unsupported == []
    # the CFL keys and the fluid's material: Akinci 2013 surface tension from
    # "Materials", the FluidModel's default Standard viscosity 0.01 (the
    # "viscosityMethod" key of the fluid block is not read by the reference)
    assert (c["cflMethod"], c["cflFactor"], c["cflMaxTimeStepSize"], c["cflMinTimeStepSize"]) == (1, 0.5, 0.005, 0.0001)
    m = sceneio.fluid_material(sf)
    assert (m["surface_tension_method"], m["surface_tension"], m["surface_tension_boundary"]) == (2, 0.5, 0.0)
    assert (m["viscosity_method"], m["viscosity"], m["viscosity_boundary"]) == (1, 0.01, 0.0)
    b = _scene("billiards-on-water-2balls.json").config
    assert b["useRigidGradientManager"] is True and b["useReleaseRigidBodyMode"] is True
    d = _scene("diff-bottle-model-360degree.json")
    assert d.config["uniformAccelerateRBTime"] == 0.25 and d.config["targetTime"] == 0.75
    assert [rb.dynamic for rb in d.rigid_bodies] == [False, True]


def test_every_reference_scene_parses():
    names = sorted(os.listdir(SCENES))
    assert len(names) == 8
    for n in names:
        sf = _scene(n)
        assert sf.rigid_bodies and (sf.fluid_blocks or sf.fluid_models), n
        assert all(rb.geometry_file.endswith(".obj") for rb in sf.rigid_bodies)


def test_contact_scenes_switch_the_contact_solver_on():
    """billiards-on-water-2balls and diff-bottle-model-collide set
    useRigidContactSolver with their own stiffness (and friction); the other
    keys keep SceneLoader.cpp:58-71's defaults; the contact solver is run, so
    nothing is reported unsupported for it."""
    for name, beta, mu in (("billiards-on-water-2balls.json", 20000.0, 0.0),
                           ("diff-bottle-model-collide.json", 100000.0, 0.7)):
        sf = _scene(name)
        c = sf.config
        assert c["useRigidContactSolver"] is True
        assert c["rigidContactBeta"] == beta and c["rigidContactFrictionCoeff"] == mu
        assert c["rigidContactGamma"] == 0.7 and c["rigidContactSupportRadiusFactor"] == 4.0
        assert not any("Contact" in u for u in sf.unsupported), sf.unsupported
    assert _scene("diff-stone-skipping.json").config["useRigidContactSolver"] is False


def test_fluid_blocks_reproduce_the_reference_state_counts():
    # billiards: the reference's saved state of this scene holds 87,374 particles
    (blk,) = _scene("billiards-on-water-2balls.json").fluid_blocks
    r = _scene("billiards-on-water-2balls.json").config["particleRadius"]
    assert sceneio.fluid_block_positions(blk, r).shape[0] == 87374
    # bottle flip: against the fixture state file itself
    sf = _scene("diff-bottle-model-360degree.json")
    x = sceneio.fluid_block_positions(sf.fluid_blocks[0], sf.config["particleRadius"])
    st = sceneio.read_state(os.path.join(BOTTLE, "state", "state_1.365999.bin"), 1)
    assert x.shape[0] == st["fluids"]["Fluid"]["position"].shape[0] == 5904


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_fluid_block_lattice_geometry(mode):
    r = 0.025
    blk = sceneio.FluidBlockData(np.array([0.0, 0.0, 0.0]), np.array([0.5, 0.4, 0.3]), mode=mode)
    x = sceneio.fluid_block_positions(blk, r)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(x).query(x, 2)
    nn = d[:, 1]
    if mode == 0:  # regular: spacing 2r
        assert np.allclose(nn, 2 * r)
        assert x.shape[0] == (round(0.5 / (2 * r)) - 1) * (round(0.4 / (2 * r)) - 1) * (round(0.3 / (2 * r)) - 1)
        np.testing.assert_allclose(x.min(0), [2 * r] * 3)
        return
    # interior number density against the stackings' closed forms: mode 1 puts
    # square 2r layers sqrt(3) r apart (1.155 x the regular 1/(2r)^3), mode 2
    # is the close packing (sqrt(2) x)
    big = sceneio.FluidBlockData(np.zeros(3), np.full(3, 3.0), mode=mode)
    x = sceneio.fluid_block_positions(big, r)
    inside = np.all((x > 0.5) & (x < 2.5), axis=1).sum()
    ratio = inside * (2 * r) ** 3 / 2.0 ** 3
    assert ratio == pytest.approx({1: 2 / math.sqrt(3), 2: math.sqrt(2)}[mode], rel=0.03)


def test_regular_triangle_sampling_of_a_square():
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    F = np.array([[0, 1, 2], [0, 2, 3]])
    d = 0.1
    s = sceneio.regular_triangle_sampling(V, F, d)
    assert np.all(np.abs(s[:, 2]) < 1e-15)
    # 4 vertices; 5 unique edges: 4 of length 1 (9 interior points each) and the diagonal
    # (ceil(sqrt 2 / 0.1) + 1 = 16 points, 14 interior)
    edges = 4 * 9 + 14
    # faces: base = the diagonal, height sqrt(2)/2 -> ceil(7.07) + 1 = 9 lines, 7 interior,
    # line i of length (1 - i/8) sqrt 2 -> ceil(len / d) + 1 points, interior ones
    face = sum(max(math.ceil((1 - i / 8) * math.sqrt(2) / d) + 1 - 2, 0) for i in range(1, 8))
    assert s.shape[0] == 4 + edges + 2 * face
    # the promised spacing: every point of the square lies within d of a sample
    g = np.stack(np.meshgrid(np.linspace(0, 1, 41), np.linspace(0, 1, 41)), -1).reshape(-1, 2)
    g = np.concatenate([g, np.zeros((g.shape[0], 1))], 1)
    from scipy.spatial import cKDTree
    assert cKDTree(s).query(g)[0].max() <= d


def test_poisson_disk_sampling_distance_and_coverage():
    V, F = sceneio.read_obj(os.path.join(BOTTLE, "models", "UnitBox.obj"), (0.3, 0.2, 0.25))
    r = 0.02
    s = sceneio.poisson_disk_sampling(V, F, r, seed=1)
    from scipy.spatial import cKDTree
    t = cKDTree(s)
    d, _ = t.query(s, 2)
    # on one face the distance is Euclidean: no two samples closer than r there
    for axis in range(3):
        for side in (V[:, axis].min(), V[:, axis].max()):
            on = s[np.abs(s[:, axis] - side) < 1e-9]
            if on.shape[0] > 1:
                dd, _ = cKDTree(on).query(on, 2)
                assert dd[:, 1].min() >= r * (1 - 1e-9)
    # maximal: random surface points are all within 2r of a sample
    rng = np.random.default_rng(0)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    tri = rng.choice(F.shape[0], 4000, p=area / area.sum())
    u, v = rng.random(4000), rng.random(4000)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    p = a[tri] + u[:, None] * (b[tri] - a[tri]) + v[:, None] * (c[tri] - a[tri])
    assert t.query(p)[0].max() < 2 * r
    # reproducible for a seed
    np.testing.assert_array_equal(s, sceneio.poisson_disk_sampling(V, F, r, seed=1))


def _write_bgeo(path, pos, attrs):
    """A BGEO version 5 file as partio's writer lays it out (independent writer for the test)."""
    n = pos.shape[0]
    out = bytearray(b"BgeoV")
    out += struct.pack(">9i", 5, n, 0, 0, 0, len(attrs), 0, 0, 0)
    for name, kind, arr in attrs:
        size = 1 if arr.ndim == 1 else arr.shape[1]
        out += struct.pack(">H", len(name)) + name.encode() + struct.pack(">Hi", size, {"f": 0, "i": 1, "v": 5}[kind])
        out += b"\0" * (4 * size)
    for i in range(n):
        out += struct.pack(">4f", *pos[i], 1.0)
        for name, kind, arr in attrs:
            vals = np.atleast_1d(arr[i])
            out += struct.pack(">%d%s" % (vals.size, "i" if kind == "i" else "f"), *vals.tolist())
    out += struct.pack(">i", 0)  # no detail attributes' values follow with nAttrib = 0
    with open(path, "wb") as f:
        f.write(gzip.compress(bytes(out)))


def test_bgeo_reader(tmp_path):
    rng = np.random.default_rng(2)
    pos = rng.normal(size=(17, 3)).astype(np.float32)
    vel = rng.normal(size=(17, 3)).astype(np.float32)
    kap = rng.normal(size=17).astype(np.float32)
    ids = rng.permutation(17).astype(np.int32)
    p = str(tmp_path / "t.bgeo")
    _write_bgeo(p, pos, [("id", "i", ids), ("velocity", "v", vel), ("kappa", "f", kap)])
    d = sceneio.read_bgeo(p)
    np.testing.assert_array_equal(d["position"], pos)
    np.testing.assert_array_equal(d["velocity"], vel)
    np.testing.assert_array_equal(d["kappa"], kap)
    np.testing.assert_array_equal(d["id"], ids)
    # the reference's own file: 5,904 particles, ids a permutation, inside the wall box
    f = sceneio.read_bgeo(os.path.join(BOTTLE, "state", "state_1.365999_particle_Fluid.bgeo"))
    assert f["position"].shape == (5904, 3) and f["velocity"].shape == (5904, 3)
    assert sorted(f["id"].tolist()) == list(range(5904))
    assert {"kappa", "kappa_v"} <= set(f)
    # the fluid fills the bottle: inside the scaled mesh's box at the bottle's pose
    sf = sceneio.read_scene_file(os.path.join(BOTTLE, "scene", "diff-bottle-model-360degree.json"))
    bottle = sf.rigid_bodies[1]
    V, _ = sceneio.read_obj(os.path.join(BOTTLE, "models", "bottle.obj"), bottle.scale)
    lo, hi = V.min(0) + bottle.translation, V.max(0) + bottle.translation
    assert np.all(f["position"] > lo) and np.all(f["position"] < hi)


def test_state_rigid_record_and_time():
    st = sceneio.read_state(os.path.join(BOTTLE, "state", "state_1.365999.bin"), 1)
    assert st["time"] == pytest.approx(1.365999)
    assert len(st["md5"]) == 32 and st["frame"] > 0
    (rec,) = st["rigid"]
    sf = sceneio.read_scene_file(os.path.join(BOTTLE, "scene", "diff-bottle-model-360degree.json"))
    bottle = sf.rigid_bodies[1]
    # the bottle rests at its scene pose (q stored as (x, y, z, w), returned as (w, x, y, z))
    np.testing.assert_allclose(rec["x"], bottle.translation, atol=1e-6)
    np.testing.assert_allclose(rec["q"], bottle.quaternion, atol=1e-6)
    np.testing.assert_allclose(rec["v"], 0.0, atol=1e-6)
