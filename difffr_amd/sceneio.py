"""The reference's scene files and state files, read into this library's scene
description (host side, numpy).

* Scene JSON (SPlisHSPlasH/Utilities/SceneLoader.cpp:40-360): the
  "Configuration" keys of the differentiable step (DiffFR's targetTime,
  uniformAccelerateRBTime, gradientMode, useRigidGradientManager,
  useReleaseRigidBodyMode and the solver settings), "RigidBodies" (mesh,
  pose, scale, density, dynamic/wall flags, initial and target state) and
  "FluidBlocks" / "FluidModels", the CFL step-size keys and the fluid's
  "Materials" entry (surface tension, viscosity; without one the fluid keeps
  the reference's default Standard viscosity 0.01).  Keys of parts this
  library does not run (other force methods, boundary viscosity, the
  iteration-count step-size rule, emitters, articulated systems, rigid
  contact) are listed in `Scene.unsupported`.
* Fluid blocks are laid out by SimulatorBase::createFluidBlocks
  (Simulator/SimulatorBase.cpp:1638-1722, dense modes 0, 1 and 2).
* Rigid bodies are OBJ meshes (scaled as loadObj does,
  RigidBody3dBoundarySimulator.cpp:32-66) sampled on their surface
  (RigidBody3dBoundarySimulator.cpp:68-248): regular triangle sampling
  (RegularTriangleSampling.cpp) at spacing 1.5 r, or Poisson disk sampling
  (PoissonDiskSampling.cpp) with minimum distance r, 10 trials and the
  curvature-corrected distance.  The reference seeds its Poisson sampler from
  std::random_device, so its sample sets differ run to run; here the
  generator is seeded and the set is reproducible.
* State files (SimulatorBase::saveState, Simulator/SimulatorBase.cpp:1899-1990):
  the partio BGEO fluid file (extern/partio/src/lib/io/BGEO.cpp: gzip,
  "BgeoV" version 5, big-endian) and, from the ".bin", its header and the
  trailing rigid-body records (x, q as (x, y, z, w), v, omega per dynamic or
  animated body, in body order).  The parameter block between them is written
  positionally by GenericParameters, which the reference does not vendor, so
  the simulated time is taken from a file name that carries it
  ("state_<time>") or given by the caller.
"""
from __future__ import annotations

import dataclasses
import gzip
import json
import math
import os
import re
import struct
from typing import List, Optional

import numpy as np

from . import scenes

# Configuration keys the library runs, with the reference's defaults
# (SceneLoader.cpp:45-87, TimeStep.cpp:24-64, TimeStepDiffDFSPH.cpp:90-111,
# FluidModel.cpp:45).
_CONFIG_DEFAULTS = {
    "timeStepSize": 0.001, "particleRadius": 0.025, "density0": 1000.0, "gravitation": [0.0, -9.81, 0.0],
    "maxIterations": 100, "maxError": 0.01, "maxIterationsV": 100, "maxErrorV": 0.1, "minIterations": 2,
    "enableDivergenceSolver": True, "targetTime": 0.8, "uniformAccelerateRBTime": 0.0, "gradientMode": 1,
    "useRigidGradientManager": False, "useReleaseRigidBodyMode": False, "useRigidContactSolver": False,
    "sim2D": False, "cflMethod": 1, "boundaryHandlingMethod": 0, "simulationMethod": 4,
    "cflFactor": 0.5, "cflMinTimeStepSize": 0.0001, "cflMaxTimeStepSize": 0.005,
    "rigidContactBeta": 1.0, "rigidContactGamma": 0.7, "rigidContactSupportRadiusFactor": 4.0,
    "rigidContactFrictionCoeff": 0.0,
}
# Fluid parameters a "Materials" entry with the fluid's id may set (read by
# SceneLoader::readParameterObject into the FluidModel), at the FluidModel's
# defaults: FluidModel::init sets the Standard viscosity (FluidModel.cpp:96-101,
# ViscosityBase.cpp:12), no surface tension (FluidModel.cpp:52,
# SurfaceTensionBase.cpp:12-13).
_MATERIAL_DEFAULTS = {
    "viscosityMethod": 1, "viscosity": 0.01, "viscosityBoundary": 0.0,
    "surfaceTensionMethod": 0, "surfaceTension": 0.05, "surfaceTensionBoundary": 0.01,
    "vorticityMethod": 0, "dragMethod": 0, "elasticityMethod": 0,
}
# Keys read but not run: the run differs from the reference's while they are on.
_UNSUPPORTED_IF = {
    "cflMethod": lambda v: int(v) not in (0, 1),     # 2: iteration-count step-size control (Simulation.cpp:528-539)
    "sim2D": bool,
    "boundaryHandlingMethod": lambda v: int(v) != 0,  # Akinci2012 only
}


@dataclasses.dataclass
class RigidBodyData:
    """One "RigidBodies" entry with the reference's defaults (SceneLoader.cpp:93-180)."""
    geometry_file: str = ""
    particle_file: str = ""
    translation: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    rotation_axis: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    rotation_angle_deg: float = 0.0
    scale: np.ndarray = dataclasses.field(default_factory=lambda: np.ones(3))
    density: float = 1000.0
    dynamic: bool = False
    is_wall: bool = False
    is_animated: bool = False
    sampling_mode: int = 0  # 0 Poisson disk, 1 regular triangle (SurfaceSampling.h:10)
    init_velocity: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    init_angular_velocity: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    target_x: np.ndarray = dataclasses.field(default_factory=lambda: np.array([0.8, 2.0, 0.0]))
    target_angle_deg: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))

    @property
    def quaternion(self) -> np.ndarray:
        """(w, x, y, z) of AngleAxis(angle, axis) (SceneLoader.cpp:115-123); identity
        unless both the axis and the angle are given."""
        a = math.radians(self.rotation_angle_deg)
        ax = np.asarray(self.rotation_axis, np.float64)
        n = np.linalg.norm(ax)
        if n == 0.0 or a == 0.0:
            return np.array([1.0, 0.0, 0.0, 0.0])
        return np.concatenate([[math.cos(a / 2)], math.sin(a / 2) * ax / n])


@dataclasses.dataclass
class FluidBlockData:
    """One "FluidBlocks" entry (SceneLoader.cpp:303-360): the box after scale and translation."""
    box_min: np.ndarray
    box_max: np.ndarray
    mode: int = 0
    fluid_id: str = "Fluid"
    initial_velocity: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    initial_angular_velocity: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))


@dataclasses.dataclass
class SceneFile:
    path: str
    config: dict
    rigid_bodies: List[RigidBodyData]
    fluid_blocks: List[FluidBlockData]
    fluid_models: List[dict]
    materials: List[dict]
    unsupported: List[str]


def _vec(v, n: int) -> np.ndarray:
    a = np.asarray(v, np.float64).reshape(-1)
    if a.shape[0] != n:
        raise ValueError(f"expected {n} values, got {a.shape[0]}")
    return a


def read_scene_file(path: str) -> SceneFile:
    with open(path) as f:
        d = json.load(f)
    cfg = dict(_CONFIG_DEFAULTS)
    cfg.update(d.get("Configuration", {}))
    unsupported = [k for k, test in _UNSUPPORTED_IF.items() if test(cfg[k])]
    bodies = []
    for b in d.get("RigidBodies", []):
        if not (b.get("geometryFile") or b.get("particleFile")):
            continue  # entries without geometry are skipped by the reference too
        rb = RigidBodyData(geometry_file=b.get("geometryFile", ""), particle_file=b.get("particleFile", ""))
        if "translation" in b:
            rb.translation = _vec(b["translation"], 3)
        if "rotationAxis" in b and "rotationAngle" in b:
            rb.rotation_axis = _vec(b["rotationAxis"], 3)
            rb.rotation_angle_deg = float(b["rotationAngle"])
        rb.density = float(b.get("density", rb.density))
        if "scale" in b:
            rb.scale = _vec(b["scale"], 3)
        rb.dynamic = bool(b.get("isDynamic", False))
        rb.is_wall = bool(b.get("isWall", False))
        rb.is_animated = bool(b.get("isAnimated", False))
        rb.sampling_mode = int(b.get("samplingMode", 0))
        for key, attr in (("initVelocity", "init_velocity"), ("initAngularVelocity", "init_angular_velocity"),
                          ("targetX", "target_x"), ("targetAngleInDegree", "target_angle_deg")):
            if key in b:
                setattr(rb, attr, _vec(b[key], 3))
        if rb.is_animated:
            unsupported.append("RigidBodies.isAnimated")
        bodies.append(rb)
    blocks = []
    for fb in d.get("FluidBlocks", []):
        if "start" not in fb or "end" not in fb:
            continue
        t = _vec(fb.get("translation", [0, 0, 0]), 3)
        s = _vec(fb.get("scale", [1, 1, 1]), 3)
        blocks.append(FluidBlockData(box_min=s * _vec(fb["start"], 3) + t, box_max=s * _vec(fb["end"], 3) + t,
                                     mode=int(fb.get("denseMode", 0)), fluid_id=str(fb.get("id", "Fluid")),
                                     initial_velocity=_vec(fb.get("initialVelocity", [0, 0, 0]), 3),
                                     initial_angular_velocity=_vec(fb.get("initialAngularVelocity", [0, 0, 0]), 3)))
        # keys such as "viscosityMethod" inside a block are not read by the
        # reference's loader (SceneLoader.cpp:303-360): fluid parameters come
        # from "Materials"
    materials = list(d.get("Materials", []))
    for m in materials:
        mat = dict(_MATERIAL_DEFAULTS)
        mat.update(m)
        if int(mat["surfaceTensionMethod"]) not in (0, 2, 3):
            unsupported.append("Materials.surfaceTensionMethod")
        if int(mat["viscosityMethod"]) not in (0, 1, 2) or float(mat["viscosityBoundary"]) != 0.0:
            unsupported.append("Materials.viscosityMethod")
        for key in ("vorticityMethod", "dragMethod", "elasticityMethod"):
            if int(mat[key]) != 0:
                unsupported.append(f"Materials.{key}")
    for key in ("Emitters", "ArticulatedSystems", "AnimationFields"):
        if d.get(key):
            unsupported.append(key)
    return SceneFile(path=os.path.abspath(path), config=cfg, rigid_bodies=bodies, fluid_blocks=blocks,
                     fluid_models=list(d.get("FluidModels", [])), materials=materials,
                     unsupported=sorted(set(unsupported)))


# ---------------------------------------------------------------------------
# Fluid blocks (SimulatorBase::createFluidBlocks, Simulator/SimulatorBase.cpp:1638-1722)
# ---------------------------------------------------------------------------
def fluid_block_positions(block: FluidBlockData, radius: float) -> np.ndarray:
    """Lattice of a fluid block: mode 0 regular (spacing 2r), 1 and 2 the
    reference's denser stackings; the first layer starts 2r inside the box."""
    diam = 2.0 * radius
    eps = 1.0e-9
    xs, ys = diam, diam
    if block.mode == 1:
        ys = math.sqrt(3.0) * radius + eps
    elif block.mode == 2:
        xs = math.sqrt(6.0) * diam / 3.0 + eps
        ys = math.sqrt(3.0) * radius + eps
    diff = block.box_max - block.box_min
    if block.mode == 1:
        diff = diff - np.array([diam, 0.0, diam])
    elif block.mode == 2:
        diff = diff - np.array([xs, 0.0, diam])
    nx = int(round(diff[0] / xs)) - 1
    ny = int(round(diff[1] / ys)) - 1
    nz = int(round(diff[2] / diam)) - 1
    if nx <= 0 or ny <= 0 or nz <= 0:
        return np.zeros((0, 3))
    start = block.box_min + diam
    j, k, l = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    j, k, l = j.ravel(), k.ravel(), l.ravel()  # loop order j (x), k (y), l (z), l fastest
    pos = np.stack([j * xs, k * ys, l * diam], axis=1) + start
    if block.mode == 1:
        even = (k % 2) == 0
        pos[even, 2] += radius
        pos[~even, 0] += radius
    elif block.mode == 2:
        pos[:, 2] += radius
        odd_j = (j % 2) == 1
        pos[odd_j, 2] += np.where(k[odd_j] % 2 == 1, -1.0, 1.0) * (diam / 2.0)
        pos[(k % 2) == 0, 0] += xs / 2.0
    return pos


def block_velocities(block: FluidBlockData, pos: np.ndarray) -> np.ndarray:
    """SimulatorBase::setInitialVelocity: v = v0 + omega x (x - centre of the block's particles)."""
    v = np.tile(block.initial_velocity, (pos.shape[0], 1))
    if pos.shape[0] and np.any(block.initial_angular_velocity != 0):
        c = pos.mean(axis=0)
        v = v + np.cross(block.initial_angular_velocity, pos - c)
    return v


# ---------------------------------------------------------------------------
# Meshes and surface sampling
# ---------------------------------------------------------------------------
def read_obj(path: str, scale=(1.0, 1.0, 1.0)):
    """Vertices (scaled per axis, as the OBJ loader applies the body's scale) and
    triangles (0-based; polygons fanned from their first vertex)."""
    s = np.asarray(scale, np.float64)
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                p = line.split()
                verts.append([float(p[1]), float(p[2]), float(p[3])])
            elif line.startswith("f "):
                idx = [int(tok.split("/")[0]) for tok in line.split()[1:]]
                idx = [i - 1 if i > 0 else len(verts) + i for i in idx]
                for t in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[t], idx[t + 1]])
    V = np.asarray(verts, np.float64).reshape(-1, 3) * s
    return V, np.asarray(faces, np.int64).reshape(-1, 3)


def regular_triangle_sampling(V: np.ndarray, F: np.ndarray, d: float) -> np.ndarray:
    """RegularTriangleSampling::sampleMesh (SPlisHSPlasH/Utilities/RegularTriangleSampling.cpp):
    every vertex; on every unique edge the interior points of ceil(l/d)+1
    equally spaced ones; on every triangle, lines parallel to its longest edge
    at ceil(h/d)+1 heights (the interior ones), each with the interior points of
    ceil(len/d)+1 equally spaced ones."""
    out = [V.copy()]
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    e = np.unique(np.sort(e, axis=1), axis=0)
    for a, b in e:
        v0 = V[a]
        direc = V[b] - v0
        n = int(math.ceil(np.linalg.norm(direc) / d)) + 1
        if n == 1:
            out.append(v0[None, :])
            continue
        i = np.arange(1, n - 1)
        out.append(v0 + (i / (n - 1))[:, None] * direc)
    for f in F:
        best = -1.0
        for c in range(3):
            v = V[f[(c + 1) % 3]] - V[f[c]]
            l2 = float(v @ v)
            if l2 > best:
                best, v0, base, top = l2, V[f[c]], v, V[f[(c + 2) % 3]] - V[f[c]]
        l = math.sqrt(best)
        bn = base / l
        hgt = np.linalg.norm(top - (bn @ top) * bn)
        nl = int(math.ceil(hgt / d)) + 1
        for line in range(1, nl - 1):
            alpha = line / (nl - 1)
            ls = v0 + alpha * top
            ll = (1.0 - alpha) * l
            ns = int(math.ceil(ll / d)) + 1
            i = np.arange(1, ns - 1)
            if i.size:
                out.append(ls + ((i * ll) / (ns - 1))[:, None] * bn)
    return np.concatenate(out, axis=0)


def poisson_disk_sampling(V: np.ndarray, F: np.ndarray, r: float, trials: int = 10, distance_norm: int = 1,
                          seed: int = 0) -> np.ndarray:
    """Poisson disk surface sampling after PoissonDiskSampling::sampleMesh
    (SPlisHSPlasH/Utilities/PoissonDiskSampling.cpp): 40 x (area / pi r^2)
    candidates, area-weighted triangles and uniform barycentric coordinates;
    candidates binned in cells of edge r/sqrt(3), sorted by cell; then, for
    trial k = 0..trials-1 and each of the 27 cell phases in turn, the k-th
    candidate of every cell of the phase is kept unless a kept sample lies
    closer than r (distance_norm 1: the chord corrected for the angle between
    the two faces' normals, which is never shorter than the chord; 0:
    Euclidean).  The cells of one phase are at least two cells (1.15 r) apart,
    so the candidates of a (trial, phase) pass are tested together against
    the samples kept before it."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    normals = cr / np.maximum(np.linalg.norm(cr, axis=1), 1e-300)[:, None]
    total = float(area.sum())
    n0 = int(40.0 * (total / (math.pi * r * r)))
    if n0 == 0:
        return np.zeros((0, 3))
    tri = rng.choice(F.shape[0], size=n0, p=area / total)
    rn1 = np.sqrt(rng.random(n0))
    bc1 = 1.0 - rn1
    bc2 = rng.random(n0) * rn1
    bc3 = 1.0 - bc1 - bc2
    pts = bc1[:, None] * a[tri] + bc2[:, None] * b[tri] + bc3[:, None] * c[tri]
    cp = np.floor((pts - V.min(axis=0)) / (r / math.sqrt(3.0))).astype(np.int64) + 1
    order = np.lexsort((cp[:, 2], cp[:, 1], cp[:, 0]))
    pts, tri, cp = pts[order], tri[order], cp[order]
    newcell = np.ones(n0, bool)
    newcell[1:] = np.any(cp[1:] != cp[:-1], axis=1)
    first = np.maximum.accumulate(np.where(newcell, np.arange(n0), 0))
    rank = np.arange(n0) - first  # position of each candidate within its cell
    phase = cp[:, 0] % 3 + 3 * (cp[:, 1] % 3) + 9 * (cp[:, 2] % 3)
    kept = np.zeros(0, np.int64)
    for k in range(trials):
        at_k = rank == k
        if not at_k.any():
            break
        for pg in range(27):
            cand = np.nonzero(at_k & (phase == pg))[0]
            if cand.size == 0:
                continue
            if kept.size:
                tree = cKDTree(pts[kept])
                lists = tree.query_ball_point(pts[cand], r)
                lens = np.fromiter((len(l) for l in lists), np.int64, count=len(lists))
                bad = np.zeros(cand.size, bool)
                if lens.sum():
                    ci = np.repeat(np.arange(cand.size), lens)
                    qi = kept[np.concatenate([np.asarray(l, np.int64) for l in lists if len(l)])]
                    pi = cand[ci]
                    dv = pts[qi] - pts[pi]
                    dist = np.linalg.norm(dv, axis=1)
                    if distance_norm == 1:
                        other = (tri[qi] != tri[pi]) & (dist > 0.0)
                        u = dv[other] / dist[other][:, None]
                        c1 = np.clip(np.einsum("ij,ij->i", normals[tri[pi[other]]], u), -1.0, 1.0)
                        c2 = np.clip(np.einsum("ij,ij->i", normals[tri[qi[other]]], u), -1.0, 1.0)
                        diff = c1 - c2
                        big = np.abs(diff) > 1e-5
                        fac = np.where(big, (np.arcsin(c1) - np.arcsin(c2)) / np.where(big, diff, 1.0),
                                       1.0 / np.sqrt(np.maximum(1.0 - c1 * c1, 1e-300)))
                        dist[other] = dist[other] * fac
                    bad = np.bincount(ci, weights=(dist < r), minlength=cand.size) > 0
                cand = cand[~bad]
            kept = np.concatenate([kept, cand])
    return pts[kept]


def body_samples(rb: RigidBodyData, radius: float, scene_dir: str, seed: int = 0) -> np.ndarray:
    """Surface samples of a rigid body in its local frame (the scaled mesh's
    coordinates), as RigidBody3dBoundarySimulator::initBoundaryData samples it."""
    if rb.particle_file:
        data = read_bgeo(os.path.join(scene_dir, rb.particle_file))
        return data["position"].astype(np.float64) * float(rb.scale[0])
    V, F = read_obj(os.path.normpath(os.path.join(scene_dir, rb.geometry_file)), rb.scale)
    if rb.sampling_mode == 0:
        return poisson_disk_sampling(V, F, radius, 10, 1, seed=seed)
    return regular_triangle_sampling(V, F, 1.5 * radius)


def fluid_material(sf: SceneFile, fluid_id: str = "Fluid") -> dict:
    """Non-pressure force parameters of the fluid `fluid_id`: the "Materials"
    entry with that id over the FluidModel's defaults (Standard viscosity
    0.01, no surface tension)."""
    mat = dict(_MATERIAL_DEFAULTS)
    for m in sf.materials:
        if str(m.get("id", "Fluid")) == fluid_id:
            mat.update(m)
    return dict(viscosity_method=int(mat["viscosityMethod"]), viscosity=float(mat["viscosity"]),
                viscosity_boundary=float(mat["viscosityBoundary"]),
                surface_tension_method=int(mat["surfaceTensionMethod"]),
                surface_tension=float(mat["surfaceTension"]),
                surface_tension_boundary=float(mat["surfaceTensionBoundary"]))


def build_scene(sf: SceneFile, steps: Optional[int] = None, seed: int = 0) -> scenes.Scene:
    """A `scenes.Scene` from a parsed scene file.  Body 0 of the reference's
    drivers is the static tank and body 1 the optimised body; the order of the
    file is kept."""
    cfg = sf.config
    r = float(cfg["particleRadius"])
    p = scenes.default_params(r, float(cfg["timeStepSize"]))
    p.update(rho0=float(cfg["density0"]), gravity=tuple(float(g) for g in cfg["gravitation"]),
             max_error=float(cfg["maxError"]), max_error_v=float(cfg["maxErrorV"]),
             max_iterations=int(cfg["maxIterations"]), max_iterations_v=int(cfg["maxIterationsV"]),
             min_iterations=int(cfg["minIterations"]),
             enable_divergence_solver=int(bool(cfg["enableDivergenceSolver"])),
             gradient_mode=int(cfg["gradientMode"]), target_time=float(cfg["targetTime"]),
             uniform_acc_time=float(cfg["uniformAccelerateRBTime"]),
             use_gradient_manager=int(bool(cfg["useRigidGradientManager"])),
             release_mode=int(bool(cfg["useReleaseRigidBodyMode"])),
             cfl_method=int(cfg["cflMethod"]), cfl_factor=float(cfg["cflFactor"]),
             cfl_min_time_step_size=float(cfg["cflMinTimeStepSize"]),
             cfl_max_time_step_size=float(cfg["cflMaxTimeStepSize"]),
             use_rigid_contact=int(bool(cfg["useRigidContactSolver"])),
             rigid_contact_beta=float(cfg["rigidContactBeta"]), rigid_contact_gamma=float(cfg["rigidContactGamma"]),
             rigid_contact_support_radius_factor=float(cfg["rigidContactSupportRadiusFactor"]),
             rigid_contact_friction=float(cfg["rigidContactFrictionCoeff"]))
    p.update(fluid_material(sf))
    scene_dir = os.path.dirname(sf.path)
    xs, vs = [], []
    for blk in sf.fluid_blocks:
        x = fluid_block_positions(blk, r)
        xs.append(x)
        vs.append(block_velocities(blk, x))
    for fm in sf.fluid_models:
        if fm.get("particleFile"):
            data = read_bgeo(os.path.join(scene_dir, fm["particleFile"]))
            s = _vec(fm.get("scale", [1, 1, 1]), 3)
            t = _vec(fm.get("translation", [0, 0, 0]), 3)
            x = data["position"].astype(np.float64) * s + t
            xs.append(x)
            vs.append(np.tile(_vec(fm.get("initialVelocity", [0, 0, 0]), 3), (x.shape[0], 1)))
    fluid_x = np.concatenate(xs, axis=0) if xs else np.zeros((0, 3))
    fluid_v = np.concatenate(vs, axis=0) if vs else np.zeros((0, 3))
    bodies = []
    for k, rb in enumerate(sf.rigid_bodies):
        x0 = body_samples(rb, r, scene_dir, seed=seed + k)
        name = os.path.splitext(os.path.basename(rb.geometry_file or rb.particle_file))[0]
        bodies.append(scenes.Body(x0, bool(rb.dynamic), rb.translation.copy(), rb.quaternion, float(rb.density),
                                  rb.init_velocity.copy(), rb.init_angular_velocity.copy(), name=name,
                                  target_x=rb.target_x.copy(), target_angle=np.radians(rb.target_angle_deg)))
    name = os.path.splitext(os.path.basename(sf.path))[0]
    n_steps = steps if steps is not None else int(round(float(cfg["targetTime"]) / float(cfg["timeStepSize"])))
    return scenes.Scene(name, p, fluid_x, fluid_v, bodies, n_steps, unsupported=list(sf.unsupported))


def load_scene(path: str, steps: Optional[int] = None, seed: int = 0) -> scenes.Scene:
    return build_scene(read_scene_file(path), steps=steps, seed=seed)


# ---------------------------------------------------------------------------
# State files
# ---------------------------------------------------------------------------
def read_bgeo(path: str) -> dict:
    """Point attributes of a Houdini BGEO version 5 file as partio writes it
    (gzip-compressed or plain): "position" (float [n][3]) and every point
    attribute (float, int or vector; indexed strings as int indices)."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:2] == b"\x1f\x8b":
        raw = gzip.decompress(raw)
    if raw[:4] != b"Bgeo" or raw[4:5] != b"V":
        raise ValueError(f"{path}: not a BGEO file")
    off = 5
    (version, n_points, n_prims, n_pgroups, n_prgroups, n_pattr, n_vattr, n_prattr, n_attr) = \
        struct.unpack_from(">9i", raw, off)
    off += 36
    if version != 5:
        raise ValueError(f"{path}: BGEO version {version}, only version 5 is read")
    attrs = [("position", "f", 3, 0)]  # x, y, z, w of every point; w is skipped
    width = 4
    for _ in range(n_pattr):
        (nl,) = struct.unpack_from(">H", raw, off)
        off += 2
        name = raw[off:off + nl].decode()
        off += nl
        size, htype = struct.unpack_from(">Hi", raw, off)
        off += 6
        if htype in (0, 1, 5):
            off += 4 * size  # default values
            attrs.append((name, "i" if htype == 1 else "f", size, width))
        elif htype == 4:
            (ni,) = struct.unpack_from(">i", raw, off)
            off += 4
            for _ in range(ni):
                (sl,) = struct.unpack_from(">H", raw, off)
                off += 2 + sl
            attrs.append((name, "i", size, width))
        else:
            raise ValueError(f"{path}: unsupported attribute type {htype} of {name!r}")
        width += size
    rec = np.frombuffer(raw, dtype=">i4", count=n_points * width, offset=off).reshape(n_points, width)
    out = {}
    for name, kind, size, col in attrs:
        block = rec[:, col:col + size].astype(np.int32)
        out[name] = (block.view(np.float32) if kind == "f" else block).copy()
        if size == 1:
            out[name] = out[name][:, 0]
    return out


def read_state(bin_path: str, n_dynamic: int, time: Optional[float] = None) -> dict:
    """A state written by SimulatorBase::saveState: the ".bin" header (scene
    MD5, next frame times, frame counter), the rigid records of the n_dynamic
    dynamic/animated bodies at the file's end (x, q as (w, x, y, z), v,
    omega; FP32 as written), the fluid fields of "<name>_particle_<id>.bgeo"
    beside it, and the time (argument, else the number in a "state_<time>"
    name; None when neither gives it)."""
    with open(bin_path, "rb") as f:
        raw = f.read()
    (n,) = struct.unpack_from("<I", raw, 0)
    md5 = raw[4:4 + n].decode(errors="replace")
    off = 4 + n
    nft, nfts, frame = struct.unpack_from("<ffI", raw, off)
    rig = []
    if n_dynamic > 0:
        tail = np.frombuffer(raw, dtype="<f4", count=13 * n_dynamic, offset=len(raw) - 52 * n_dynamic)
        for k in range(n_dynamic):
            rec = tail[13 * k:13 * k + 13].astype(np.float64)
            q_xyzw = rec[3:7]
            rig.append(dict(x=rec[0:3].copy(), q=np.array([q_xyzw[3], q_xyzw[0], q_xyzw[1], q_xyzw[2]]),
                            v=rec[7:10].copy(), omega=rec[10:13].copy()))
    base = bin_path[:-4] if bin_path.endswith(".bin") else bin_path
    fluids = {}
    d = os.path.dirname(base) or "."
    stem = os.path.basename(base) + "_particle_"
    for fn in sorted(os.listdir(d)):
        if fn.startswith(stem) and fn.endswith(".bgeo"):
            fluids[fn[len(stem):-5]] = read_bgeo(os.path.join(d, fn))
    if time is None:
        m = re.fullmatch(r"state_(\d+\.\d+)", os.path.basename(base))
        time = float(m.group(1)) if m else None
    return dict(md5=md5, next_frame_time=nft, next_frame_time_state=nfts, frame=frame, rigid=rig, fluids=fluids,
                time=time)
