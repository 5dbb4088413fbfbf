"""Copies the reference's scene, mesh and state files that the scene/state
tests read into tests/golden/reference_files/ (data only; run where the
reference tree is present).  Layout: scenes/ holds the experiments' scene
JSONs; bottle_flip_robot/ keeps that experiment's relative layout
(scene/, models/, state/) so its scene's mesh paths resolve."""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_files")
EXP = os.path.join(REF, "experiments")

SCENES = [
    "bottle_flip_robot/scene/diff-bottle-model-360degree.json",
    "on_water_inverted_pendulum/scene/cartpole-diff-controller.json",
    "rigid_body_trajectory_optimization/scene/billiards-on-water-2balls.json",
    "rigid_body_trajectory_optimization/scene/diff-bottle-model-collide.json",
    "rigid_body_trajectory_optimization/scene/diff-bottle-model.json",
    "rigid_body_trajectory_optimization/scene/diff-high-diving-duck.json",
    "rigid_body_trajectory_optimization/scene/diff-stone-skipping.json",
    "rigid_body_trajectory_optimization/scene/diff-water-rafting-bunny.json",
]
BOTTLE = [
    "bottle_flip_robot/scene/diff-bottle-model-360degree.json",
    "bottle_flip_robot/models/UnitBox.obj",
    "bottle_flip_robot/models/bottle.obj",
    "bottle_flip_robot/state/state_1.365999.bin",
    "bottle_flip_robot/state/state_1.365999_particle_Fluid.bgeo",
]

os.makedirs(os.path.join(OUT, "scenes"), exist_ok=True)
for s in SCENES:
    shutil.copyfile(os.path.join(EXP, s), os.path.join(OUT, "scenes", os.path.basename(s)))
for f in BOTTLE:
    dst = os.path.join(OUT, f)
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    shutil.copyfile(os.path.join(EXP, f), dst)
print("copied", len(SCENES) + len(BOTTLE), "files into", OUT)
