"""Writes the small fixtures under tests/golden/ (run once, in the build container).

Data files of the reference (SPH/data/models/*.ply, SPH/data/scenes/*.json) are
read as data with this package's own PLY and JSON readers and stored as .npy/.json
fixtures; nothing of the reference's code is imported or run.  The oracle-made
regression vectors (oracle_*.npz) come from this repository's own oracle.

    python tests/golden/make_fixtures.py
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/SPH/data"

from nonnewton_amd import ply, scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402


def reference_data():
    for name in ("test_neighbor", "onepar", "test_deformable", "torus"):
        pts = ply.read_ply_particles(os.path.join(REF, "models", name + ".ply"))
        np.save(os.path.join(HERE, f"{name}_points.npy"), pts)
    for name in ("test_neighbor", "test_nonNewton", "test_deformable", "dambreak", "rigidbody", "test_oneway_coupling"):
        shutil.copyfile(os.path.join(REF, "scenes", name + ".json"), os.path.join(HERE, f"scene_{name}.json"))


def oracle_vectors():
    sc = scenes.tiny_box()
    out = {"x0": sc.x, "kind": sc.kind}
    for prec in ("f32", "f64"):
        s = O.OracleSim.from_scene(sc, precision=prec)
        s.init()
        for _ in range(3):
            st = s.step()
        out[f"{prec}_x"] = s.get(O.F_X)
        out[f"{prec}_v"] = s.get(O.F_V)
        out[f"{prec}_density"] = s.get(O.F_DENSITY)
        out[f"{prec}_iters"] = np.array([st["iterations_v"], st["iterations_p"]])
    np.savez_compressed(os.path.join(HERE, "oracle_tiny_box_3steps.npz"), **out)


if __name__ == "__main__":
    if os.path.isdir(REF):
        reference_data()
    oracle_vectors()
    print("fixtures written to", HERE)
