"""The committed golden fixtures (tests/golden/*.json, written by tests/golden/make_golden.py
from the reference's own solver, oracle/_ref): hashes of the configs' layouts and of the
reference's receiver signals, checked without the reference.

CPU: this package's set-up lays each config's grid out to the same bytes (every layout
array's SHA-256 and dtype), the scenes' drives are the same, and the C oracle's signals
over the first steps hash the same. GPU (tests/test_gpu_golden.py): the HIP stepper's."""

import hashlib
import json
import os

import numpy as np
import pytest

from acoustic_spaces_amd import scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("shoebox", "hall", "outdoor")


def golden(name):
    with open(os.path.join(HERE, f"{name}.json")) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scene_of(name):
    return {"shoebox": scenes.shoebox, "hall": lambda: scenes.hall()[0], "outdoor": scenes.outdoor}[name]()


@pytest.mark.parametrize("name", NAMES)
def test_layout_and_drive_equal_the_golden(name):
    from acoustic_spaces_amd.solver import WaveSolver

    g = golden(name)
    scene = scene_of(name)
    assert list(scene.grid.shape) == g["grid"]
    assert sha(np.asarray(scene.drive[: g["steps"]], np.float32)) == g["drive_sha256"]
    solver = WaveSolver.from_scene(scene)
    layout = solver.layout()
    assert sorted(layout) == sorted(g["layout"])
    for k, v in g["layout"].items():
        a = layout[k]
        assert str(a.dtype) == v["dtype"] and a.size == v["size"], k
        assert sha(a) == v["sha256"], k
    assert solver.air_cells == g["air_cells"]
    assert solver.boundary_cells == g["boundary_cells"]


@pytest.mark.parametrize("name", ["shoebox", "hall"])
def test_oracle_signals_equal_the_golden(name):
    """The C oracle (oracle/step.c) on the golden layout's grid: the first steps' signals
    hash as the reference's."""
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver

    g = golden(name)
    scene = scene_of(name)
    cpu = OracleSolver(WaveSolver.from_scene(scene).layout())
    y = cpu.run(np.asarray(scene.drive[: g["short"]], np.float32))
    assert sha(y) == g["short_signals_sha256"]
    assert np.array_equal(y[:, :16], np.array(g["head"]))
