"""The HIP stepper against the golden fixtures (tests/golden/*.json): on this package's
set-up of each config's grid, the receivers' float64 signals over the fixture's steps (in
two calls) hash as the reference's own run's, for the 6 x 4 x 3 m shoebox (all 5,941
steps), the hall (1,200 steps) and the outdoor grid with its matched layer (240 steps)."""

import numpy as np
import pytest

from test_golden import golden, scene_of, sha

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["shoebox", "hall", "outdoor"])
def test_signals_equal_the_golden(name):
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    g = golden(name)
    scene = scene_of(name)
    print(f"golden {name}: {g['steps']} steps", flush=True)
    gpu = HipWaveSolver(WaveSolver.from_scene(scene))
    drive = np.asarray(scene.drive[: g["steps"]], np.float32)
    first = gpu.run(drive[: g["short"]])
    rest = gpu.run(drive[g["short"]:])
    assert sha(first) == g["short_signals_sha256"]
    assert sha(np.concatenate([first, rest], axis=1)) == g["signals_sha256"]
