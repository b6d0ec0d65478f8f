"""The hall with fitted walls, laid out by this package's own fit (csrc/fitted.cpp) and
stepped by the HIP kernels: bit for bit the oracle's run of the same layout, and the
reference's own receivers (tests/golden/hall_fitted*, from its fitted set-up and run,
make_fitted_hall.py) within the north star's bars: 1e-4 relative L2 of the signals, 0.05 dB
in each octave band's strength."""

import json
from pathlib import Path

import numpy as np
import pytest

from acoustic_spaces_amd import scenes

GOLDEN = Path(__file__).parent / "golden"


def _fitted(scene):
    from acoustic_spaces_amd.solver import WaveSolver

    return WaveSolver.from_scene(scene, fitted_walls=True).layout()


@pytest.mark.gpu
def test_fitted_hall_steps_equal_the_oracle():
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from oracle.stepper import OracleSolver

    scene = scenes.hall()[0]
    layout = _fitted(scene)
    print(f"fitted hall: {len(layout['special'])} specials", flush=True)
    gpu = HipWaveSolver(layout)
    # The library's choice: the kinds' tables (486,785 specials of 5,003 kinds).
    assert 0 < gpu.info()["special_kinds"] <= len(layout["special"]) // 4
    cpu = OracleSolver(layout)
    d = scene.drive[40:190]
    first, second = cpu.run(d[:75]), cpu.run(d[75:])
    assert np.array_equal(gpu.run(d[:75]), first)
    assert np.array_equal(gpu.run(d[75:]), second)
    assert np.array_equal(gpu.field(), cpu.field())
    # Each special's own slabs: the same run.
    slabs = HipWaveSolver(layout, special_kinds=1)
    assert slabs.info()["special_kinds"] == 0
    assert np.array_equal(slabs.run(d), np.concatenate([first, second], axis=1))
    assert np.array_equal(slabs.field(), cpu.field())


@pytest.mark.gpu
@pytest.mark.parametrize("source", [1, 2])
def test_fitted_hall_receivers_match_the_reference(source):
    from acoustic_spaces_amd.analysis import octave_results
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.wave import high_pass

    scene = scenes.hall()[source - 1]
    want = json.loads((GOLDEN / "hall_fitted.json").read_text())[f"source{source}"]
    prefix = np.load(GOLDEN / "hall_fitted_signals.npz")[f"source{source}"]
    y = HipWaveSolver.run_once(_fitted(scene), scene.drive)
    assert y.shape[1] == want["steps"]
    head = y[:, : prefix.shape[1]]
    rel = np.linalg.norm(head - prefix, axis=1) / np.linalg.norm(prefix, axis=1)
    print(f"source {source}: largest relative L2 over the first {prefix.shape[1]} steps {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-4
    rows = [octave_results(high_pass(r, scene.rate), scene.rate, bands=3) for r in y]
    dg = np.abs(np.array([r["g_db"] for r in rows]) - np.array(want["g_db"]))
    dt = np.abs(np.array([r["t30_s"] for r in rows]) - np.array(want["t30_s"]))
    print(f"source {source}: |dG| <= {dg.max():.2e} dB, |dT30| <= {dt.max():.2e} s", flush=True)
    assert dg.max() < 0.05
    assert dt.max() < 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("kinds", [1, 2])
def test_special_orders_step_alike(order, split, kinds):
    """The general specials in each order (by counts, by cell, by tile), fused and split,
    from their slabs (kinds 1) or their kinds' tables (2): the coarse hall's fitted layout,
    300 steps in two calls, bit for bit the oracle's."""
    import sys

    sys.path.insert(0, str(Path(__file__).parent))
    import cases
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from oracle.stepper import OracleSolver

    scene = cases.coarse_hall()
    layout = _fitted(scene)
    gpu = HipWaveSolver(layout, special_order=order, split=split, special_kinds=kinds)
    assert (gpu.info()["special_kinds"] > 0) == (kinds == 2)
    cpu = OracleSolver(layout)
    d = scene.drive[:300]
    assert np.array_equal(gpu.run(d[:151]), cpu.run(d[:151]))
    assert np.array_equal(gpu.run(d[151:]), cpu.run(d[151:]))
    assert np.array_equal(gpu.field(), cpu.field())
