"""Batches of rooms: the merged grid steps each room as the room alone, to the bit.
CPU: the oracle on the merged layout against the oracle on each room. GPU: the batched
HIP stepper against the oracle on each room."""

import numpy as np
import pytest

import cases
from acoustic_spaces_amd import scenes


def _rooms(count=4, seed=21):
    rooms = scenes.rooms_batch(count=count, cell=0.1, seed=seed, length_s=0.03, size_low=(1.2, 1.1, 1.0),
                               size_high=(2.6, 2.0, 1.6))
    return [cases.with_branches(r) if k % 2 else r for k, r in enumerate(rooms)]


def _layouts(rooms):
    from acoustic_spaces_amd.solver import WaveSolver

    return [WaveSolver.from_scene(r).layout() for r in rooms]


def test_merged_layout_steps_each_room_as_alone():
    from acoustic_spaces_amd.batch import merge_layouts
    from oracle.stepper import OracleSolver

    rooms = _rooms()
    layouts = _layouts(rooms)
    merged, placements = merge_layouts(layouts)
    drive = rooms[0].drive[:90]
    together = OracleSolver(merged).run(drive)
    for L, p in zip(layouts, placements):
        alone = OracleSolver(L).run(drive)
        assert np.array_equal(together[p.probes], alone)


def test_batch_refuses_mixed_stencils():
    from acoustic_spaces_amd.batch import merge_layouts

    a = _layouts([cases.small_box()])[0]
    b = _layouts([cases.small_box(stencil=19)])[0]
    with pytest.raises(ValueError, match="stencil"):
        merge_layouts([a, b])


@pytest.mark.gpu
def test_batched_gpu_equals_each_room_alone():
    from acoustic_spaces_amd.batch import BatchedWaveSolver
    from oracle.stepper import OracleSolver

    rooms = _rooms(count=6, seed=22)
    layouts = _layouts(rooms)
    batch = BatchedWaveSolver(layouts)
    drive = rooms[0].drive[:200]
    first = batch.run(drive[:77])
    second = batch.run(drive[77:])
    for L, a, b in zip(layouts, first, second):
        cpu = OracleSolver(L)
        assert np.array_equal(a, cpu.run(drive[:77]))
        assert np.array_equal(b, cpu.run(drive[77:]))


@pytest.mark.gpu
def test_autotuner_picks_a_measured_shape(tmp_path, monkeypatch):
    from acoustic_spaces_amd import tune

    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    layout = _layouts([cases.small_box()])[0]
    shapes = ({"tile_z": 4, "graph_steps": 32}, {"tile_z": 8, "graph_steps": 64})
    best = tune.autotune(layout, steps=64, candidates=shapes)
    assert (best["tile_z"], best["graph_steps"]) in {(s["tile_z"], s["graph_steps"]) for s in shapes}
    assert best["us_per_step"] == min(t["us_per_step"] for t in best["timings"])
    assert tune.autotune(layout, steps=64, candidates=shapes) == best  # from the cache
