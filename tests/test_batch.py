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


def test_groups_cover_each_room_once_within_budget():
    from acoustic_spaces_amd.batch import BYTES_PER_CELL, plan_groups

    layouts = _layouts(_rooms(count=7, seed=23))
    assert len(plan_groups(layouts)) == 1 and sorted(plan_groups(layouts)[0]) == list(range(len(layouts)))
    budget = 2 * max(int(np.prod(L["shape"])) for L in layouts) * BYTES_PER_CELL[32]
    groups = plan_groups(layouts, budget)
    assert sorted(i for g in groups for i in g) == list(range(len(layouts)))
    assert len(groups) > 1
    for g in groups:
        nx = max(int(layouts[i]["shape"][0]) for i in g)
        ny = max(int(layouts[i]["shape"][1]) for i in g)
        nz = sum(int(layouts[i]["shape"][2]) for i in g)
        assert len(g) == 1 or nx * ny * nz * BYTES_PER_CELL[32] <= budget


@pytest.mark.gpu
@pytest.mark.parametrize("grouped", [False, True])
def test_batched_gpu_equals_each_room_alone(grouped):
    from acoustic_spaces_amd.batch import BYTES_PER_CELL, BatchedWaveSolver
    from oracle.stepper import OracleSolver

    rooms = _rooms(count=6, seed=22)
    layouts = _layouts(rooms)
    budget = 2 * max(int(np.prod(L["shape"])) for L in layouts) * BYTES_PER_CELL[32] if grouped else None
    batch = BatchedWaveSolver(layouts, group_bytes=budget)
    print(f"batch grouped={grouped} groups={len(batch.groups)}", flush=True)
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
    # The first shape, in the candidates' order, within TIE of the fastest.
    fastest = min(t["us_per_step"] for t in best["timings"])
    first = next(t for t in best["timings"] if t["us_per_step"] <= fastest * (1.0 + tune.TIE))
    assert (best["tile_z"], best["graph_steps"], best["us_per_step"]) == (
        first["tile_z"], first["graph_steps"], first["us_per_step"])
    assert tune.autotune(layout, steps=64, candidates=shapes) == best  # from the cache


def test_tuning_cache_key_follows_the_library(tmp_path, monkeypatch):
    """The autotuner's cache key changes with the library's bytes (a rebuilt library times
    its shapes again), with the precision and with live fields, and not otherwise."""
    from acoustic_spaces_amd import native, tune

    layout = _layouts(_rooms(count=1))[0]
    lib = tmp_path / "libasfdtd.so"
    lib.write_bytes(b"one build")
    monkeypatch.setattr(native, "LIBRARY", lib)
    key = tune.cache_key(layout)
    assert tune.cache_key(layout) == key
    assert tune.cache_key(layout, precision=64) != key
    assert tune.cache_key(layout, live=True) != key
    lib.write_bytes(b"another build")
    assert tune.cache_key(layout) != key


@pytest.mark.gpu
def test_sixty_four_rooms_at_5_cm_batched_equal_their_plain_solves():
    """The batching workload itself: 64 shoeboxes U(3,12)xU(3,10)xU(2.5,5) m at 5 cm
    (0.2M-4.8M cells each, Monte Carlo wall absorption), 0.5 s each. Batched (the default
    grouping), every room's 4 receivers equal its own plain solve's bit for bit over all
    steps; the first 300 steps of three rooms (smallest, median, largest) equal the oracle's."""
    from acoustic_spaces_amd.batch import BatchedWaveSolver
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver

    rooms = scenes.rooms_batch(count=64)
    layouts = [WaveSolver.from_scene(r).layout() for r in rooms]
    cells = [int(np.prod(L["shape"])) for L in layouts]
    assert len(rooms) == 64 and rooms[0].grid.cell == 0.05
    drive = rooms[0].drive
    print(f"64 rooms: {min(cells):,} to {max(cells):,} cells, {len(drive)} steps", flush=True)
    batch = BatchedWaveSolver(layouts, group_bytes=BatchedWaveSolver.DEFAULT_GROUP_BYTES)
    together = batch.run(drive)
    batch.close()
    print(f"batched in {len(batch.groups)} groups", flush=True)
    for k, L in enumerate(layouts):
        alone = HipWaveSolver.run_once(L, drive)
        assert np.array_equal(together[k], alone), f"room {k}"
    order = np.argsort(cells)
    for k in (order[0], order[len(order) // 2], order[-1]):
        cpu = OracleSolver(layouts[k]).run(drive[:300])
        assert np.array_equal(together[k][:, :300], cpu), f"room {k} against the oracle"
        print(f"room {k} ({cells[k]:,} cells): 300 steps equal to the oracle", flush=True)
