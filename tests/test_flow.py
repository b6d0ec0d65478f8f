"""Moving air (the reference's set_up_flow and its carrying, fdtd.hpp:1680-2120): the set-up's
shifts, row masks, the wind solved round the solids and its loops, equal to the reference's
own (oracle/_ref) bit for bit, plain and round, on both stencils; its refusals; the oracle's
carrying against the reference's run; and on the GPU, the carrying's three stages before each
step, the probes equal to the reference's run bit for bit (fused, split and host-launched
steps, runs in pieces, the device set-up's layout)."""

import numpy as np
import pytest

import cases
from acoustic_spaces_amd import scenes


def ground_scene(cell=0.1, layer=4, stencil=7, n=(22, 20, 20)):
    """A ground block inside a padded grid with a matched layer: a wind over it that changes
    with height only moves each row as a whole (the plain case). ``n``: the grid's cells."""
    pad = layer + 2
    grid = scenes.Grid(-cell * (pad + scenes.GRID_SHIFT), cell, n)
    gx, gy = (n[0] - 2 * pad) * cell, (n[1] - 2 * pad) * cell
    top = 0.3
    mesh = scenes.box_mesh((gx, gy, top), low=(0.0, 0.0, 0.0), inward=False)
    courant = scenes.COURANT if stencil == 7 else scenes.ISOTROPIC_COURANT
    rate = scenes.C / (courant * cell)
    sc = scenes.Scene(name="ground under wind", mesh=mesh, beta=np.full(mesh.materials, 0.1),
                      branches=np.zeros((0, 3)), branch_material=np.zeros(0, np.int64), grid=grid,
                      source=np.array([0.41, 0.52, top + 0.27]),
                      probes=np.array([[0.8, 0.6, top + 0.3], [0.3, 0.7, top + 0.12]]), rate=rate, steps=300,
                      drive=scenes.gaussian_drive(rate, 300), layer=layer, layer_damping=0.5, stencil=stencil)
    return cases.with_branches(sc, per_material=((0.0, 4.0, 0.2),)), top


def wind_profile(sc, top):
    nx, ny, nz = sc.grid.shape
    z = sc.grid.origin[2] + (np.arange(nz) + 0.5) * sc.grid.cell
    prof = np.where(z > top, 0.05 + 0.1 * np.clip(z - top, 0, None), 0.0)
    fx = np.broadcast_to(prof[:, None, None], (nz, ny, nx)).astype(np.float32).ravel()
    fy = np.broadcast_to((-0.5 * prof)[:, None, None], (nz, ny, nx)).astype(np.float32).ravel()
    return fx, fy


def vortex(sc, strength=0.15):
    """The air turning about the grid's vertical centre line, as a body turns."""
    nx, ny, nz = sc.grid.shape
    c = [sc.grid.origin[a] + (np.arange(sc.grid.shape[a]) + 0.5) * sc.grid.cell for a in range(2)]
    x, y = c[0][None, :] - c[0].mean(), c[1][:, None] - c[1].mean()
    reach = max(np.abs(x).max(), np.abs(y).max())
    vx = np.broadcast_to(-strength * y / reach, (nz, ny, nx)).astype(np.float32).ravel()
    vy = np.broadcast_to(strength * x / reach, (nz, ny, nx)).astype(np.float32).ravel()
    return vx, vy


def uniform(sc, vx=0.1, vy=0.05):
    n = sc.grid.cells
    return np.full(n, vx, np.float32), np.full(n, vy, np.float32)


CASES = {
    "plain": lambda: (lambda sc_top: (sc_top[0], wind_profile(*sc_top)))(ground_scene()),
    "plain_wide": lambda: (lambda sc_top: (sc_top[0], wind_profile(*sc_top)))(ground_scene(n=(150, 40, 20))),
    "round": lambda: (lambda sc: (sc, uniform(sc)))(cases.open_block()),
    "round_box": lambda: (lambda sc: (sc, uniform(sc, 0.08, -0.06)))(cases.small_box()),
    "round19": lambda: (lambda sc: (sc, vortex(sc)))(cases.baffled_room(stencil=19)),
    "round_vortex": lambda: (lambda sc: (sc, vortex(sc)))(cases.baffled_room()),
}


def _pair(name, threads=4):
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.reference import ReferenceSolver

    sc, (fx, fy) = CASES[name]()
    return (sc, ReferenceSolver.from_scene(sc, threads=threads, flow_x=fx, flow_y=fy),
            WaveSolver.from_scene(sc, threads=threads, flow_x=fx, flow_y=fy))


@pytest.mark.parametrize("name", sorted(CASES))
def test_flow_layout_equals_the_reference(reference_available, name):
    sc, ref, mine = _pair(name)
    A, B = ref.layout(), mine.layout()
    assert sorted(A) == sorted(B)
    for k in A:
        assert A[k].dtype == B[k].dtype and np.array_equal(A[k], B[k]), k
    assert len(A["flow_mask"]) == sc.grid.cells
    if name.startswith("plain"):
        assert len(A["flow_x"]) == sc.grid.cells and len(A["wind_x"]) == 0
    elif name != "round_box":
        assert len(A["wind_x"]) == sc.grid.cells and len(A["flow_x"]) == 0
        assert np.any(A["wind_x"] != 0) and np.any(A["flow_mask"] == 1)
    else:
        # A closed room holds no wind: the given one, made free of sources inside the walls,
        # is none, and the air stands still to the bit.
        assert not any(np.any(A[f"wind_{ax}"]) for ax in "xyz") and not np.any(A["flow_mask"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_wind_set_up_equals_the_references(reference_available, name):
    """Whether the wind was made to flow round the solids, how fast it then moves and the
    sweeps its solve took: the reference's own; and the air over which the mean square is
    taken is the grid's cells that are not solid."""
    sc, ref, mine = _pair(name)
    want = ref.flow()
    assert want["flowing"] and mine.flowing
    assert (mine.wind_round, mine.wind_fastest, mine.wind_sweeps) == (
        want["wind_round"], want["wind_fastest"], want["wind_sweeps"])
    assert mine.wind_round == (not name.startswith("plain"))
    assert (mine.wind_fastest == 0.0) == (name == "round_box") and mine.wind_fastest < 0.5
    assert ref.air_cells == mine.air_cells == np.count_nonzero(mine.layout()["types"])


def test_flow_refusals_match_the_reference(reference_available):
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.reference import ReferenceSolver

    sc = cases.open_block()
    fx, fy = uniform(sc)
    with pytest.raises(ValueError, match="both of its components"):
        WaveSolver.from_scene(sc, flow_x=fx)
    fast = np.full(sc.grid.cells, 0.6, np.float32)
    for maker in (ReferenceSolver, WaveSolver):
        with pytest.raises(ValueError, match="less than half a cell per step"):
            maker.from_scene(sc, flow_x=fast, flow_y=fy)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_carries_as_the_reference(reference_available, name):
    """The oracle's carrying (oracle/step.c) against the reference's own run, compiled from its
    sources: the probes equal bit for bit over 160 steps, and the air does move them (the
    run differs from the same grid's in still air), except in the closed room, which holds no
    wind."""
    from oracle.stepper import OracleSolver

    sc, ref, _ = _pair(name)
    layout = ref.layout()
    drive = sc.drive[:160]
    mine = OracleSolver(layout)
    a = np.concatenate([mine.run(drive[:61]), mine.run(drive[61:])], axis=1)
    assert np.array_equal(a, ref.run(drive))
    still = {k: v for k, v in layout.items() if k != "flow_mask"}
    b = OracleSolver(still).run(drive)
    assert np.array_equal(a, b) == (name == "round_box")


@pytest.mark.gpu
@pytest.mark.parametrize("name,shape", [("plain", "fused"), ("round", "fused"), ("round_vortex", "fused"),
                                        ("round19", "fused"), ("round", "split"), ("plain", "split"),
                                        ("plain", "no graphs"), ("plain_wide", "fused")])
def test_moving_air_steps_equal_the_reference(reference_available, name, shape):
    """The carrying's three stages for both pressures before each step, then the step: the
    probes equal the reference's CPU run bit for bit, over a run split in two pieces (its
    state, the carrying's included, kept across calls)."""
    from acoustic_spaces_amd.gpu import HipWaveSolver

    from acoustic_spaces_amd.solver import WaveSolver

    print(f"moving air: {name}, {shape}", flush=True)
    sc, ref, mine = _pair(name)
    # The set-up's passes on the device give the host's layout, the air's arrays included.
    fx, fy = CASES[name]()[1]
    layout = WaveSolver.from_scene(sc, threads=4, flow_x=fx, flow_y=fy, on_device=True).layout()
    host = mine.layout()
    for k in host:
        assert np.array_equal(host[k], layout[k]), k
    drive = sc.drive[:240]
    tuning = {"fused": {}, "split": {"split": True}, "no graphs": {"graph_steps": -1}}[shape]
    pieces = (drive[:97], drive[97:])
    a, b, ms_gpu, ms_ref = [], [], [], []
    with HipWaveSolver(layout, **tuning) as gpu:
        for piece in pieces:
            a.append(gpu.run(piece))
            ms_gpu.append(gpu.mean_square)
            b.append(ref.run(piece))
            ms_ref.append(ref.flow()["mean_square"])
        gpu.reset()
        assert gpu.mean_square == 0.0
        again = gpu.run(drive[:40])
    a, b = np.concatenate(a, axis=1), np.concatenate(b, axis=1)
    assert np.array_equal(a, b) and np.any(a != 0)
    assert np.array_equal(again, b[:, :40])
    # The pressure squared over the air, each run's own: the reference's to rounding (its
    # sums run in another order).
    assert all(m > 0 for m in ms_ref)
    np.testing.assert_allclose(ms_gpu, ms_ref, rtol=1e-12)


@pytest.mark.gpu
def test_moving_air_refusals():
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    sc = cases.open_block()
    layout = WaveSolver.from_scene(sc, flow_x=uniform(sc)[0], flow_y=uniform(sc)[1]).layout()
    with pytest.raises(ValueError, match="float only"):
        HipWaveSolver(layout, precision=64)
    with pytest.raises(ValueError, match="still air only"):
        HipWaveSolver(layout, no_faces=True).gradient(sc.drive[:40], np.zeros((3, 40)))
