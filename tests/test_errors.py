"""Edge cases and refusals of the stepper, as the reference's stepper meets them: a layout
it cannot step raises ValueError (engine/gpu.py:440-441, 495-504), as does an adjoint asked
of a grid it does not cover; runs of no steps, grids with no probes and no special cells,
and runs after reset behave as the reference's run() does (fdtd.hpp:786-915)."""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _layout(scene=None):
    from acoustic_spaces_amd.solver import WaveSolver

    return WaveSolver.from_scene(scene or cases.small_box()).layout()


def _bad(layout, key, fix):
    L = {k: np.array(v, copy=True) for k, v in layout.items()}
    fix(L[key]) if key in L else None
    return L


@pytest.mark.parametrize("what,key,fix,match", [
    ("a class past 3", "types", lambda a: a.__setitem__(5, 7), "type"),
    ("a special on the edge", "special", lambda a: a.__setitem__(0, 0), "edge"),
    ("a link with no direction", "link_direction", lambda a: a.__setitem__(0, 18), "direction"),
    ("a state with no branch", "state_branch", lambda a: a.__setitem__(0, 1 << 20), "branch"),
    ("a probe off the grid", "probe_cells", lambda a: a.__setitem__(0, 1 << 40), "probe"),
])
def test_refused_layouts_raise_value_error(what, key, fix, match):
    from acoustic_spaces_amd.gpu import HipWaveSolver

    L = _bad(_layout(), key, fix)
    print(f"refused: {what}", flush=True)
    with pytest.raises(ValueError, match=match):
        HipWaveSolver(L)


def test_refused_arguments():
    from acoustic_spaces_amd.gpu import HipWaveSolver

    L = _layout()
    with pytest.raises(ValueError, match="precision"):
        HipWaveSolver(L, precision=16)
    with pytest.raises(ValueError, match="row alignment"):
        HipWaveSolver(L, row_align=8)
    small = dict(L)
    small["shape"] = np.array([2, 5, 5], np.uint64)
    with pytest.raises(ValueError):
        HipWaveSolver(small)


def test_no_steps_no_probes_no_specials():
    """A run of no steps returns (probes, 0); a grid with no probes runs and returns
    (0, steps); a closed box of rigid walls with no branch states and a run after reset
    repeat themselves."""
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from oracle.stepper import OracleSolver

    scene = cases.small_box(branches=False)
    L = _layout(scene)
    gpu = HipWaveSolver(L)
    assert gpu.run(np.zeros(0, np.float32)).shape == (gpu.probes, 0)
    quiet = dict(L)
    quiet["probe_start"] = np.zeros(1, np.uint64)
    quiet["probe_cells"] = np.zeros(0, np.uint64)
    quiet["probe_weights"] = np.zeros(0, np.float64)
    g = HipWaveSolver(quiet)
    assert g.run(scene.drive[:70]).shape == (0, 70)
    cpu = OracleSolver(quiet)
    cpu.run(scene.drive[:70])
    assert np.array_equal(g.field(), cpu.field())
    g.reset()
    assert not np.any(g.field())


def test_adjoint_refusals():
    from acoustic_spaces_amd.gpu import HipWaveSolver

    L = _layout()
    drive = cases.small_box().drive[:40]
    c = np.zeros((len(L["probe_start"]) - 1, len(drive)))
    with pytest.raises(ValueError, match="no_faces"):
        HipWaveSolver(L).gradient(drive, c)
    # A grid with a matched layer keeps the record's history: refused without it.
    open_layout = _layout(cases.open_block())
    with pytest.raises(ValueError, match="matched layer keeps the record's history"):
        HipWaveSolver(open_layout, no_faces=True, keep_history=False).gradient(cases.open_block().drive[:40],
                                                                               np.zeros((3, 40)))
    with pytest.raises(ValueError, match="every 0"):
        HipWaveSolver(L, no_faces=True).gradient(drive, c, every=0)
    with pytest.raises(ValueError, match="record"):
        HipWaveSolver(L, no_faces=True).adjoint(c)


def _probes_with(layout, start):
    L = dict(layout)
    L["probe_start"] = np.asarray(start, np.uint64)
    return L


@pytest.mark.parametrize("what,start,match", [
    ("ranges not from 0", lambda s: np.r_[1, s[1:]], "start at 0"),
    ("a range running backwards", lambda s: np.r_[s[0], s[2], s[1], s[3:]], "backwards"),
    ("ranges past the taps", lambda s: np.r_[s[:-1], s[-1] + 1], "end at the taps"),
])
def test_refused_probe_ranges(what, start, match):
    """probe_start must run from 0 to the taps, never decreasing: a range running
    backwards would make a probe's wavefront loop about 2^32 times, one past the taps read
    past its cells. Refused on the host, before any device work."""
    from acoustic_spaces_amd.gpu import HipWaveSolver

    L = _layout()
    s = np.asarray(L["probe_start"], np.int64)
    assert len(s) >= 4 and s[1] < s[2]
    print(f"refused: {what}", flush=True)
    with pytest.raises(ValueError, match=match):
        HipWaveSolver(_probes_with(L, start(s)))
