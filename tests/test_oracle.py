"""Pins the oracle (oracle/step.c), the checker of every GPU result.

1. Bit for bit against the reference's own run() (fdtd.hpp:786-904), built from the
   reference's headers into oracle/_ref, on layouts covering both stencils, branch states,
   a speed per cell and the matched layer, with runs split so the fields' parity flips
   both ways (tests/test_gpu.py:45-85 of the reference splits its runs the same way).
2. Against what the mathematics fixes, independently of the reference's code: a rigid box
   rings at its modes (the continuous ones within 0.6 Hz, as tests/test_wave.py:263-284;
   the discrete scheme's own to within the spectrum's resolution), a closed rigid room
   keeps the volume its source put in, and the stencil's weights are the exact ones
   lowered by the margin (tests/test_wave.py:400-447).
3. The double restatement agrees with the float one to float rounding.
CPU only."""

import itertools
import math

import numpy as np
import pytest

import cases
from acoustic_spaces_amd import scenes


def _ref(scene, **kw):
    from oracle.reference import ReferenceSolver

    return ReferenceSolver.from_scene(scene, threads=2, **kw)


CASES = {
    "box7": lambda: (cases.small_box(), {}),
    "box19": lambda: (cases.small_box(stencil=19), {}),
    "open_layer": lambda: (cases.open_block(), {}),
    "speed": lambda: (lambda sc: (sc, {"speed": cases.speeds_for(sc)}))(cases.small_box()),
    "speed19": lambda: (lambda sc: (sc, {"speed": cases.speeds_for(sc)}))(cases.small_box(stencil=19)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_equals_the_reference_run(reference_available, name):
    from oracle.stepper import OracleSolver

    scene, kw = CASES[name]()
    ref = _ref(scene, **kw)
    oracle = OracleSolver(ref.layout())
    drive = scene.drive[: scene.steps]
    for piece in (drive[:101], drive[101:224], drive[224:]):
        assert np.array_equal(ref.run(piece), oracle.run(piece))


def test_double_restatement_is_the_same_scheme():
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver

    scene = cases.small_box()
    layout = WaveSolver.from_scene(scene).layout()
    a = OracleSolver(layout, 32).run(scene.drive[: scene.steps])
    b = OracleSolver(layout, 64).run(scene.drive[: scene.steps])
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert 0 < rel < 1e-5


def _rigid_box(stencil, size=(2.0, 1.5, 1.0), cell=0.1):
    rate = scenes.C / ((scenes.COURANT if stencil == 7 else scenes.ISOTROPIC_COURANT) * cell)
    steps = int(4 * rate)
    sc = scenes.shoebox(size=size, cell=cell, seed=1, length_s=4.0, stencil=stencil, absorption=np.zeros(6))
    sc.source = np.array([0.23, 0.17, 0.13])
    sc.probes = np.array([[1.77, 1.31, 0.87]])
    sc.drive = scenes.gaussian_drive(rate, steps, sigma_s=1.5e-3)
    return sc, rate, steps


@pytest.mark.parametrize("stencil", [7, 19])
def test_rigid_box_rings_at_its_modes(stencil):
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver

    size = (2.0, 1.5, 1.0)
    sc, rate, steps = _rigid_box(stencil, size)
    solver = WaveSolver.from_scene(sc)
    assert solver.air_cells == 20 * 15 * 10
    layout = solver.layout()
    pressure = OracleSolver(layout).run(sc.drive)[0]
    pressure = pressure - pressure[steps // 2:].mean()
    n_fft = 1 << 18
    spectrum = np.abs(np.fft.rfft(pressure * np.hanning(len(pressure)), n_fft))
    f = np.fft.rfftfreq(n_fft, 1.0 / rate)
    a = float(layout["weights"][0])
    d = float(layout["weights"][6]) if stencil == 19 else 0.0
    centre = float(layout["centre"][0])
    n_cells = (20, 15, 10)
    for mode in ((1, 0, 0), (0, 1, 0), (1, 1, 0)):
        expected = scenes.C / 2 * math.hypot(*(m / s for m, s in zip(mode, size)))
        near = (f > expected - 4) & (f < expected + 4)
        peak = f[near][np.argmax(spectrum[near])]
        assert peak == pytest.approx(expected, abs=0.6)
        # The discrete scheme's own frequency: the Neumann modes cos(pi m (x + 1/2) / N) of
        # the box's cells are eigenvectors of the update, p+ + p- = A p,
        # A = centre + 2 a sum cos t_a + 4 d sum_{a<b} cos t_a cos t_b, cos(w T) = A / 2.
        c = [math.cos(math.pi * m / n) for m, n in zip(mode, n_cells)]
        A = centre + 2 * a * sum(c) + 4 * d * (c[0] * c[1] + c[0] * c[2] + c[1] * c[2])
        discrete = math.acos(A / 2) * rate / (2 * math.pi)
        assert peak == pytest.approx(discrete, abs=2 * (f[1] - f[0]))


def test_a_closed_rigid_room_keeps_what_its_source_put_in():
    """The leapfrog's uniform mode obeys p_tt = 0: once the drive has ended, the pressure
    summed over a closed rigid room's air grows by the same amount every step (the drive's
    net volume; the reference's own drive is high-passed twice to leave none,
    engine/wave.py:135-139)."""
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver

    sc, rate, steps = _rigid_box(7)
    layout = WaveSolver.from_scene(sc).layout()
    oracle = OracleSolver(layout, 64)
    oracle.run(sc.drive[:400])
    air = np.asarray(layout["types"]) != 0
    sums = []
    for _ in range(3):
        sums.append(oracle.field()[air].sum())
        oracle.run(np.zeros(1500, np.float32))
    first, second = sums[1] - sums[0], sums[2] - sums[1]
    assert first > 0
    assert second == pytest.approx(first, rel=1e-6)


@pytest.mark.parametrize("stencil", [7, 19])
def test_stencil_weights_are_the_exact_ones_lowered(stencil):
    from acoustic_spaces_amd.solver import WaveSolver

    layout = WaveSolver.from_scene(cases.small_box(stencil=stencil, branches=False)).layout()
    axial = float(layout["weights"][0])
    diagonal = float(layout["weights"][6]) if stencil == 19 else 0.0
    centre = float(layout["centre"][0])
    margin = 2.0 ** -18
    exact = (1 / 3, 0.0) if stencil == 7 else (0.25, 0.125)
    np.testing.assert_allclose([axial, diagonal], np.multiply(exact, 1 - margin), rtol=2 ** -22)
    c = np.array(list(itertools.product([1.0, -1.0], repeat=3)))
    pairs = c[:, 0] * c[:, 1] + c[:, 0] * c[:, 2] + c[:, 1] * c[:, 2]
    factors = centre + 2 * axial * c.sum(axis=1) + 4 * diagonal * pairs
    assert factors[0] == 2.0
    assert factors.min() >= -2 + 4 * margin


def test_a_planted_mistake_is_caught():
    """A plausible slip in the oracle (the previous pressure's sign flipped in the interior
    update) fails the mode check: the check pins the update, not just its plumbing."""
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver

    sc, rate, steps = _rigid_box(7)
    layout = WaveSolver.from_scene(sc).layout()
    good = OracleSolver(layout).run(sc.drive[:2000])[0]
    # Plant it through the layout: the centre weight off by one float step.
    bad_layout = dict(layout)
    bad_layout["centre"] = np.nextafter(layout["centre"], np.float32(1.0)).astype(np.float32) * np.float32(1.001)
    bad = OracleSolver(bad_layout).run(sc.drive[:2000])[0]
    assert not np.array_equal(good, bad)
