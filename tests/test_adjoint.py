"""The adjoint (time-reversed) scheme's gradient.

CPU: the adjoint's oracle (oracle/adjoint.py, a double NumPy restatement of the step with
each special's wall term g and each state's gain given) equals the C oracle's double run.
GPU: dJ/dg and dJ/dgain from HipWaveSolver.gradient equal central finite differences of
the NumPy forward (h = 1e-6 on g and gain, J linear in the probes with seeded weights),
are independent of the checkpoint interval, and the float adjoint agrees with the double.
"""

import numpy as np
import pytest

import cases


def _layout(stencil=7, scene=None):
    from acoustic_spaces_amd.solver import WaveSolver

    sc = scene if scene is not None else cases.small_box(stencil=stencil)
    return sc, WaveSolver.from_scene(sc).layout()


def _speed_layout(stencil=7):
    """The small box with a speed of sound per cell in (0.9, 1]."""
    from acoustic_spaces_amd.solver import WaveSolver

    sc = cases.small_box(stencil=stencil)
    return sc, WaveSolver.from_scene(sc, speed=cases.speeds_for(sc)).layout()


def _layer_layout(speed=False):
    """The rigid block in open air with a matched layer of 4 cells round the grid, and a
    fourth probe inside the layer (read from eight layer cells)."""
    from acoustic_spaces_amd.solver import WaveSolver

    sc = cases.open_block()
    sc.probes = np.vstack([sc.probes, [[-0.33, 0.61, 0.52]]])
    kw = {"speed": cases.speeds_for(sc)} if speed else {}
    return sc, WaveSolver.from_scene(sc, **kw).layout()


@pytest.mark.parametrize("case", ["box", "speed", "layer", "layer_speed"])
def test_numpy_forward_equals_the_double_oracle(case):
    from oracle.adjoint import Forward
    from oracle.stepper import OracleSolver

    sc, layout = {"box": _layout, "speed": _speed_layout, "layer": _layer_layout,
                  "layer_speed": lambda: _layer_layout(True)}[case]()
    f = Forward(layout)
    g, gain = f.layout_coefficients(layout)
    drive = sc.drive[:200]
    a = f.run(drive, g, gain, coef_b=np.asarray(layout["coef_b"], np.float64),
              coef_inv=np.asarray(layout["coef_inv"], np.float64))
    b = OracleSolver(layout, 64).run(drive)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


def _fd(f, layout, drive, gain, c, which, index, h):
    """Central differences of J = sum(c y) in a special's wall term g or a state's gain,
    about the layout's own float coefficients: g moves coef_b by -h and 1 / coef_inv by +h
    (coef_b = 1 - g, coef_inv = 1 / (1 + g)), as the adjoint differentiates them."""
    cb = np.asarray(layout["coef_b"], np.float64)
    ci = np.asarray(layout["coef_inv"], np.float64)

    def J(dg, gn):
        return float(np.sum(c * f.run(drive, None, gn, coef_b=cb - dg, coef_inv=1.0 / (1.0 / ci + dg))))
    zero = np.zeros(len(cb))
    if which == "g":
        up, dn = zero.copy(), zero.copy()
        up[index] = h
        dn[index] = -h
        return (J(up, gain) - J(dn, gain)) / (2 * h)
    up, dn = gain.copy(), gain.copy()
    up[index] += h
    dn[index] -= h
    return (J(zero, up) - J(zero, dn)) / (2 * h)


@pytest.mark.gpu
@pytest.mark.parametrize("stencil,listeners", [(7, False), (19, False), (7, True)])
def test_gradient_equals_finite_differences(stencil, listeners):
    """On the small box's probes (every tap an interior cell), and on two listeners' 13
    points each, one listener a cell from a wall: taps on special cells and interior cells,
    cells shared by several probes' taps. The specials checked include the tap cells."""
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from oracle.adjoint import Forward

    sc, layout = _layout(stencil, cases.listeners13() if listeners else None)
    drive = sc.drive[:300]
    rng = np.random.default_rng(5)
    gpu = HipWaveSolver(layout, precision=64, no_faces=True)
    c = rng.normal(size=(gpu.probes, len(drive)))
    y, sens_g, sens_gain = gpu.gradient(drive, c, every=64)
    f = Forward(layout)
    g, gain = f.layout_coefficients(layout)
    # The forward the adjoint differentiates is the stepper's own.
    assert np.allclose(y, f.run(drive, g, gain, coef_b=np.asarray(layout["coef_b"], np.float64),
                                coef_inv=np.asarray(layout["coef_inv"], np.float64)), rtol=0, atol=1e-9 * np.abs(y).max())
    scale_g = np.abs(sens_g).max()
    scale_gain = np.abs(sens_gain).max()
    specials = rng.choice(len(g), 5, replace=False)
    tapped = np.flatnonzero(np.isin(np.asarray(layout["special"]), np.asarray(layout["probe_cells"])))
    if listeners:
        assert len(tapped) > 0
        specials = np.concatenate([specials, rng.choice(tapped, 3, replace=False)])
    for k in specials:
        fd = _fd(f, layout, drive, gain, c, "g", k, 1e-6)
        assert abs(fd - sens_g[k]) <= 1e-6 * scale_g + 1e-5 * abs(fd)
    for j in rng.choice(len(gain), 5, replace=False):
        fd = _fd(f, layout, drive, gain, c, "gain", j, 1e-6)
        assert abs(fd - sens_gain[j]) <= 1e-6 * scale_gain + 1e-5 * abs(fd)


@pytest.mark.gpu
@pytest.mark.parametrize("stencil", [7, 19])
def test_gradient_with_a_speed_per_cell_equals_finite_differences(stencil):
    """A speed of sound per cell (scales in (0.9, 1]): the adjoint steps R = s Q, the
    scaled interior's own update, and each special reads Q = R / s."""
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from oracle.adjoint import Forward

    sc, layout = _speed_layout(stencil)
    drive = sc.drive[:300]
    rng = np.random.default_rng(7)
    gpu = HipWaveSolver(layout, precision=64, no_faces=True)
    c = rng.normal(size=(gpu.probes, len(drive)))
    y, sens_g, sens_gain = gpu.gradient(drive, c, every=64)
    f = Forward(layout)
    g, gain = f.layout_coefficients(layout)
    assert np.allclose(y, f.run(drive, g, gain, coef_b=np.asarray(layout["coef_b"], np.float64),
                                coef_inv=np.asarray(layout["coef_inv"], np.float64)), rtol=0, atol=1e-9 * np.abs(y).max())
    scale_g, scale_gain = np.abs(sens_g).max(), np.abs(sens_gain).max()
    for k in rng.choice(len(g), 6, replace=False):
        fd = _fd(f, layout, drive, gain, c, "g", k, 1e-6)
        assert abs(fd - sens_g[k]) <= 1e-6 * scale_g + 1e-5 * abs(fd), (k, fd, sens_g[k])
    for j in rng.choice(len(gain), 6, replace=False):
        fd = _fd(f, layout, drive, gain, c, "gain", j, 1e-6)
        assert abs(fd - sens_gain[j]) <= 1e-6 * scale_gain + 1e-5 * abs(fd), (j, fd, sens_gain[j])


@pytest.mark.gpu
@pytest.mark.parametrize("speed", [False, True])
def test_gradient_through_the_matched_layer_equals_finite_differences(speed):
    """Open air with a matched layer (its parts and face velocities stepped back by their
    own adjoint kernels, coupled to the cells off the layer through the faces between),
    with and without a speed of sound per cell; a probe reads layer cells."""
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from oracle.adjoint import Forward

    sc, layout = _layer_layout(speed)
    drive = sc.drive[:300]
    rng = np.random.default_rng(11)
    gpu = HipWaveSolver(layout, precision=64, no_faces=True)
    c = rng.normal(size=(gpu.probes, len(drive)))
    y, sens_g, sens_gain = gpu.gradient(drive, c, every=64)
    f = Forward(layout)
    g, gain = f.layout_coefficients(layout)
    assert np.allclose(y, f.run(drive, g, gain, coef_b=np.asarray(layout["coef_b"], np.float64),
                                coef_inv=np.asarray(layout["coef_inv"], np.float64)), rtol=0, atol=1e-9 * np.abs(y).max())
    scale_g, scale_gain = np.abs(sens_g).max(), np.abs(sens_gain).max()
    # Steps of 3e-5: open air leaves J small against the signals, and differences over 1e-6
    # steps wander by rounding (1e-5 of a gain's sensitivity); over 3e-5 they settle to
    # within 1e-7 of those over 1e-4.
    for k in rng.choice(len(g), 8, replace=False):
        fd = _fd(f, layout, drive, gain, c, "g", k, 3e-5)
        assert abs(fd - sens_g[k]) <= 1e-6 * scale_g + 1e-5 * abs(fd), (k, fd, sens_g[k])
    for j in rng.choice(len(gain), 6, replace=False):
        fd = _fd(f, layout, drive, gain, c, "gain", j, 3e-5)
        assert abs(fd - sens_gain[j]) <= 1e-6 * scale_gain + 1e-5 * abs(fd), (j, fd, sens_gain[j])


@pytest.mark.gpu
def test_a_layer_without_the_history_is_refused():
    from acoustic_spaces_amd.gpu import HipWaveSolver

    sc, layout = _layer_layout()
    gpu = HipWaveSolver(layout, precision=64, no_faces=True, keep_history=False)
    with pytest.raises(ValueError, match="matched layer keeps the record's history"):
        gpu.record(sc.drive[:100], every=32)


@pytest.mark.gpu
def test_checkpoint_interval_does_not_change_the_gradient():
    from acoustic_spaces_amd.gpu import HipWaveSolver

    sc, layout = _layout()
    drive = sc.drive[:250]
    gpu = HipWaveSolver(layout, precision=64, no_faces=True)
    c = np.random.default_rng(6).normal(size=(gpu.probes, len(drive)))
    a = gpu.gradient(drive, c, every=7)
    b = gpu.gradient(drive, c, every=1000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_float_adjoint_agrees_with_double():
    from acoustic_spaces_amd.gpu import HipWaveSolver

    sc, layout = _layout()
    drive = sc.drive[:300]
    c = np.random.default_rng(7).normal(size=(4, len(drive)))
    _, g64, n64 = HipWaveSolver(layout, precision=64, no_faces=True).gradient(drive, c)
    _, g32, n32 = HipWaveSolver(layout, precision=32, no_faces=True).gradient(drive, c)
    assert np.linalg.norm(g32 - g64) <= 1e-4 * np.linalg.norm(g64)
    assert np.linalg.norm(n32 - n64) <= 1e-4 * np.linalg.norm(n64)


@pytest.mark.gpu
def test_record_then_adjoint_equals_gradient():
    """The gradient's two halves, as an autograd function calls them (the forward run
    keeping its checkpoints, then the backward pass), give the one call's results; the
    handle is at rest after either."""
    from acoustic_spaces_amd.gpu import HipWaveSolver

    sc, layout = _layout()
    drive = sc.drive[:260]
    gpu = HipWaveSolver(layout, precision=32, no_faces=True)
    c = np.random.default_rng(8).normal(size=(gpu.probes, len(drive)))
    y, g, n = gpu.gradient(drive, c, every=64)
    y2 = gpu.record(drive, every=64)
    g2, n2 = gpu.adjoint(c)
    assert np.array_equal(y, y2) and np.array_equal(g, g2) and np.array_equal(n, n2)
    assert not np.any(gpu.field())
    assert np.array_equal(gpu.run(drive[:50]), y[:, :50])


@pytest.mark.gpu
@pytest.mark.parametrize("precision,stencil", [(32, 7), (64, 7), (32, 19)])
def test_kept_history_equals_recompute(precision, stencil):
    """The record that keeps every step's history of the special cells (so that the
    backward pass recomputes nothing) gives the gradients of the record that keeps only
    checkpoints and recomputes each segment, bit for bit: both are the same kernels'
    arithmetic. Whole segments and a partial last one (250 steps, every 64 and 48), and
    the probes' signals too."""
    from acoustic_spaces_amd.gpu import HipWaveSolver

    sc, layout = _layout(stencil)
    drive = sc.drive[:250]
    rng = np.random.default_rng(17)
    probes = len(layout["probe_start"]) - 1
    c = rng.normal(size=(probes, len(drive)))
    for every in (64, 48):
        out = {}
        for keep in (True, False):
            with HipWaveSolver(layout, precision=precision, no_faces=True, keep_history=keep) as s:
                out[keep] = s.gradient(drive, c, every=every)
        print(f"precision {precision}, stencil {stencil}, every {every}", flush=True)
        for a, b in zip(out[True], out[False]):
            assert np.array_equal(a, b)
        assert np.any(out[True][1] != 0)
