"""The responses' Python side (acoustic_spaces_amd.wave) against what fixes it: the
listener's 13 points; the source signal's filters (the low-pass's unit gain, the double
high-pass's zero, the 19-point taps' gain (5 + 4 cos wT) / 9); the first-order gradient's
exactness on linear fields and its choice of differences; the trapezoidal integral of a
band-limited pulse; WaveResponse.fed, energy, field and high_pass (the wave bands' gains and
the resampled run). GPU: scene_responses' W, X, Y, Z equal the
oracle's run of the same layout and drive, to the bit."""

import math

import numpy as np
import pytest

import cases
from acoustic_spaces_amd import analysis
from acoustic_spaces_amd import wave as wv


def test_probe_points():
    p = wv.probe_points([1.0, 2.0, 3.0], 0.1)
    assert p.shape == (13, 3)
    d = p - p[0]
    assert np.allclose(np.linalg.norm(d[1:7], axis=1), 0.1)
    assert np.allclose(np.linalg.norm(d[7:], axis=1), 0.2)
    assert np.allclose(d[1:7:2] + d[2:7:2], 0.0)  # +axis, then -axis
    assert np.allclose(d[7:], 2 * d[1:7])


@pytest.mark.parametrize("stencil", [7, 19])
def test_source_signal_filters(stencil):
    courant = wv.COURANT if stencil == 7 else wv.ISOTROPIC_COURANT
    rate, cell = 2000.0, 0.1
    s, delay = wv.source_signal(rate, cell, courant)
    bare, bare_delay = wv.source_signal(rate, cell, courant, level=False)
    scale = 4 * math.pi * courant ** 2 / cell
    # Twice high-passed: no DC, and no first moment (it leaves no air in the room).
    assert abs(bare.sum()) < 1e-12 * scale
    assert abs(s.sum()) < 1e-12 * scale
    if stencil == 7:
        assert delay == bare_delay == wv.LOW_PASS_TAPS // 2
        assert np.array_equal(s, bare)
        return
    assert delay == bare_delay + 1
    # The taps' gain is (5 + 4 cos wT) / 9, the inverse of the scheme's at the source.
    n = 4096
    S, B = np.fft.rfft(s, n), np.fft.rfft(bare, n)
    w = 2 * math.pi * np.fft.rfftfreq(n)
    phase = np.exp(-1j * w)  # the extra step of delay
    mask = np.abs(B) > 1e-6 * np.abs(B).max()
    assert np.allclose(S[mask] / (B[mask] * phase[mask]), (5 + 4 * np.cos(w[mask])) / 9, rtol=1e-9, atol=1e-12)


def _field(f, g, q=None):
    """p at the 13 points of a listener, one cell = 1: f(t) + g(t) x (+ q(t) x^2)."""
    x = wv.probe_points([0.0, 0.0, 0.0], 1.0)[:, 0]
    rows = f[None, :] + x[:, None] * g[None, :]
    if q is not None:
        rows = rows + (x ** 2)[:, None] * q[None, :]
    return rows


def test_first_order_gradient_choices():
    rng = np.random.default_rng(3)
    n = 400
    f, g, q = rng.normal(size=(3, n))
    # A uniform field has no gradient.
    out = wv.first_order(_field(f, 0 * g), [True] * 13)
    assert np.array_equal(out[:, 0], f) and np.allclose(out[:, 1:], 0.0, atol=1e-12)
    # On a linear field every difference is exact: the fourth-order, the second-order and
    # the one-sided choices agree.
    rows = _field(f, g)
    full = wv.first_order(rows, [True] * 13)
    second = wv.first_order(rows, [True] * 7 + [False] * 6)
    plus = wv.first_order(rows, [True, True, False] + [True] * 10)
    minus = wv.first_order(rows, [True, False, True] + [True] * 10)
    for other in (second, plus, minus):
        assert np.allclose(other[:, 1], full[:, 1], rtol=0, atol=1e-9 * np.abs(full[:, 1]).max())
    # A quadratic one tells them apart: the central differences see no gradient at the
    # listener, the one-sided ones h q and -h q.
    rows = _field(f, 0 * g, q)
    assert np.allclose(wv.first_order(rows, [True] * 13)[:, 1], 0.0, atol=1e-9)
    assert np.allclose(wv.first_order(rows, [True] * 7 + [False] * 6)[:, 1], 0.0, atol=1e-9)
    a = wv.first_order(rows, [True, True, False] + [True] * 10)[:, 1]
    b = wv.first_order(rows, [True, False, True] + [True] * 10)[:, 1]
    assert np.allclose(a, -b, atol=1e-9) and np.abs(a).max() > 1.0
    with pytest.raises(ValueError, match="not in the air"):
        wv.first_order(rows, [False] + [True] * 12, "far")


def test_first_order_integrates_a_pulse():
    n = np.arange(2000)
    d = np.exp(-(((n - 400) / 25.0) ** 2))  # its integral is 25 sqrt(pi)
    f = np.zeros(len(n))
    out = wv.first_order(_field(f, d), [True] * 13, courant=1.0)
    # After the pulse the integral holds its area (away from the record's end, where the
    # trapezoid's gain correction rings on the cut-off); halfway through it, half; before
    # it, nothing.
    area = 25 * math.sqrt(math.pi)
    assert out[800:1500, 1] == pytest.approx(np.full(700, area), rel=1e-4)
    assert out[400, 1] == pytest.approx(area / 2, rel=1e-3)
    assert abs(out[250, 1]) < 1e-2


def test_response_fed_and_energy():
    rate = 4000.0
    t = np.arange(int(rate)) / rate
    p = np.sin(2 * math.pi * analysis.octave_midbands()[1] * t)
    sig = np.column_stack([p, p, 0 * p, 0 * p])
    r = wv.WaveResponse(rate, sig, 3, 0.1)
    e = r.energy(1e-3, 1000)
    assert np.allclose(e, analysis.band_energy(p, rate, 3, 1e-3, 1000))
    # A figure of eight along x, where X equals W: the same energy either way.
    assert np.allclose(r.energy(1e-3, 1000, axis=[1.0, 0.0, 0.0]), e)
    assert np.allclose(r.energy(1e-3, 1000, axis=[1.0, 0.0, 0.0], exponent=1), e)
    g = r.fed([1.0, 4.0, 9.0], 0.0)
    assert np.allclose(g.energy(1e-3, 1000), e * np.array([1.0, 4.0, 9.0])[:, None])
    later = g.fed([1.0, 1.0, 1.0], 0.25)
    assert later.delay_s == 0.25 and np.array_equal(later.band_gain, g.band_gain)
    # 0.25 s later: the same energy in all, its centre of gravity 0.25 s later (less what
    # the filters spread ahead of the tone's start, which undelayed counts in the first
    # slot, and a sample on a slot's edge, which may fall either side).
    a, b = later.energy(1e-3, 1250), g.energy(1e-3, 1000)
    assert np.allclose(a.sum(axis=1), b.sum(axis=1), rtol=1e-9)
    centroid = lambda x: (x * (np.arange(x.shape[1]) + 0.5)).sum(axis=1) / x.sum(axis=1) * 1e-3  # noqa: E731
    assert np.allclose(centroid(a) - centroid(b), 0.25, atol=2e-3)


def test_wave_gains_rise_from_the_low_edge():
    """The wave bands' gains are the synthesis's band gains, the lowest one rising from 0
    below LOW_EDGE_HZ / 2^(1/4) to its band gain above LOW_EDGE_HZ 2^(1/4) (a raised cosine
    half an octave wide, 1/2 at the edge). All eight together are that rise alone."""
    rate, n_fft = 6000.0, 1 << 16
    f = np.fft.rfftfreq(n_fft, 1.0 / rate)
    g = wv.wave_gains(n_fft, rate, 8)
    band = analysis.band_gains(n_fft, rate, 8)
    assert np.array_equal(g[1:], band[1:])
    rise = g.sum(axis=0)
    assert np.max(rise[f < wv.LOW_EDGE_HZ * 2 ** -0.25]) == 0.0
    assert np.allclose(rise[f > wv.LOW_EDGE_HZ * 2 ** 0.25], 1.0, atol=1e-12)
    assert abs(np.interp(wv.LOW_EDGE_HZ, f, rise) - 0.5) < 1e-3
    assert np.allclose(wv.wave_gains(n_fft, rate, 3), g[:3])


def test_response_field_and_high_pass_split_the_run():
    """field() and high_pass() (1 - g) of the same resampled run add up to it (the wave
    bands' gains g and 1 - g); late=True filters at sqrt(1 - g^2); the run's start sits
    (delay_s - lead_s) from emission, an impulse's height scaled by rate / sample_rate; a
    band's feed scales its share."""
    rng = np.random.default_rng(8)
    rate, sr, n = 2000.0, 8000.0, 6000
    sig = rng.standard_normal((1200, 4))
    r = wv.WaveResponse(rate, sig, 3, 0.1, lead_s=0.01, delay_s=0.05)
    run, before = r._timeline(sr, n)
    start = int(round((0.05 - 0.01) * sr))
    assert before == 0 and run.shape == (max(n, start + 4800), 4)
    assert np.allclose(run[start:start + 4800], analysis.resample(sig, rate, sr) * (rate / sr))
    field = r.field(sr, n)
    assert field.shape == (n, 4)
    assert np.allclose(field + r.high_pass(run, sr)[:n], run[:n], atol=1e-12)
    n_fft = analysis.fft_size(len(run), sr)
    g = wv.wave_gains(n_fft, sr, 3).sum(axis=0)
    late = r.high_pass(run, sr, late=True)
    assert np.allclose(late, analysis.zero_phase(run, np.sqrt(1.0 - g ** 2), n_fft))
    # A lead longer than the delay: samples before emission, filtered with the rest and
    # left out of the field, which is the same run's 720 samples later.
    early = wv.WaveResponse(rate, sig, 3, 0.1, lead_s=0.05)
    run2, before2 = early._timeline(sr, n)
    assert before2 == int(round(0.05 * sr))
    shift = start + before2
    assert np.allclose(early.field(sr, n), r.field(sr, n + shift)[shift:shift + n], atol=1e-9)
    # Feeding each band: the field of one band alone scales with its amplitude.
    one = wv.WaveResponse(rate, sig, 1, 0.1)
    assert np.allclose(one.fed([4.0], 0.0).field(sr, n), 2.0 * one.field(sr, n))


@pytest.mark.gpu
def test_wave_responses_equal_the_oracle():
    from acoustic_spaces_amd.solver import WaveSolver
    from oracle.stepper import OracleSolver
    import dataclasses

    scene = cases.small_box(seed=31)
    listeners = cases.listeners13().probes[[0, 13]]
    print("wave_responses small box", flush=True)
    responses, info = wv.scene_responses(scene, listeners, ["a", "b"], length_s=0.05)
    points = np.vstack([wv.probe_points(p, scene.grid.cell) for p in listeners])
    solver = WaveSolver.from_scene(dataclasses.replace(scene, probes=points), probes_per_listener=13)
    signal, delay = wv.source_signal(scene.rate, scene.grid.cell)
    steps = delay + int(math.ceil(0.05 * scene.rate))
    drive = np.zeros(steps, np.float32)
    drive[: min(len(signal), steps)] = signal[:steps]
    recorded = OracleSolver(solver.layout()).run(drive)
    found = [solver.probe_found(p) and solver.probe_in_sight(p) for p in range(26)]
    found = wv.off_the_source(found, points, scene.source, scene.grid.cell, 7).tolist()
    for i, r in enumerate(responses):
        expect = wv.first_order(recorded[13 * i:13 * (i + 1)], found[13 * i:13 * (i + 1)])
        assert np.array_equal(r.signals, expect)
        assert r.lead_s == delay / scene.rate and r.rate == scene.rate
    assert info["wave_steps"] == steps and info["wave_gpu"] == 1.0


def test_high_pass_is_the_source_signals():
    """Driving with a pulse and high-passing what is received equals driving with the
    high-passed pulse (one linear filter either side): checked on the filter itself, the
    bare source signal being the low-pass followed by this high-pass."""
    rate, cell = 2000.0, 0.1
    bare, delay = wv.source_signal(rate, cell, level=False)
    half = wv.LOW_PASS_TAPS // 2
    k = np.arange(-half, half + 1)
    c = wv.cutoff(wv.COURANT)
    low = 2 * c * np.sinc(2 * c * k) * np.blackman(wv.LOW_PASS_TAPS)
    low /= low.sum()
    x = np.zeros(len(bare))
    x[: len(low)] = low * 4 * math.pi * wv.COURANT ** 2 / cell
    assert np.allclose(wv.high_pass(x, rate), bare, atol=1e-12 * np.abs(bare).max())
    # No DC passes.
    assert abs(wv.high_pass(np.ones(5000), rate)[-1]) < 1e-9
