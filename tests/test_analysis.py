"""Octave-band levels and decay times (acoustic_spaces_amd.analysis) against what the
mathematics fixes: the Butterworth band-pass's values at mid-band, at its -3 dB edges and
its geometric symmetry; the fixed-order dot's exact sums and its order; the Schroeder
curve and T30 of an exponential decay; the strength of a known energy."""

import math

import numpy as np
import pytest

from acoustic_spaces_amd import analysis as an


def test_band_pass_closed_form():
    fm = an.octave_midbands()
    assert np.allclose(fm, 1000.0 * 10 ** (0.3 * np.arange(-4, 4)), rtol=1e-12)
    g = an.band_pass(fm)
    assert np.allclose(np.diag(g), 1.0, rtol=0, atol=1e-15)
    half = math.sqrt(an.OCTAVE_RATIO)
    for b in range(8):
        edges = an.band_pass(np.array([fm[b] / half, fm[b] * half]))[b]
        assert np.allclose(edges, 1 / math.sqrt(2.0), rtol=1e-12)
        # Geometric symmetry about fm, and about -20 dB at the neighbouring mid-bands.
        f = np.array([fm[b] * 1.37, fm[b] / 1.37])
        h = an.band_pass(f)[b]
        assert h[0] == pytest.approx(h[1], rel=1e-12)
        nb = 20 * math.log10(an.band_pass(np.array([fm[b] * an.OCTAVE_RATIO]))[b, 0])
        assert -20.5 < nb < -19.0


def test_fixed_dot_sums_exactly_and_in_its_order():
    rng = np.random.default_rng(1)
    x = rng.integers(-1000, 1000, 1003).astype(float)
    y = rng.integers(-1000, 1000, 1003).astype(float)
    assert an.fixed_dot(x, y) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))
    # Lane 0 takes 1e16, then 1 (lost to rounding), then -1e16: 0, where an exact sum is 1.
    z = np.zeros(17)
    z[0], z[8], z[16] = 1e16, 1.0, -1e16
    assert an.fixed_dot(z, np.ones(17)) == 0.0
    assert math.fsum(z) == 1.0
    a, b, c = an.fixed_dots(x, y)
    assert (a, b, c) == (an.fixed_dot(x, x), an.fixed_dot(y, y), an.fixed_dot(x, y))


@pytest.mark.parametrize("t60", [0.4, 1.2, 2.5])
def test_t30_of_an_exponential_decay(t60):
    dt = 1e-3
    k = np.arange(int(3.0 / dt))
    e = np.exp(-k * dt * 6 * math.log(10) / t60)
    assert an.band_t30(e, dt) == pytest.approx(t60, rel=1e-5)
    curve = an.decay_curve(e, dt)
    # The curve of an exponential, tail included, is a straight line of -60 / T60 dB/s,
    # closely down to -40 dB; near the end it rests on the estimated tail.
    line = -60.0 / t60 * k * dt
    upper = line > -40.0
    assert np.allclose(curve[upper], line[upper], rtol=0, atol=2e-3)
    assert np.allclose(curve, line, rtol=0, atol=0.5)


def test_tail_of_a_truncated_exponential():
    """A record that has decayed far past its fit range: the line through its curve is the
    exponential's, and extended past the end it gives the energy left to come. (A short
    record's curve bends within the fit range, so its estimate is only near.)"""
    dt = 1e-3
    for t60, n, rel in ((0.3, 600, 1e-3), (1.0, 600, 0.15)):
        k = np.arange(n)
        decay = 6 * math.log(10) / t60
        e = np.exp(-k * dt * decay)
        extra, rate = an.tail(e, dt)
        exact = np.exp(-n * dt * decay) / (1 - np.exp(-dt * decay))
        assert extra == pytest.approx(exact, rel=rel)
        assert rate == pytest.approx(decay, rel=rel)


def test_strength_of_a_known_energy():
    e = np.zeros(100)
    e[0] = 0.01
    assert an.band_strength(e, 1e-3) == pytest.approx(0.0, abs=1e-12)
    e[0] = 1.0
    assert an.band_strength(e, 1e-3) == pytest.approx(20.0, abs=1e-12)


def test_band_energy_of_a_tone_sits_in_its_band():
    rate = 8000.0
    t = np.arange(int(2 * rate)) / rate
    fm = an.octave_midbands()[1]
    e = an.band_energy(np.sin(2 * math.pi * fm * t), rate, 3, 1e-3, 2000)
    totals = e.sum(axis=1)
    # The neighbouring bands pass it at about -20 dB (and count over their own shares).
    assert totals[1] > 20 * totals[0] and totals[1] > 20 * totals[2]
    # The band's energy: the tone's sum of squares (N / 2, the filter passing it at 1)
    # over 4 pi and its filter's share of the spectrum.
    n_fft = an.fft_size(len(t) + int(round(an.PAD_S * rate)), rate)
    g = an.band_filter_gains(n_fft, rate)[1]
    w = np.full(len(g), 2.0)
    w[0] = w[-1] = 1.0
    share = float((g ** 2 * w).sum() / n_fft)
    assert totals[1] == pytest.approx(0.5 * len(t) / (an.ENERGY_TO_IR * share), rel=0.01)


# Resampling and the synthesis's band split (render.py:105-162, synthesis.py:50-71).

def _tone(rate, n, hz, phase=0.3):
    return np.sin(2 * np.pi * hz * np.arange(n) / rate + phase)


@pytest.mark.parametrize("rate_in,rate_out,hz", [
    (2000.0, 44100.0, 100.0),     # up, 441 phases (exact)
    (5940.93, 48000.0, 180.0),    # up, no short ratio (rounded phases)
    (44100.0, 2000.0, 150.0),     # down, 20 / 441 (exact)
    (11881.87, 8000.0, 400.0),    # down, rounded phases
])
def test_resample_keeps_a_band_limited_tone(rate_in, rate_out, hz):
    """A tone well inside both Nyquist bands comes out as the same tone at the new rate:
    the sample times (output k at input position k rate_in / rate_out), the phases and the
    filter's unit gain in its pass band. A shift of one sample or one phase would leave an
    error of order 2 pi hz / rate; rounding a position to one of the 4096 phases leaves at
    most 2 pi hz / rate / 8192."""
    from fractions import Fraction

    n = int(2.0 * rate_in)
    y = an.resample(_tone(rate_in, n, hz), rate_in, rate_out)
    assert len(y) == math.ceil(n * rate_out / rate_in)
    want = _tone(rate_out, len(y), hz)
    inner = slice(len(y) // 5, 4 * len(y) // 5)  # away from the zero padding at the ends
    exact = (Fraction(rate_in) / Fraction(rate_out)).limit_denominator(1 << 20).denominator <= 4096
    bound = 1e-5 if exact else 1e-5 + 1.2 * 2 * math.pi * hz / rate_in / (2 * an.RESAMPLE_MAX_PHASES)
    assert np.max(np.abs(y[inner] - want[inner])) < bound


def test_resample_removes_what_would_alias():
    """Down to 2 kHz, a tone at 1.6 kHz (above the new Nyquist frequency) is filtered out,
    and one at 300 Hz kept."""
    rate_in, rate_out = 8000.0, 2000.0
    n = 16000
    inner = slice(1000, 3000)
    high = an.resample(_tone(rate_in, n, 1600.0), rate_in, rate_out)
    low = an.resample(_tone(rate_in, n, 300.0), rate_in, rate_out)
    assert np.max(np.abs(high[inner])) < 1e-4
    assert abs(np.max(np.abs(low[inner])) - 1.0) < 1e-3


def test_resample_shapes_and_identity():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((500, 3))
    assert np.array_equal(an.resample(x, 1000.0, 1000.0), x)
    two = an.resample(x, 1000.0, 2205.0)
    assert two.shape == (1103, 3)
    for c in range(3):
        assert np.array_equal(two[:, c], an.resample(x[:, c], 1000.0, 2205.0))
    with pytest.raises(ValueError):
        an.resample(x, 0.0, 1000.0)


def test_ordered_rows_dot_adds_in_tap_order():
    """Each row's terms from the first tap to the last, one addition at a time: (1e16 + 1) -
    1e16 is 0, (1e16 - 1e16) + 1 is 1."""
    ones = np.ones(3)
    assert an.ordered_rows_dot(np.array([[1e16, 1.0, -1e16]]), ones)[0] == 0.0
    assert an.ordered_rows_dot(np.array([[1e16, -1e16, 1.0]]), ones)[0] == 1.0
    assert an.ordered_rows_dot(np.array([[2.0, 3.0], [1.0, -1.0]]), np.array([0.5, 2.0])).tolist() == [7.0, -1.5]


@pytest.mark.parametrize("bands", [8, 24])
def test_band_gains_split_the_spectrum(bands):
    """The bands' gains sum to one at every frequency (perfect reconstruction); each is 1
    at its nominal centre, 1/2 at each of its edges (a raised-cosine crossover, half a band
    wide in log frequency), and 0 a crossover's half width beyond an edge."""
    rate, n_fft = 48000.0, 1 << 16
    g = an.band_gains(n_fft, rate, bands)
    f = np.fft.rfftfreq(n_fft, 1.0 / rate)
    assert g.shape == (bands, n_fft // 2 + 1)
    assert np.max(np.abs(g.sum(axis=0) - 1.0)) < 1e-12
    centres = an.OCTAVE_BANDS_HZ if bands == 8 else an.THIRD_OCTAVE_BANDS_HZ
    edges = an.crossover_frequencies(bands)
    width = an.crossover_width(bands)
    for b in range(1, bands - 1):
        assert abs(np.interp(centres[b], f, g[b]) - 1.0) < 1e-3
        for e in (edges[b - 1], edges[b]):
            assert abs(np.interp(e, f, g[b]) - 0.5) < 2e-3
        assert np.interp(edges[b] * 2 ** (0.6 * width), f, g[b]) < 1e-12
    assert abs(edges[0] - (63.0 * math.sqrt(2.0) if bands == 8 else 1000.0 * 2 ** (-12.5 / 3))) < 1e-9
