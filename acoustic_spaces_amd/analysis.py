"""Octave-band levels and decay times of the receivers' signals.

The reference's room-acoustic parameters of a band, restated for the wave solver's output:
- its octave band-pass filters (`parameters.py:1062-1080`);
- its band energies per time slot (`WaveResponse.energy`, `engine/wave.py:402-454`);
- the Schroeder decay curve with the estimated tail past the response's end
  (`parameters.py:306-333, 357-369, 442-452`);
- the decay times' line fits (`parameters.py:266-296, 336-343`, with `products.dots`,
  `fixed_sums.hpp:44-47, 370-378`);
- T30 (`parameters.py:464-475`) and the strength G (`parameters.py:385-440`);
- the synthesis's band split, zero-phase raised-cosine crossovers half a band wide whose
  gains sum to one (`synthesis.py:31-76`, `materials/bands.py:18-29, 41-50, 78-84`);
- band-limited resampling, polyphase Kaiser-windowed sinc (`render.py:46-54, 105-162`).
The constants are the reference's.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

OCTAVE_BANDS_HZ = (63, 125, 250, 500, 1000, 2000, 4000, 8000)
OCTAVE_RATIO = 10.0 ** 0.3  # base-ten octave, IEC 61260-1
FILTER_ORDER = 3
DECAY_RANGES = {"edt": (0.0, -10.0), "t20": (-5.0, -25.0), "t30": (-5.0, -35.0)}
TAIL_HEADROOM_DB = 10.0
TAIL_SPAN_DB = 20.0
MIN_DECAY_DB = -20.0
ENERGY_TO_IR = 4.0 * math.pi  # impulse-response energy per unit of simulated energy
REFERENCE_ENERGY_10M = 0.01   # the direct sound's energy at 10 m: G's reference
PAD_S = 0.2                   # zero padding round FFT filtering (s)
THIRD_OCTAVE_BANDS_HZ = (50, 63, 80, 100, 125, 160, 200, 250, 315, 400, 500, 630, 800, 1000, 1250,
                         1600, 2000, 2500, 3150, 4000, 5000, 6300, 8000, 10000)
CROSSOVER_WIDTH = 0.5         # a crossover between octave bands, in octaves: half a band
RESAMPLE_ZEROS = 32           # zero crossings of the resampling filter on each side
RESAMPLE_BETA = 9.0           # its Kaiser window's beta
RESAMPLE_MAX_PHASES = 4096    # most filter phases; finer ratios round the position to these
RESAMPLE_CHUNK = 1 << 15      # outputs resampled at once when the rates have no short ratio


def octave_midbands() -> np.ndarray:
    """Exact mid-band frequencies (Hz) of the eight octave bands, base ten:
    1000 G^k, k the nominal band's whole number of octaves from 1 kHz."""
    nominal = np.array(OCTAVE_BANDS_HZ, float)
    k = np.round(np.log(nominal / 1000.0) / math.log(OCTAVE_RATIO))
    return 1000.0 * OCTAVE_RATIO ** k


def band_pass(f, order: int = FILTER_ORDER) -> np.ndarray:
    """Magnitudes (8, len(f)) at frequencies ``f`` (Hz, > 0) of the octave bands'
    Butterworth band-pass filters of ``order`` (parameters.py:1062-1085): with the
    detuning v = (f / fm - fm / f) / (G^(1/2) - G^(-1/2)), which is +-1 at the band's edges
    fm G^(+-1/2), |H| = 1 / sqrt(1 + v^(2 order)): 1 at mid-band, -3 dB at the edges."""
    f = np.asarray(f, float)
    mid = octave_midbands()[:, None]
    edge = OCTAVE_RATIO ** 0.5
    detuning = (f / mid - mid / f) / (edge - 1.0 / edge)
    return 1.0 / np.sqrt(1.0 + detuning ** (2 * order))


def band_filter_gains(n_fft: int, sample_rate: float, order: int = FILTER_ORDER) -> np.ndarray:
    """band_pass on the rFFT grid of n_fft at ``sample_rate`` (8, n_fft // 2 + 1)."""
    return band_pass(np.maximum(np.fft.rfftfreq(n_fft, 1.0 / sample_rate), 1e-9), order)


def fft_size(n: int, sample_rate: float) -> int:
    """The FFT length for filtering n samples with PAD_S of zeros on both sides."""
    return 1 << int(math.ceil(math.log2(n + 2 * int(PAD_S * sample_rate) + 1)))


def zero_phase(x: np.ndarray, gain: np.ndarray, n_fft: int) -> np.ndarray:
    """x (n, ...) through a zero-phase gain on the rFFT grid of n_fft."""
    shape = (-1,) + (1,) * (x.ndim - 1)
    return np.fft.irfft(np.fft.rfft(x, n_fft, axis=0) * gain.reshape(shape), n_fft, axis=0)[: len(x)]


def fixed_dot(x, y) -> float:
    """x . y in eight partial sums, sum L over the i with i mod 8 == L in ascending order
    from +0.0, combined as ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)): an order that
    does not depend on the machine."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    n = len(x)
    m = -(-n // 8) * 8
    p = np.zeros(m + 8)  # a first row of +0.0: each lane starts from it
    p[8:8 + n] = x * y
    s = np.cumsum(p.reshape(-1, 8), axis=0)[-1]  # each lane summed in order
    return float(((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])))


def fixed_dots(x, y) -> tuple[float, float, float]:
    """(x . x, y . y, x . y), each as fixed_dot."""
    return fixed_dot(x, x), fixed_dot(y, y), fixed_dot(x, y)


def _line_window(level: np.ndarray, top: float, bottom: float):
    """The samples a decay line is fitted over: from the first at or below ``top`` to the
    first at or below ``bottom`` (parameters.py:266-296). None where the curve never
    reaches ``bottom``, reaches it only where no energy is left (-inf: the end of the data,
    not a decay), or leaves fewer than three samples between the two."""
    reached = np.flatnonzero(level <= bottom)
    if reached.size == 0 or not np.isfinite(level[reached[0]]):
        return None
    stop = int(reached[0])
    above = np.flatnonzero(level <= top)
    start = int(above[0]) if above.size else 0
    return (start, stop + 1) if stop - start >= 2 else None


def _least_squares(x: np.ndarray, y: np.ndarray):
    """The closed-form straight line through (x, y): (slope, intercept, r^2), the sums of
    squares and products from fixed_dots (a machine-independent order)."""
    n = len(x)
    sum_x, sum_y = float(x.sum()), float(y.sum())
    xx, yy, xy = fixed_dots(x, y)
    var_x = n * xx - sum_x * sum_x
    var_y = n * yy - sum_y * sum_y
    cov = n * xy - sum_x * sum_y
    slope = cov / var_x
    r2 = min(cov * cov / (var_x * var_y), 1.0) if var_y > 0 else 1.0
    return slope, (sum_y - slope * sum_x) / n, r2


def fit_line(level: np.ndarray, t: np.ndarray, top: float, bottom: float):
    """The least-squares line of a decay curve between ``top`` and ``bottom`` dB:
    (slope dB/s, intercept dB, r^2), or None without a window (see _line_window) or where
    the line does not fall."""
    window = _line_window(level, top, bottom)
    if window is None:
        return None
    a, b = window
    line = _least_squares(t[a:b], level[a:b])
    return line if line[0] < 0 else None


def _still_to_come(e: np.ndarray, extra: float = 0.0, closing_edge: bool = False) -> np.ndarray:
    """The energy still to come at the start of each slot of ``e`` (the backward, or
    Schroeder, integral), ``extra`` added after the last slot; with ``closing_edge``, one more
    value at the end of the last slot."""
    backward = np.cumsum(e[::-1])[::-1]
    if closing_edge:
        backward = np.append(backward, 0.0)
    return backward + extra


def _level_db(remaining: np.ndarray, reference: float) -> np.ndarray:
    """10 log10(remaining / reference), -inf where nothing remains."""
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.maximum(remaining / reference, 0.0))


def tail(e: np.ndarray, dt: float) -> tuple[float, float]:
    """The energy after the end of ``e`` (per-slot energies) and its decay rate (1/s)
    (parameters.py:306-333): a line through the data's own backward-integrated curve,
    first between -5 and -15 dB, then twice moved down to end TAIL_HEADROOM_DB above where
    the previous line reaches at the end of the data (at most -15 dB, at least -60 dB) and
    span TAIL_SPAN_DB above that (at most up to -5 dB); the last line, extended to the end,
    gives the energy still to come. (0, inf) where any fit fails."""
    remaining = _still_to_come(e)
    if remaining.size == 0 or remaining[0] <= 0:
        return 0.0, math.inf
    total = float(remaining[0])
    t = np.arange(len(e)) * dt
    end = len(e) * dt
    level = _level_db(remaining, total)
    window = (-5.0, -15.0)
    for refinement in range(3):
        line = fit_line(level, t, *window)
        if line is None:
            return 0.0, math.inf
        if refinement == 2:
            break
        bottom = min(max(line[0] * end + line[1] + TAIL_HEADROOM_DB, -60.0), -15.0)
        window = (min(bottom + TAIL_SPAN_DB, -5.0), bottom)
    slope, intercept = line[0], line[1]
    return float(total * 10.0 ** ((slope * end + intercept) / 10.0)), float(-slope * math.log(10.0) / 10.0)


def decay_curve(energy: np.ndarray, time_step: float) -> np.ndarray:
    """The Schroeder decay in dB at the start of each slot, the estimated tail included,
    0 dB at the start (parameters.py:357-369); -inf throughout without energy."""
    e = np.asarray(energy, float)
    remaining = _still_to_come(e, tail(e, time_step)[0])
    if remaining.size == 0 or remaining[0] <= 0:
        return np.full(len(e), -np.inf)
    return _level_db(remaining, remaining[0])


def decay_level(e: np.ndarray, extra: float) -> tuple[np.ndarray, np.ndarray]:
    """The energy still to come at each slot edge (``extra`` after the last) and its level
    in dB against the curve's own start (parameters.py:442-452)."""
    remaining = _still_to_come(e, extra, closing_edge=True)
    return remaining, _level_db(remaining, remaining[0])


def decay_time(level: np.ndarray, t: np.ndarray, top: float, bottom: float) -> float:
    """-60 dB over the slope of the line fitted between ``top`` and ``bottom``, or NaN."""
    line = fit_line(level, t, top, bottom)
    return math.nan if line is None else -60.0 / line[0]


def too_short(level: np.ndarray, edges: np.ndarray, edt: float) -> bool:
    """Whether the decay stops short of MIN_DECAY_DB (parameters.py:455-461): its curve's
    end (the tail included) still above it, and the early decay's slope not reaching it by
    the last edge either (a NaN EDT never does)."""
    end_above = level[-1] > MIN_DECAY_DB
    slope_reaches = 60.0 * edges[-1] / edt >= -MIN_DECAY_DB
    return bool(end_above and not slope_reaches)


def band_t30(e: np.ndarray, dt: float, offset: float = 0.0) -> float:
    """T30 (s) of one band's slot energies ``e`` from the slot of the onset, which lies
    ``offset`` seconds into it (parameters.py:464-475); NaN where the decay has no energy,
    is too short, or gives no line between -5 and -35 dB."""
    extra = tail(e, dt)[0]
    if float(e.sum()) + extra <= 0:
        return math.nan
    edges = np.arange(len(e) + 1) * dt - offset
    level = decay_level(e, extra)[1]
    if too_short(level, edges, decay_time(level, edges, *DECAY_RANGES["edt"])):
        return math.nan
    return decay_time(level, edges, *DECAY_RANGES["t30"])


def band_strength(e: np.ndarray, dt: float, reference: float = REFERENCE_ENERGY_10M) -> float:
    """The strength G (dB) of one band's slot energies: its energy with the estimated tail,
    against the direct sound's at 10 m in free field."""
    extra, _ = tail(e, dt)
    total = float(e.sum()) + extra
    return 10.0 * math.log10(total / reference) if total > 0 else -math.inf


def band_energy(signal: np.ndarray, rate: float, bands: int, time_step: float, slots: int) -> np.ndarray:
    """Energy per octave band and time slot (bands, slots) of a pressure signal from
    emission, as the ray tracer counts it: the band-filtered pressure's energy over its
    filter's share of the spectrum; what the filters spread before emission counts in the
    first slot."""
    signal = np.asarray(signal, float)
    pad = int(round(PAD_S * rate))
    n_fft = fft_size(len(signal) + pad, rate)
    gains = band_filter_gains(n_fft, rate)[:bands]
    weights = np.full(gains.shape[1], 2.0)
    weights[0] = weights[-1] = 1.0
    time = (np.arange(len(signal) + pad) - pad) / rate
    slot = np.maximum(time / time_step, 0).astype(np.int64)
    keep = slot < slots
    padded = np.concatenate([np.zeros(pad), signal])
    out = np.zeros((bands, slots))
    for b, gain in enumerate(gains):
        pressure = zero_phase(padded, gain, n_fft)
        share = float((gain ** 2 * weights).sum() / n_fft)
        energy = pressure[keep] ** 2 / (ENERGY_TO_IR * share)
        out[b] = np.bincount(slot[keep], energy, minlength=slots)[:slots]
    return out


def octave_results(signal: np.ndarray, rate: float, bands: int = 3, time_step: float = 1e-3) -> dict:
    """The lowest ``bands`` octave bands' strength G (dB) and T30 (s) of one receiver's
    pressure signal from emission, read from the slot of its loudest arrival on."""
    slots = int(math.ceil(len(signal) / rate / time_step))
    e = band_energy(signal, rate, bands, time_step, slots)
    out = {"bands_hz": list(OCTAVE_BANDS_HZ[:bands]), "g_db": [], "t30_s": []}
    for row in e:
        first = int(np.argmax(row))
        out["g_db"].append(band_strength(row, time_step))
        out["t30_s"].append(band_t30(row[first:], time_step))
    return out


def bands_per_octave(bands: int) -> int:
    """1 for the 8 octave bands, 3 for the 24 third-octave bands (materials/bands.py:41-50)."""
    if bands == len(OCTAVE_BANDS_HZ):
        return 1
    if bands == len(THIRD_OCTAVE_BANDS_HZ):
        return 3
    raise ValueError(f"expected {len(OCTAVE_BANDS_HZ)} octave or {len(THIRD_OCTAVE_BANDS_HZ)} "
                     f"third-octave bands, got {bands}")


def crossover_frequencies(bands: int = len(OCTAVE_BANDS_HZ)) -> np.ndarray:
    """The edges between adjacent bands (Hz): the nominal octave centres times sqrt(2), or
    between third octaves 1 kHz times 2^((k + 1/2) / 3) (materials/bands.py:78-84)."""
    if bands_per_octave(bands) == 1:
        return np.array(OCTAVE_BANDS_HZ[:-1], float) * math.sqrt(2.0)
    k = np.arange(len(THIRD_OCTAVE_BANDS_HZ) - 1) - THIRD_OCTAVE_BANDS_HZ.index(1000)
    return 1000.0 * 2.0 ** ((k + 0.5) / 3.0)


def crossover_width(bands: int = len(OCTAVE_BANDS_HZ)) -> float:
    """A crossover's width in octaves: half a band (synthesis.py:47-49)."""
    return CROSSOVER_WIDTH / bands_per_octave(bands)


def band_gains(n_fft: int, sample_rate: float, bands: int = len(OCTAVE_BANDS_HZ)) -> np.ndarray:
    """The synthesis's zero-phase band gains (bands, n_fft // 2 + 1) on the rFFT grid: at
    each edge a raised-cosine step sin^2(pi x / 2) over the crossover's width in log2 f,
    each band the difference of the steps at its edges, so that they sum to one
    (synthesis.py:50-71)."""
    f = np.fft.rfftfreq(int(n_fft), 1.0 / float(sample_rate))
    log_f = np.log2(np.maximum(f, 1e-6))
    width = crossover_width(bands)
    steps = np.array([np.sin(0.5 * np.pi * np.clip((log_f - (edge - width / 2)) / width, 0.0, 1.0)) ** 2
                      for edge in np.log2(crossover_frequencies(bands))])
    gains = np.empty((bands, len(f)))
    gains[0] = 1.0 - steps[0]
    gains[1:-1] = steps[:-1] - steps[1:]
    gains[-1] = steps[-1]
    return gains


def kaiser_sinc(t: np.ndarray, cutoff: float) -> np.ndarray:
    """The resampling low-pass at offsets ``t`` (input samples), ``cutoff`` a share of the
    input's Nyquist frequency: the ideal low-pass's impulse response cutoff sinc(cutoff t)
    under a Kaiser window of shape RESAMPLE_BETA that closes at RESAMPLE_ZEROS of its zero
    crossings (the reference's filter, render.py:105-109)."""
    support = RESAMPLE_ZEROS / cutoff
    inside = np.sqrt(np.clip(1.0 - (t / support) ** 2, 0.0, 1.0))
    return cutoff * np.sinc(cutoff * t) * (np.i0(RESAMPLE_BETA * inside) / np.i0(RESAMPLE_BETA))


def ordered_rows_dot(rows: np.ndarray, taps: np.ndarray) -> np.ndarray:
    """rows (n, taps) times taps, each row's terms added in ascending tap order from the
    first (the reference's products.product, products.py:39-56)."""
    return _tap_major(lambda k: rows[:, k], taps)


def _tap_major(column, taps):
    """sum over k of column(k) * taps[k], for every row at once, added in ascending k from
    the first term: the order of a plain loop over the taps."""
    acc = column(0) * taps[0]
    for k in range(1, len(taps)):
        acc = acc + column(k) * taps[k]
    return acc


def _output_positions(count: int, rate_in: float, rate_out: float):
    """Where each of ``count`` outputs lies among the inputs: (the input sample at or
    below it, its phase, the number of phases). Output k sits at k rate_in / rate_out;
    where that ratio is p / q with q <= RESAMPLE_MAX_PHASES the phase is exact integer
    arithmetic (k p mod q of q); otherwise the position's fraction rounds to one of
    RESAMPLE_MAX_PHASES phases, a round up to the next sample carried into it."""
    ratio = (Fraction(rate_in) / Fraction(rate_out)).limit_denominator(1 << 20)
    k = np.arange(count, dtype=np.int64)
    if ratio.denominator <= RESAMPLE_MAX_PHASES:
        p, q = ratio.numerator, ratio.denominator
        return (k * p) // q, (k * p) % q, q
    position = k * (rate_in / rate_out)
    whole = np.floor(position).astype(np.int64)
    phase = np.round((position - whole) * RESAMPLE_MAX_PHASES).astype(np.int64)
    return whole + phase // RESAMPLE_MAX_PHASES, phase % RESAMPLE_MAX_PHASES, RESAMPLE_MAX_PHASES


def resample(signal, rate_in: float, rate_out: float) -> np.ndarray:
    """``signal`` (samples,) or (samples, channels) at ``rate_out``, ceil(n rate_out /
    rate_in) samples, by band-limited interpolation (the reference's render.resample,
    render.py:112-162): the cutoff just below the lower Nyquist frequency, 0.97 of it;
    a bank of the filter's phases (one row of taps per phase, taps at whole-sample offsets
    round the output's position); every output the dot product of its phase's row with
    the inputs round it (zeros past either end), in ascending tap order. The outputs are
    formed all at once, tap after tap, in CHUNK-sized blocks."""
    x = np.asarray(signal, np.float64)
    if rate_in <= 0 or rate_out <= 0:
        raise ValueError("sample rates must be positive")
    if math.isclose(rate_in, rate_out) or len(x) == 0:
        return x.copy()
    cutoff = 0.97 * min(1.0, rate_out / rate_in)
    reach = int(math.ceil(RESAMPLE_ZEROS / cutoff))
    count = int(math.ceil(len(x) * rate_out / rate_in))
    whole, phase, phases = _output_positions(count, rate_in, rate_out)
    offsets = np.arange(1 - reach, reach + 1)  # taps' inputs, relative to the one at or below
    bank = kaiser_sinc(offsets[None, :] - (np.arange(phases) / phases)[:, None], cutoff)
    columns = x.reshape(len(x), -1)
    # Input i + reach of `padded` is input i: every tap of every output lies inside it.
    padded = np.zeros((len(x) + 2 * reach + 1, columns.shape[1]))
    padded[reach: reach + len(x)] = columns
    out = np.empty((count, columns.shape[1]))
    for a in range(0, count, RESAMPLE_CHUNK):
        b = min(a + RESAMPLE_CHUNK, count)
        rows = bank[phase[a:b]]
        first = whole[a:b] + 1  # padded index of the first tap (offset 1 - reach)
        acc = padded[first] * rows[:, 0, None]
        for k in range(1, len(offsets)):
            acc += padded[first + k] * rows[:, k, None]
        out[a:b] = acc
    return out[:, 0] if x.ndim == 1 else out
