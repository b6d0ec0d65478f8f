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
from numpy.lib.stride_tricks import sliding_window_view

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
    Butterworth band-pass filters of ``order``: 1 at mid-band, -3 dB at fm G^(+-1/2),
    |H| = 1 / sqrt(1 + ((f / fm - fm / f) / (G^(1/2) - G^(-1/2)))^(2 order))."""
    f = np.asarray(f, float)
    fm = octave_midbands()[:, None]
    half = OCTAVE_RATIO ** 0.5
    width = half - 1.0 / half
    return 1.0 / np.sqrt(1.0 + ((f / fm - fm / f) / width) ** (2 * order))


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


def fit_line(level: np.ndarray, t: np.ndarray, top: float, bottom: float):
    """Least-squares line through ``level`` from where it first falls to ``top`` to where it
    first falls to ``bottom``: (slope dB/s, intercept dB, r^2), or None when the curve does
    not fall that far, or reaches ``bottom`` only where no energy is left (-inf), or does
    not fall."""
    below = level <= bottom
    if not below.any():
        return None
    stop = int(np.argmax(below))
    if not np.isfinite(level[stop]):
        return None
    start = int(np.argmax(level <= top))
    if stop - start < 2:
        return None
    x, y = t[start:stop + 1], level[start:stop + 1]
    n = len(x)
    sx, sy = float(x.sum()), float(y.sum())
    xx, yy, xy = fixed_dots(x, y)
    sxx, syy, sxy = n * xx - sx * sx, n * yy - sy * sy, n * xy - sx * sy
    slope = sxy / sxx
    if slope >= 0:
        return None
    intercept = (sy - slope * sx) / n
    r2 = min(sxy * sxy / (sxx * syy), 1.0) if syy > 0 else 1.0
    return slope, intercept, r2


def tail(e: np.ndarray, dt: float) -> tuple[float, float]:
    """The energy after the end of ``e`` (per-slot energies) and its decay rate (1/s): a
    line through the data's own backward-integrated curve, fitted low on the decay but
    TAIL_HEADROOM_DB above where it would reach by the end, extended to the end."""
    remaining = np.cumsum(e[::-1])[::-1]
    if len(remaining) == 0 or remaining[0] <= 0:
        return 0.0, math.inf
    total = float(remaining[0])
    t = np.arange(len(e)) * dt
    end = len(e) * dt
    with np.errstate(divide="ignore"):
        level = 10.0 * np.log10(np.maximum(remaining / total, 0.0))
    line = fit_line(level, t, -5.0, -15.0)
    for _ in range(2):
        if line is None:
            return 0.0, math.inf
        bottom = min(max(line[0] * end + line[1] + TAIL_HEADROOM_DB, -60.0), -15.0)
        line = fit_line(level, t, min(bottom + TAIL_SPAN_DB, -5.0), bottom)
    if line is None:
        return 0.0, math.inf
    slope, intercept = line[:2]
    return float(total * 10.0 ** ((slope * end + intercept) / 10.0)), float(-slope * math.log(10.0) / 10.0)


def decay_curve(energy: np.ndarray, time_step: float) -> np.ndarray:
    """The backward-integrated (Schroeder) decay in dB at the start of each slot, the
    estimated tail included; 0 dB at the start."""
    e = np.asarray(energy, float)
    extra, _ = tail(e, time_step)
    remaining = np.cumsum(e[::-1])[::-1] + extra
    if len(remaining) == 0 or remaining[0] <= 0:
        return np.full(len(e), -np.inf)
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.maximum(remaining, 0.0) / remaining[0])


def decay_level(e: np.ndarray, extra: float) -> tuple[np.ndarray, np.ndarray]:
    """The energy still to come at each slot edge (``extra`` after the last) and its level
    in dB against the curve's own start; -inf where none is left."""
    remaining = np.concatenate([np.cumsum(e[::-1])[::-1], [0.0]]) + extra
    with np.errstate(divide="ignore"):
        level = 10.0 * np.log10(np.maximum(remaining / remaining[0], 0.0))
    return remaining, level


def decay_time(level: np.ndarray, t: np.ndarray, top: float, bottom: float) -> float:
    """-60 dB over the slope of the line fitted between ``top`` and ``bottom``, or NaN."""
    fit = fit_line(level, t, top, bottom)
    return math.nan if fit is None else -60.0 / fit[0]


def too_short(level: np.ndarray, edges: np.ndarray, edt: float) -> bool:
    """Whether the response stops before its decay has fallen MIN_DECAY_DB, at its end or
    along the early decay's slope."""
    return bool(level[-1] > MIN_DECAY_DB and not 60.0 * edges[-1] / edt >= -MIN_DECAY_DB)


def band_t30(e: np.ndarray, dt: float, offset: float = 0.0) -> float:
    """T30 (s) of one band's slot energies ``e``, starting at the slot of the onset, which
    lies ``offset`` seconds into it; NaN where it cannot be read."""
    edges = np.arange(len(e) + 1) * dt - offset
    extra, _ = tail(e, dt)
    if float(e.sum()) + extra <= 0:
        return math.nan
    _, level = decay_level(e, extra)
    edt = decay_time(level, edges, *DECAY_RANGES["edt"])
    if too_short(level, edges, edt):
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
    """Low-pass taps at offsets ``t`` (input samples), ``cutoff`` a share of the input's
    Nyquist frequency: sinc times a Kaiser window reaching RESAMPLE_ZEROS zero crossings
    (render.py:105-109)."""
    half = RESAMPLE_ZEROS / cutoff
    window = np.i0(RESAMPLE_BETA * np.sqrt(np.clip(1.0 - (t / half) ** 2, 0.0, 1.0)))
    return cutoff * np.sinc(cutoff * t) * window / np.i0(RESAMPLE_BETA)


def ordered_rows_dot(rows: np.ndarray, taps: np.ndarray) -> np.ndarray:
    """rows (n, taps) times taps, each row's terms added in ascending tap order from the
    first (the reference's products.product, products.py:39-56)."""
    acc = rows[:, 0] * taps[0]
    for k in range(1, len(taps)):
        acc = acc + rows[:, k] * taps[k]
    return acc


def resample(signal, rate_in: float, rate_out: float) -> np.ndarray:
    """``signal`` (samples,) or (samples, channels) at ``rate_out``, ceil(n rate_out /
    rate_in) samples: polyphase windowed-sinc interpolation, the cutoff just below the
    lower Nyquist frequency (render.py:112-162). Output k sits at input position
    k rate_in / rate_out; where that ratio is p / q with q <= RESAMPLE_MAX_PHASES, each of
    the q phases is one ordered product over the outputs that share it; otherwise the
    position rounds to one of RESAMPLE_MAX_PHASES phases."""
    x = np.asarray(signal, np.float64)
    if rate_in <= 0 or rate_out <= 0:
        raise ValueError("sample rates must be positive")
    if math.isclose(rate_in, rate_out) or len(x) == 0:
        return x.copy()
    ratio = (Fraction(rate_in) / Fraction(rate_out)).limit_denominator(1 << 20)
    exact = ratio.denominator <= RESAMPLE_MAX_PHASES
    step, phases = (ratio.numerator, ratio.denominator) if exact else (0, RESAMPLE_MAX_PHASES)
    cutoff = 0.97 * min(1.0, rate_out / rate_in)
    reach = int(math.ceil(RESAMPLE_ZEROS / cutoff))
    offsets = np.arange(-reach + 1, reach + 1)  # taps relative to the position's floor
    table = kaiser_sinc(offsets[None, :] - (np.arange(phases) / phases)[:, None], cutoff)
    columns = x.reshape(len(x), -1)
    count = int(math.ceil(len(x) * rate_out / rate_in))
    out = np.zeros((count, columns.shape[1]))
    for c in range(columns.shape[1]):
        padded = np.concatenate([np.zeros(reach), columns[:, c], np.zeros(reach + 1)])
        window = sliding_window_view(padded, len(offsets))  # row r: the taps round input r - 1
        if exact:
            inverse = pow(step, -1, phases)
            for p in range(phases):
                k0 = p * inverse % phases  # the first output of phase p
                if k0 >= count:
                    continue
                n = len(range(k0, count, phases))
                first = (k0 * step - p) // phases + 1
                out[k0::phases, c] = ordered_rows_dot(window[first: first + (n - 1) * step + 1: step], table[p])
            continue
        for start in range(0, count, RESAMPLE_CHUNK):
            k = np.arange(start, min(start + RESAMPLE_CHUNK, count))
            position = k * (rate_in / rate_out)
            whole = np.floor(position).astype(np.int64)
            phase = np.round((position - whole) * phases).astype(np.int64)
            whole, phase = whole + phase // phases, phase % phases
            out[k, c] = np.einsum("kt,kt->k", table[phase], window[whole + 1])
    return out[:, 0] if x.ndim == 1 else out

