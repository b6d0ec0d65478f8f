"""The wave solver's responses: first-order (W, X, Y, Z) signals at each listener.

The Python side of the reference's `wave_responses` (`engine/wave.py:805-1271`) around the
stepping seam, for a scene of this package's (`scenes.Scene`):
- each listener is read at the 13 points round it (`_probe_points`, 677-682);
- the source is fed the reference's drive (`source_signal`, 281-327) with its lead
  (1118-1128);
- the grid is stepped on the GPU (`HipWaveSolver`);
- each listener's pressure and gradient become W, X, Y, Z (`_first_order`, 757-802), in a
  `WaveResponse` (330-454) with the reference's fields, `fed`, `field` (resampled to a
  render rate), `high_pass` and `energy`.

The reference's model types, band grids from materials, moving air and its CPU stepper are
not here: a scene carries its own grid and admittances, and the stepping is the GPU's.
"""

from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

from acoustic_spaces_amd import analysis

COURANT = 1.0 / math.sqrt(3.0)      # the 7-point grid's
ISOTROPIC_COURANT = math.sqrt(0.75)  # the 19-point grid's
HIGH_PASS_S = 0.12                  # the source's high-pass window (s)
LOW_PASS_TAPS = 33
LOW_PASS_CUTOFF = 0.2               # of the sample rate, on the 7-point grid
PROBES = 13                         # points a listener is read at
SOURCE_CELLS = {7: 0.0, 19: 1.5}    # a listener's points this near the source are left out
LOW_EDGE_HZ = analysis.OCTAVE_BANDS_HZ[0] / (2.0 * math.sqrt(2.0))  # where the wave bands start


def wave_gains(n_fft: int, sample_rate: float, bands: int, total: int = len(analysis.OCTAVE_BANDS_HZ)) -> np.ndarray:
    """Zero-phase gains (bands, n_fft // 2 + 1) of the lowest ``bands`` of ``total`` bands as
    the wave solver covers them: the synthesis's band gains, the lowest one rising from
    LOW_EDGE_HZ over a crossover's width (engine/wave.py:195-206)."""
    gains = analysis.band_gains(n_fft, sample_rate, total)[:bands]
    f = np.fft.rfftfreq(n_fft, 1.0 / sample_rate)
    width = analysis.CROSSOVER_WIDTH
    x = (np.log2(np.maximum(f, 1e-6)) - math.log2(LOW_EDGE_HZ) + width / 2) / width
    gains[0] *= np.sin(0.5 * np.pi * np.clip(x, 0.0, 1.0)) ** 2
    return gains


def cutoff(courant: float) -> float:
    """The source's low-pass cutoff (of the grid's sample rate) on the grid of Courant
    number ``courant``."""
    return LOW_PASS_CUTOFF * courant / COURANT


def source_signal(rate: float, cell: float, courant: float = COURANT, speed: float = 1.0, *,
                  level: bool = True) -> tuple[np.ndarray, int]:
    """What the source adds to the grid each step, and its delay in steps: an impulse,
    low-passed (a Blackman-windowed sinc of LOW_PASS_TAPS taps at ``cutoff(courant)``) and
    high-passed twice (one minus a Hann window of HIGH_PASS_S, so it leaves no air in the
    room), scaled to 4 pi courant^2 / cell so that the free-field pressure at r metres is
    that impulse over r. On the 19-point grid, with ``level``, three taps 2/9, 5/9, 2/9
    (the side taps over speed^2) undo the scheme's gain 9 / (5 + 4 cos wT) at the source,
    with a step more of delay."""
    half = LOW_PASS_TAPS // 2
    k = np.arange(-half, half + 1)
    c = cutoff(courant)
    low = 2 * c * np.sinc(2 * c * k) * np.blackman(LOW_PASS_TAPS)
    low /= low.sum()
    length = max(3, int(round(HIGH_PASS_S * rate)))
    window = np.hanning(length + 2)[1:-1]
    high = -window / window.sum()
    high[0] += 1.0
    signal = np.convolve(np.convolve(low, high), high)
    delay = half
    if level and math.isclose(courant, ISOTROPIC_COURANT, rel_tol=1e-9):
        side = 2.0 / (9.0 * speed ** 2)
        signal = np.convolve(signal, [side, 1.0 - 2.0 * side, side])
        delay += 1
    return 4.0 * math.pi * courant ** 2 / cell * signal, delay


def high_pass(x, rate: float) -> np.ndarray:
    """``x`` (samples, ...) through the source signal's high-pass twice (one minus a Hann
    window of HIGH_PASS_S): what a drive that carries air into a closed room (a Gaussian
    pulse) would give had it been high-passed as source_signal's is, the scheme being
    linear; same length."""
    x = np.asarray(x, float)
    length = max(3, int(round(HIGH_PASS_S * rate)))
    window = np.hanning(length + 2)[1:-1]
    high = -window / window.sum()
    high[0] += 1.0
    h2 = np.convolve(high, high)
    n_fft = 1 << int(math.ceil(math.log2(len(x) + len(h2))))
    shape = (-1,) + (1,) * (x.ndim - 1)
    y = np.fft.irfft(np.fft.rfft(x, n_fft, axis=0) * np.fft.rfft(h2, n_fft).reshape(shape), n_fft, axis=0)
    return y[: len(x)]


def probe_points(listener, cell: float) -> np.ndarray:
    """The listener, the points a cell to either side along each axis (+x, -x, +y, -y,
    +z, -z), then two cells to either side in the same order: PROBES points (13, 3)."""
    listener = np.asarray(listener, float)
    steps = np.repeat(np.eye(3), 2, axis=0) * np.tile([1, -1], 3)[:, None]
    return listener[None, :] + cell * np.vstack([np.zeros(3), steps, 2 * steps])


def off_the_source(found, probes, source, cell: float, stencil: int) -> np.ndarray:
    """``found`` without the points among the source's own cells: nearer the source than
    SOURCE_CELLS cells and nearer it than their listener. An axis that loses both its
    points a cell away keeps its four."""
    limit = (SOURCE_CELLS[stencil] - 1e-9) * cell
    away = np.linalg.norm(np.asarray(probes) - np.asarray(source), axis=1).reshape(-1, PROBES)
    near = (away < limit) & (away < away[:, :1] - 1e-9 * cell)
    before = np.asarray(found, dtype=bool).reshape(-1, PROBES)
    after = before & ~near
    for axis in range(3):
        points = [1 + 2 * axis, 2 + 2 * axis, 7 + 2 * axis, 8 + 2 * axis]
        empty = ~(after[:, points[0]] | after[:, points[1]])
        after[np.ix_(empty, points)] = before[np.ix_(empty, points)]
    return after.reshape(-1)


def first_order(probes: np.ndarray, found, name: str = "", courant: float = COURANT,
                cut: float = LOW_PASS_CUTOFF) -> np.ndarray:
    """W, X, Y, Z (samples, 4) from the pressure at the listener and round it (the 13
    rows of ``probes``, at probe_points). X, Y, Z are c times the time integral of the
    pressure gradient, X = -rho c u_x (AmbiX: towards the sound):
    - the gradient is a fourth-order central difference over two cells either side,
      second-order where a point two cells away is not ``found``, one-sided where a
      point a cell away is not;
    - the integral is the trapezoidal rule, its gain (wT/2) / tan(wT/2) divided out up
      to ``cut`` of the sample rate.
    ``courant`` is the Courant number of the speed of sound at the listener."""
    if not found[0]:
        raise ValueError(f"listener {name!r} is not in the air of the wave grid")
    out = np.zeros((probes.shape[1], 4))
    out[:, 0] = probes[0]
    for axis in range(3):
        plus, minus = 1 + 2 * axis, 2 + 2 * axis
        if found[plus] and found[minus]:
            difference = 0.5 * (probes[plus] - probes[minus])
            if found[plus + 6] and found[minus + 6]:
                difference = difference * (4 / 3) - (probes[plus + 6] - probes[minus + 6]) / 12
        elif found[plus]:
            difference = probes[plus] - probes[0]
        elif found[minus]:
            difference = probes[0] - probes[minus]
        else:
            continue
        out[:, 1 + axis] = courant * (np.cumsum(difference) - 0.5 * difference)
    n_fft = 1 << int(math.ceil(math.log2(2 * len(out) + 1)))
    turn = np.pi * np.fft.rfftfreq(n_fft)  # wT / 2
    gain = np.ones(len(turn))
    below = (turn > 0) & (turn <= np.pi * cut)
    gain[below] = np.tan(turn[below]) / turn[below]
    spectrum = np.fft.rfft(out[:, 1:], n_fft, axis=0) * gain[:, None]
    out[:, 1:] = np.fft.irfft(spectrum, n_fft, axis=0)[: len(out)]
    return out


@dataclass
class WaveResponse:
    """The lowest octave bands of one response, from the wave solver: ``signals`` (samples,
    4) holds W, X, Y, Z at ``rate`` from ``lead_s`` before emission on."""

    rate: float
    signals: np.ndarray
    bands: int
    cell: float
    total_bands: int = len(analysis.OCTAVE_BANDS_HZ)
    lead_s: float = 0.0
    band_gain: np.ndarray | None = None
    delay_s: float = 0.0
    _cache: dict = field(default_factory=dict, init=False, repr=False, compare=False)

    def fed(self, gain, delay_s: float) -> "WaveResponse":
        """These bands with the energy ``gain`` of each and ``delay_s`` later."""
        gain = np.asarray(gain, float)
        total = gain if self.band_gain is None else self.band_gain * gain
        return dataclasses.replace(self, band_gain=total, delay_s=self.delay_s + delay_s)

    def _amplitudes(self) -> np.ndarray:
        return np.ones(self.bands) if self.band_gain is None else np.sqrt(self.band_gain)

    def _resampled(self, sample_rate: float, n: int) -> tuple:
        """The run at ``sample_rate`` (at least before + n samples, 4 channels), still
        broadband, and ``before``: how many of its samples precede emission. Its start is
        rounded to a sample once, and an impulse's height scales with the sample period
        (engine/wave.py:366-387)."""
        key = (float(sample_rate), int(n))
        if key not in self._cache:
            start = int(round((self.delay_s - self.lead_s) * sample_rate))
            before = max(0, -start)
            x = analysis.resample(self.signals, self.rate, sample_rate)
            out = np.zeros((before + max(n, start + len(x)), 4))
            out[before + start: before + start + len(x)] = x * (self.rate / sample_rate)
            self._cache = {key: (out, before)}
        return self._cache[key]

    def field(self, sample_rate: float, n: int) -> np.ndarray:
        """The first-order sound field (n, 4) of the wave bands together from emission, at
        ``sample_rate``: the resampled run through the bands' summed zero-phase gains (each
        band's times its feed's amplitude), whose spread before emission is left out
        (engine/wave.py:389-399)."""
        run, before = self._resampled(sample_rate, n)
        n_fft = analysis.fft_size(len(run), sample_rate)
        gains = wave_gains(n_fft, sample_rate, self.bands, self.total_bands)
        gain = (gains * self._amplitudes()[:, None]).sum(axis=0)
        spectrum = np.fft.rfft(run, n_fft, axis=0) * gain[:, None]
        return np.fft.irfft(spectrum, n_fft, axis=0)[before: before + n]

    def high_pass(self, signals, sample_rate: float, late: bool = False) -> np.ndarray:
        """``signals`` (n, channels) without what the wave bands cover, which take them at a
        gain g: at 1 - g (early paths, adding in pressure) or, with ``late``, at
        sqrt(1 - g^2) (a tail, adding in energy) (engine/wave.py:401-409)."""
        signals = np.asarray(signals, float)
        n_fft = analysis.fft_size(len(signals), sample_rate)
        g = wave_gains(n_fft, sample_rate, self.bands, self.total_bands).sum(axis=0)
        return analysis.zero_phase(signals, np.sqrt(1.0 - g ** 2) if late else 1.0 - g, n_fft)

    def energy(self, time_step: float, slots: int, axis=None, exponent: int = 2) -> np.ndarray:
        """Energy per band and time slot (bands, slots), as the ray tracer counts it: each
        band's filtered pressure's energy over its filter's share of the spectrum. With
        ``axis`` (a unit vector), a figure of eight facing it, from X, Y, Z; with
        ``exponent`` 1, the magnitude of its product with the pressure."""
        out = np.zeros((self.bands, slots))
        pad = int(round(analysis.PAD_S * self.rate))
        n_fft = analysis.fft_size(len(self.signals) + pad, self.rate)
        gains = analysis.band_filter_gains(n_fft, self.rate)[: self.bands]
        weights = np.full(gains.shape[1], 2.0)
        weights[0] = weights[-1] = 1.0
        lead = int(round(self.lead_s * self.rate))
        time = (np.arange(len(self.signals) + pad) - pad - lead) / self.rate + self.delay_s
        slot = np.maximum(time / time_step, 0).astype(np.int64)
        amplitude = self._amplitudes()
        keep = slot < slots
        padded = np.concatenate([np.zeros(pad), self.signals[:, 0]])
        if axis is not None:
            figure8 = np.concatenate([np.zeros(pad), self.signals[:, 1:] @ np.asarray(axis, float)])
        for b, gain in enumerate(gains):
            pressure = analysis.zero_phase(padded, gain, n_fft)
            share = float((gain ** 2 * weights).sum() / n_fft)
            power = pressure[keep] ** 2
            if axis is not None:
                side = analysis.zero_phase(figure8, gain, n_fft)[keep]
                power = side ** 2 if exponent == 2 else np.abs(side * pressure[keep])
            energy = power * amplitude[b] ** 2 / (analysis.ENERGY_TO_IR * share)
            out[b] = np.bincount(slot[keep], energy, minlength=slots)[:slots]
        return out


def wave_responses(scene, listeners, names, *, length_s: float, bands: int = 3, precision: int = 32,
                   say=None, label: str = "") -> tuple[list, dict]:
    """The responses from ``scene``'s source to each of ``listeners`` (L, 3) on the scene's
    own grid (its cell and rate), ``length_s`` seconds from emission, and run details:
    the scene's set-up with 13 points a listener, the reference's drive with its lead, the
    GPU stepper, then each listener's W, X, Y, Z. Raises ValueError for a listener out of
    the source's air."""
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    listeners = np.asarray(listeners, float).reshape(-1, 3)
    h = scene.grid.cell
    courant = COURANT if scene.stencil == 7 else ISOTROPIC_COURANT
    points = np.vstack([probe_points(p, h) for p in listeners])
    run = dataclasses.replace(scene, probes=points)
    solver = WaveSolver.from_scene(run, probes_per_listener=PROBES)
    signal, delay = source_signal(scene.rate, h, courant)
    samples = int(math.ceil(length_s * scene.rate))
    lead = delay
    steps = lead + samples
    drive = np.zeros(steps, np.float32)
    drive[lead - delay: lead - delay + len(signal)] = signal[: steps - lead + delay]
    found = np.array([solver.probe_found(p) and solver.probe_in_sight(p) for p in range(len(points))])
    found = off_the_source(found, points, scene.source, h, scene.stencil).tolist()
    missing = [n for i, n in enumerate(names) if not found[PROBES * i]]
    if missing:
        raise ValueError(f"{label or 'wave'}: listeners out of the source's air: {', '.join(missing)}")
    stepper = HipWaveSolver(solver, precision=precision)
    try:
        if say is not None:
            say(f"{label}: wave solver, {steps} steps of {solver.air_cells:,} cells")
        recorded = stepper.run(drive)
    finally:
        stepper.close()
    if not np.all(np.isfinite(recorded)):
        raise ValueError(f"{label or 'wave'}: the wave solver's run did not stay finite")
    info = {"wave_cells": float(solver.air_cells), "wave_cell_m": h, "wave_steps": float(steps),
            "wave_stencil": float(scene.stencil), "wave_gpu": 1.0}
    if solver.open:
        info["wave_open"] = 1.0
    responses = []
    for i, name in enumerate(names):
        part = slice(PROBES * i, PROBES * (i + 1))
        sig = first_order(recorded[part], found[part], name, courant, cutoff(courant))
        responses.append(WaveResponse(scene.rate, sig, bands, h, lead_s=lead / scene.rate))
    return responses, info
