"""The wave solver's responses: first-order (W, X, Y, Z) signals at each listener.

The Python side of the reference's `wave_responses` (`engine/wave.py:805-1271`) around the
stepping seam, for a scene of this package's (`scenes.Scene`):
- each listener is read at the 13 points round it (`_probe_points`, 677-682);
- the source is fed the reference's drive (`source_signal`, 281-327) with its lead
  (1118-1128);
- the grid is stepped on the GPU (`HipWaveSolver`);
- each listener's pressure and gradient become W, X, Y, Z (`_first_order`, 757-802), in a
  `WaveResponse` (330-454) with the reference's fields, `fed`, `field` (resampled to a
  render rate), `high_pass` and `energy`;
and the entry point itself, `wave_responses`, from a model read as the reference's
TraceModel: band grids from its materials, the admittance fits, the weather's speeds of
sound, moving air and the watch over it. The stepping is the GPU's: the reference's CPU
and WebGPU steppers are not here.
"""

from __future__ import annotations

import dataclasses
import math
import time
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
    number ``courant``: LOW_PASS_CUTOFF on the 7-point grid, in proportion on others."""
    return LOW_PASS_CUTOFF * courant / COURANT


def _low_pass_taps(courant: float) -> np.ndarray:
    """LOW_PASS_TAPS taps of a Blackman-windowed sinc at ``cutoff(courant)``, unit gain at
    DC (engine/wave.py:281-327)."""
    c = cutoff(courant)
    k = np.arange(LOW_PASS_TAPS) - LOW_PASS_TAPS // 2
    taps = np.blackman(LOW_PASS_TAPS) * (2.0 * c) * np.sinc(2.0 * c * k)
    return taps / taps.sum()


def _hann_high_pass_taps(rate: float) -> np.ndarray:
    """An impulse less a unit-area Hann window of HIGH_PASS_S (at least 3 samples): it
    passes no DC (engine/wave.py:281-327)."""
    length = max(3, int(round(HIGH_PASS_S * rate)))
    hann = np.hanning(length + 2)[1:-1]
    taps = hann * (-1.0 / hann.sum())
    taps[0] += 1.0
    return taps


def _isotropic_level_taps(speed: float) -> np.ndarray:
    """Three taps (2/9, 5/9, 2/9) with the side taps over speed^2: their gain
    (5 + 4 cos wT) / 9 undoes the 19-point scheme's at the source."""
    side = 2.0 / (9.0 * speed ** 2)
    return np.array([side, 1.0 - 2.0 * side, side])


def source_signal(rate: float, cell: float, courant: float = COURANT, speed: float = 1.0, *,
                  level: bool = True) -> tuple[np.ndarray, int]:
    """What the source adds to the grid each step, and its delay in steps (the
    reference's drive, engine/wave.py:281-327): an impulse low-passed by _low_pass_taps and
    high-passed twice by _hann_high_pass_taps (twice, so that it leaves no air in a closed
    room), scaled by 4 pi courant^2 / cell, so that the free-field pressure r metres away is
    that impulse over r. On the 19-point grid, with ``level``, also through
    _isotropic_level_taps, a step more of delay."""
    high = _hann_high_pass_taps(rate)
    chain = [_low_pass_taps(courant), high, high]
    delay = LOW_PASS_TAPS // 2
    if level and math.isclose(courant, ISOTROPIC_COURANT, rel_tol=1e-9):
        chain.append(_isotropic_level_taps(speed))
        delay += 1
    signal = chain[0]
    for taps in chain[1:]:
        signal = np.convolve(signal, taps)
    return (4.0 * math.pi * courant ** 2 / cell) * signal, delay


def high_pass(x, rate: float) -> np.ndarray:
    """``x`` (samples, ...) through the source signal's high-pass twice, same length: by
    the scheme's linearity, what a drive that carries air into a closed room (a Gaussian
    pulse) would give had it been high-passed as source_signal's is."""
    x = np.asarray(x, float)
    twice = np.convolve(*([_hann_high_pass_taps(rate)] * 2))
    n_fft = 1 << int(math.ceil(math.log2(len(x) + len(twice))))
    response = np.fft.rfft(twice, n_fft).reshape((-1,) + (1,) * (x.ndim - 1))
    return np.fft.irfft(np.fft.rfft(x, n_fft, axis=0) * response, n_fft, axis=0)[: len(x)]


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


def _cell_difference(rows: np.ndarray, found, axis: int):
    """The change of pressure over one cell along ``axis`` at the listener, from the 13
    rows (probe_points' order), or None: the fourth-order central difference over two
    cells either side; the second-order one where a point two cells away is missing; the
    one-sided one where a point a cell away is (engine/wave.py:757-802)."""
    up, down = 1 + 2 * axis, 2 + 2 * axis
    if found[up] and found[down]:
        near = 0.5 * (rows[up] - rows[down])
        if not (found[up + 6] and found[down + 6]):
            return near
        far = rows[up + 6] - rows[down + 6]
        return near * (4 / 3) - far / 12
    if found[up]:
        return rows[up] - rows[0]
    if found[down]:
        return rows[0] - rows[down]
    return None


def _trapezoid_integral(d: np.ndarray) -> np.ndarray:
    """The running trapezoidal integral of d (one sample = 1), from a zero before the
    first sample: I[n] = I[n - 1] + (d[n - 1] + d[n]) / 2."""
    halves = 0.5 * d
    return np.cumsum(halves + np.concatenate(([0.0], halves[:-1])))


def _undo_trapezoid_gain(x: np.ndarray, cut: float) -> np.ndarray:
    """``x`` (samples, channels) with the trapezoidal rule's gain (wT/2) / tan(wT/2)
    divided out at frequencies up to ``cut`` of the sample rate (zero-padded FFT)."""
    n_fft = 1 << int(math.ceil(math.log2(2 * len(x) + 1)))
    half_turn = np.pi * np.fft.rfftfreq(n_fft)
    correction = np.ones(len(half_turn))
    inside = (half_turn > 0) & (half_turn <= np.pi * cut)
    correction[inside] = np.tan(half_turn[inside]) / half_turn[inside]
    return np.fft.irfft(np.fft.rfft(x, n_fft, axis=0) * correction[:, None], n_fft, axis=0)[: len(x)]


def first_order(probes: np.ndarray, found, name: str = "", courant: float = COURANT,
                cut: float = LOW_PASS_CUTOFF) -> np.ndarray:
    """W, X, Y, Z (samples, 4) from the pressure at a listener's 13 points (the rows of
    ``probes``, in probe_points' order; the reference's _first_order, engine/wave.py:
    757-802). W is the pressure at the listener. X, Y, Z are c times the time integral of
    the pressure gradient, X = -rho c u_x (AmbiX: towards the sound): per axis the cell
    difference (_cell_difference) integrated by the trapezoidal rule, times the Courant
    number ``courant`` of the speed of sound at the listener (c dt / h), then the rule's
    gain undone up to ``cut`` of the sample rate. An axis with no point a cell away either
    side reads 0."""
    if not found[0]:
        raise ValueError(f"listener {name!r} is not in the air of the wave grid")
    out = np.zeros((probes.shape[1], 4))
    out[:, 0] = probes[0]
    for axis in range(3):
        d = _cell_difference(probes, found, axis)
        if d is not None:
            out[:, 1 + axis] = courant * _trapezoid_integral(d)
    out[:, 1:] = _undo_trapezoid_gain(out[:, 1:], cut)
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

    def _timeline(self, sample_rate: float, n: int) -> tuple:
        """(run, before): the run resampled to ``sample_rate`` (4 channels, still broadband)
        and laid on a timeline that starts ``before`` samples ahead of emission and holds
        at least before + n samples; the run's start rounds to a sample once, and an
        impulse's height scales with the sample period (engine/wave.py:366-387). The last
        timeline asked for is kept."""
        key = (float(sample_rate), int(n))
        if key not in self._cache:
            at = int(round((self.delay_s - self.lead_s) * sample_rate))  # the run's start, from emission
            before = max(0, -at)
            x = analysis.resample(self.signals, self.rate, sample_rate) * (self.rate / sample_rate)
            run = np.zeros((before + max(n, at + len(x)), 4))
            run[before + at: before + at + len(x)] = x
            self._cache = {key: (run, before)}
        return self._cache[key]

    def _covered(self, n_fft: int, sample_rate: float) -> np.ndarray:
        """The wave bands' summed zero-phase gain on the rFFT grid, each band's at its
        feed's amplitude."""
        return self._amplitudes() @ wave_gains(n_fft, sample_rate, self.bands, self.total_bands)

    def field(self, sample_rate: float, n: int) -> np.ndarray:
        """The first-order sound field (n, 4) of the wave bands together, from emission, at
        ``sample_rate`` (engine/wave.py:389-399): the timeline through _covered, what the
        filter spreads ahead of emission dropped."""
        run, before = self._timeline(sample_rate, n)
        n_fft = analysis.fft_size(len(run), sample_rate)
        return analysis.zero_phase(run, self._covered(n_fft, sample_rate), n_fft)[before: before + n]

    def high_pass(self, signals, sample_rate: float, late: bool = False) -> np.ndarray:
        """``signals`` (n, channels) without what the wave bands cover at their gain g:
        through 1 - g (early paths, which add in pressure) or, with ``late``, through
        sqrt(1 - g^2) (a tail, which adds in energy) (engine/wave.py:401-409)."""
        signals = np.asarray(signals, float)
        n_fft = analysis.fft_size(len(signals), sample_rate)
        g = wave_gains(n_fft, sample_rate, self.bands, self.total_bands).sum(axis=0)
        rest = np.sqrt(1.0 - g * g) if late else 1.0 - g
        return analysis.zero_phase(signals, rest, n_fft)

    def energy(self, time_step: float, slots: int, axis=None, exponent: int = 2) -> np.ndarray:
        """Energy per band and time slot (bands, slots), as the ray tracer counts it
        (engine/wave.py:411-454): every band's zero-phase filtered pressure squared, over
        ENERGY_TO_IR times the filter's share of the spectrum, times the band's feed, into
        the slot of its time from emission (what the filters spread before emission into
        the first). With ``axis`` (a unit vector), a figure of eight facing it (X, Y, Z
        along it) squared, or with ``exponent`` 1 the magnitude of its product with the
        pressure."""
        pad = int(round(analysis.PAD_S * self.rate))
        n = len(self.signals) + pad
        n_fft = analysis.fft_size(n, self.rate)
        gains = analysis.band_filter_gains(n_fft, self.rate)[: self.bands]
        one_sided = np.full(gains.shape[1], 2.0)  # bins counted twice on the one-sided grid
        one_sided[0] = one_sided[-1] = 1.0
        per_energy = self._amplitudes() ** 2 / (analysis.ENERGY_TO_IR * (gains ** 2 @ one_sided) / n_fft)

        def banded(series):  # (n, bands): the series after `pad` zeros, through each band
            spectrum = np.fft.rfft(np.concatenate([np.zeros(pad), series]), n_fft)
            return np.fft.irfft(spectrum[:, None] * gains.T, n_fft, axis=0)[:n]

        lead = int(round(self.lead_s * self.rate))
        t = (np.arange(n) - pad - lead) / self.rate + self.delay_s
        slot = np.maximum(t / time_step, 0).astype(np.int64)
        keep = slot < slots
        pressure = banded(self.signals[:, 0])
        if axis is None:
            power = pressure * pressure
        else:
            side = banded(self.signals[:, 1:] @ np.asarray(axis, float))
            power = side * side if exponent == 2 else np.abs(side * pressure)
        power = power[keep] * per_energy
        out = np.zeros((self.bands, slots))
        for b in range(self.bands):
            out[b] = np.bincount(slot[keep], power[:, b], minlength=slots)[:slots]
        return out


def scene_responses(scene, listeners, names, *, length_s: float, bands: int = 3, precision: int = 32,
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
    solver = WaveSolver.from_scene(run, probes_per_listener=PROBES, on_device=True)
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


# --- the reference's entry point -----------------------------------------------------------

#: The stencils' points and Courant numbers, by scheme (engine/wave.py:116-130).
STENCILS = {"standard": (7, COURANT), "isotropic": (19, ISOTROPIC_COURANT)}
#: Link weights of an axial and an edge link, by the stencil's points.
LINK_WEIGHTS = {7: (1.0 / 3.0, 0.0), 19: (0.25, 0.125)}
MAX_CELLS = 40_000_000                      # grids of more cells are refused
WAVE_BANDS_HZ = analysis.OCTAVE_BANDS_HZ[:4]  # the highest band a run may reach
LAYER_CELLS = 8                              # the matched layer round an open model
# The configuration's refusals of fitted walls (scene/model.py:3679-3686).
FITTED_WALLS_STENCIL = ("simulation.wave.walls: fitted works on the standard (7-point) stencil only: set "
                        "simulation.wave.scheme to standard or auto, or simulation.wave.walls to cell_faces")
FITTED_WALLS_OUTDOORS = ("simulation.wave.walls: fitted is not available outdoors: set simulation.wave.walls to "
                         "cell_faces")
LAYER_DAMPING = 0.5
LAYER_ROOM = 0.15   # room past the points, a share of their farthest spread across each axis
LONG_RUN = 2e9      # cell updates beyond which a run reports its tenths
#: Moving air is watched for sound that grows (engine/wave.py:177-187): the pressure squared
#: over the grid's air is averaged over pieces of the run at least WATCH_S long (s); from the
#: end of the source's pulse on, a piece more than WATCH_RISE times the lowest before it, and
#: more than WATCH_FLOOR of the largest, stops the run.
WATCH_S = 0.25
WATCH_RISE = 2.0
WATCH_FLOOR = 1e-6


def band_top(band: int, total: int = len(analysis.OCTAVE_BANDS_HZ)) -> float:
    """The highest frequency band ``band`` covers (Hz): its upper crossover's top."""
    edge = analysis.crossover_frequencies(total)[band]
    return float(edge * 2.0 ** (analysis.crossover_width(total) / 2))


@dataclass(frozen=True)
class BandGrid:
    """A band's grid: its cell (m) and its sample rate (Hz)."""

    band: int
    cell: float
    rate: float


def band_grid(band: int, sample_rate: float, speed_of_sound: float, points_per_wavelength: float = 8.0,
              total: int = len(analysis.OCTAVE_BANDS_HZ), courant: float = COURANT) -> BandGrid:
    """The coarsest grid with ``points_per_wavelength`` cells a wavelength at the top of
    ``band`` whose step is a whole number of output samples, or one over a whole number
    (engine/wave.py:253-272): the rate speed / (courant cell) moved to the nearest such
    rate at or above it, the cell then taken from the rate."""
    cell = speed_of_sound / (points_per_wavelength * band_top(band, total))
    rate = speed_of_sound / (courant * cell)
    ratio = sample_rate / rate
    rate = sample_rate / math.floor(ratio) if ratio >= 1.0 else sample_rate * math.ceil(1.0 / ratio)
    return BandGrid(band, speed_of_sound / (courant * rate), rate)


def slow_by(stencil: int, per_wavelength: float, ratio: float = 1.0) -> float:
    """How much slower than sound (a fraction) a wave of ``per_wavelength`` cells a
    wavelength runs along the grid's axes: from the stencil's dispersion relation
    2 - 2 cos(w T) = ratio^2 sum_k g_k (1 - cos(k . d_k)) (engine/wave.py:216-245)."""
    axial, edge = LINK_WEIGHTS[stencil]
    courant = STENCILS["standard" if stencil == 7 else "isotropic"][1]
    theta = 2.0 * math.pi / per_wavelength
    c = math.cos(theta)
    total = 2.0 * axial * (1.0 - c) + 4.0 * edge * (2.0 * (1.0 - c))  # two edge pairs hold the axis
    turned = math.acos(1.0 - ratio ** 2 * total / 2.0)
    return 1.0 - turned / (ratio * courant * theta)


def _spread_across(source, listeners) -> np.ndarray:
    """(3,): the farthest a listener lies from the source in the plane across each axis."""
    d = np.asarray(listeners, float) - np.asarray(source, float)
    if len(d) == 0:
        return np.zeros(3)
    square = (d * d).sum(axis=1)[:, None] - d * d
    return np.sqrt(square.max(axis=0))


def _band_centres(bands_kind: str) -> list:
    return list(analysis.OCTAVE_BANDS_HZ) if bands_kind == "octave" else list(analysis.THIRD_OCTAVE_BANDS_HZ)


def wave_responses(model, source, listeners, names, *, max_band_hz: int, length_s: float, sample_rate: float,
                   speed_of_sound: float, points_per_wavelength: float = 8.0, threads: int = 0,
                   device: str = "gpu", scheme: str = "auto", say=None, label: str = "",
                   wind: str = "effective", walls: str = "cell_faces") -> tuple[list, dict]:
    """The wave bands of the responses from ``source`` to each of ``listeners`` in the room
    of ``model``, and run details: the reference's entry point (engine/wave.py:805-1271),
    its arguments, checks, results (``WaveResponse`` each) and ``info`` keys, stepped on the
    GPU.

    ``model`` is read as the reference's TraceModel is: ``room.mesh`` (``vertices``,
    ``faces``), ``room.face_materials``, ``room.material_names``, ``room.outdoor`` (None
    indoors), ``room.volume_m3``, ``absorption`` (faces x bands), ``transparent`` (faces),
    ``transmission`` (faces x bands, or None), ``face_impedance`` (None), ``bands``
    ("octave" or "third_octave"), ``ground`` (outdoors) and, outdoors,
    ``room.scene.atmosphere`` (an acoustic_spaces_amd.weather.Atmosphere, or None) and
    ``room.scene.air.temperature_c``.

    As the reference: each material's admittance is fitted to its absorption in the wave
    bands at the grid's rate; the grid is ``band_grid``'s for the highest band; the stencil
    is the isotropic one (``scheme`` "auto") unless the model is open, where the standard
    one with its 8-cell matched layer steps it (outdoors at once); outdoors under a weather,
    each cell's effective speed of sound from the source (the wind's component along the
    horizontal added), or with ``wind`` "moving" the still air's speed and the wind moving
    the air (each step's carrying, the run watched for sound that grows), a closed grid's
    air still under any wind. Moving air is stepped on the GPU too (the reference steps it
    on the CPU only). Differences, each refused or documented: ``device`` "cpu" (this
    package steps on the GPU), materials with an impedance (a porous ground's fit), and the
    room round an open model's points taken on every side (the reference leaves it off
    sides its surfaces close, ``_open_sides``). ``walls`` "fitted" fits the model to the
    grid with this package's fit (closed models off the grid's planes; the standard
    stencil; refused outdoors, as the reference refuses it)."""
    from acoustic_spaces_amd.admittance import MAX_ABSORPTION, fit_admittance
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import Bvh, WaveSolver
    from acoustic_spaces_amd.weather import Weather

    say = say or (lambda _: None)
    if device not in ("gpu", "auto", "hip", "cpu"):
        raise ValueError(f"unknown device {device!r}; use cpu, gpu or auto")
    if device == "cpu":
        raise ValueError("this package steps the wave grid on the GPU: set simulation.wave.device to gpu or auto")
    if max_band_hz not in WAVE_BANDS_HZ:
        raise ValueError(f"max_band_hz must be one of {WAVE_BANDS_HZ}, got {max_band_hz}")
    if scheme not in ("auto", "standard", "isotropic"):
        raise ValueError(f"unknown scheme {scheme!r}; use auto, standard or isotropic")
    if wind not in ("effective", "moving"):
        raise ValueError(f"unknown wind {wind!r}; use effective or moving")
    if walls not in ("cell_faces", "fitted"):
        raise ValueError(f"unknown walls {walls!r}; use cell_faces or fitted")
    fitted = walls == "fitted"
    if fitted and scheme == "isotropic":
        raise ValueError(FITTED_WALLS_STENCIL)
    for what, value in (("points_per_wavelength", points_per_wavelength), ("length_s", length_s),
                        ("sample_rate", sample_rate), ("speed_of_sound", speed_of_sound)):
        if not (math.isfinite(value) and value > 0.0):
            raise ValueError(f"{what} must be a positive number, not {value!r}")
    listeners = np.asarray(listeners, float).reshape(-1, 3)
    if len(names) != len(listeners):
        raise ValueError(f"{len(names)} names for {len(listeners)} listeners")
    room = model.room
    mesh = room.mesh
    solid = ~np.asarray(model.transparent, bool)
    if not solid.any():
        raise ValueError("the wave solver needs a surface that reflects: every face of the model lets all "
                         "sound through (an opening)")
    face_materials = np.asarray(room.face_materials)
    if getattr(model, "transmission", None) is not None:
        partial = np.unique(face_materials[solid & np.any(np.asarray(model.transmission) > 0, 1)])
        if len(partial):
            kinds = ", ".join(repr(room.material_names[m]) for m in partial)
            raise ValueError(f"the wave solver cannot pass sound through a surface; {kinds} let some through "
                             "(openings, with transmission 1, are fine): set simulation.wave.enabled to false or "
                             "give them no transmission")
    if getattr(model, "face_impedance", None) is not None:
        raise ValueError("materials with an impedance (a porous ground's) are not fitted by this package")
    faces = np.ascontiguousarray(np.asarray(mesh.faces)[solid], np.int64)
    vertices = np.asarray(mesh.vertices, float)
    bvh = Bvh(vertices, faces)
    source = np.asarray(source, float).reshape(3)
    points = np.vstack([source[None, :], listeners])
    centres = _band_centres(model.bands)
    per_octave = 1 if model.bands == "octave" else 3
    bands = (analysis.OCTAVE_BANDS_HZ.index(max_band_hz) + 1) * per_octave
    total = len(centres)
    rows, face_material = np.unique(np.round(np.asarray(model.absorption, float)[solid, :bands], 6), axis=0,
                                    return_inverse=True)
    face_material = np.asarray(face_material, np.int64).reshape(-1)
    outdoor = getattr(room, "outdoor", None) is not None
    weather = None
    if outdoor:
        scene = getattr(room, "scene", None)
        atmosphere = getattr(scene, "atmosphere", None)
        if atmosphere is not None and not atmosphere.still:
            extent = getattr(room.outdoor, "extent", None)  # the ground's, for the zones' polygons
            weather = Weather(atmosphere, float(scene.air.temperature_c), float(model.ground), source, extent)
    # Whether the wind blows somewhere over the scene: then with wind "moving" the air moves
    # (temperature alone needs no moving air), and a closed grid's air is still.
    windy = weather is not None and weather.winds_over(
        max(float(vertices[:, 2].max()), float(points[:, 2].max())) + 100.0)
    moving = windy and wind == "moving"
    if fitted and outdoor:
        raise ValueError(FITTED_WALLS_OUTDOORS)
    fit_seconds = []  # with fitted walls, the set-up's time on each grid made

    def layout(courant, layer, speed, per_wavelength):
        grid = band_grid(bands - 1, sample_rate, speed, per_wavelength, total, courant)
        h = grid.cell
        low = np.minimum(vertices.min(axis=0), points.min(axis=0) - 2.0 * h)
        high = np.maximum(vertices.max(axis=0), points.max(axis=0) + 2.0 * h)
        offset = h * (2.0 + layer + np.array([0.0137, 0.0219, 0.0291]))
        if layer:
            room_past = LAYER_ROOM * _spread_across(points[0], points[1:])
            below = np.ceil(np.maximum(low - (points.min(axis=0) - room_past), 0.0) / h)
            if outdoor:
                below[2] = 0.0
            low = low - h * below
            high = np.maximum(high, points.max(axis=0) + room_past)
        origin = low - offset
        shape = np.ceil((high - origin) / h).astype(np.int64) + 2 + layer
        if int(np.prod(shape)) > MAX_CELLS:
            what = " with the matched layer round this open model" if layer else ""
            raise ValueError(f"the wave grid up to the {max_band_hz} Hz band would have {int(np.prod(shape)):,} "
                             f"cells{what} (at most {MAX_CELLS:,}); lower max_band_hz or points_per_wavelength")
        return grid, origin, shape

    def prepare(stencil, layer, still_air=False):
        points_n, courant = STENCILS[stencil]
        speeds, extremes, still, flow = None, None, speed_of_sound, (None, None)
        if weather is None:
            grid, origin, shape = layout(courant, layer, speed_of_sound, points_per_wavelength)
        elif moving or still_air:
            # Fine enough for sound going against the fastest wind in the slowest still air,
            # in steps for the fastest still air (the air's motion is carried apart; still
            # air, ``still_air``, has none).
            slowest = fastest = float(weather.at(source)[0])
            for _ in range(8):
                grid, origin, shape = layout(courant, layer, fastest, points_per_wavelength * fastest / slowest)
                low, high, top = weather.moving_extremes(origin, grid.cell, shape)
                if still_air:
                    top = 0.0
                if low - top >= slowest and high <= fastest:
                    break
                slowest, fastest = min(slowest, low - top), max(fastest, high)
            speeds, *flow = weather.moving_fields(origin, grid.cell, shape, courant * grid.cell * grid.rate,
                                                  1.0 / (grid.cell * grid.rate))
            if still_air:
                flow = (None, None)
            extremes, still = (low - top, high + top), low
        else:
            # Fine enough for the slowest sound, in steps for the fastest, over the grid
            # those give (the reference's eight rounds).
            slowest = fastest = float(weather.at(source)[0])
            for _ in range(8):
                grid, origin, shape = layout(courant, layer, fastest, points_per_wavelength * fastest / slowest)
                low_speed, high_speed = weather.extremes(origin, grid.cell, shape)
                if low_speed >= slowest and high_speed <= fastest:
                    break
                slowest, fastest = min(slowest, low_speed), max(fastest, high_speed)
            speeds = weather.ratios(origin, grid.cell, shape, courant * grid.cell * grid.rate)
            extremes, still = (low_speed, high_speed), low_speed
        fits = [fit_admittance(row, total, grid.rate) for row in rows]
        branches = [f.branches(grid.rate) for f in fits]
        branch_material = np.concatenate([np.full(len(b), m, np.int64) for m, b in enumerate(branches)]) \
            if branches else np.zeros(0, np.int64)
        probes = np.vstack([probe_points(p, grid.cell) for p in listeners])
        try:
            began = time.perf_counter()
            solver = WaveSolver(bvh, face_material, np.array([f.beta for f in fits]),
                                np.vstack(branches) if branches else np.zeros((0, 3)), branch_material, origin,
                                grid.cell, int(shape[0]), int(shape[1]), int(shape[2]), source, probes, 1.0, layer,
                                LAYER_DAMPING, threads, points_n, speeds, probes_per_listener=PROBES,
                                flow_x=flow[0], flow_y=flow[1], on_device=not fitted, fitted_walls=fitted)
            if fitted:
                fit_seconds.append(time.perf_counter() - began)
        except ValueError as error:
            if "the wind round the model's solids" not in str(error):
                raise
            raise ValueError(f"the wave solver's moving air: {error} (on cells of {100 * grid.cell:.0f} cm; a sharp "
                             "edge speeds the wind up, the more the finer the grid): set simulation.wave.wind to "
                             "effective, or lower max_band_hz or points_per_wavelength") from error
        return solver, grid, fits, probes, courant, extremes, still

    if scheme == "isotropic" and outdoor:
        raise ValueError("outdoors the isotropic stencil is not available (it has no matched layer for the sky): "
                         "set simulation.wave.scheme to auto or standard")
    built = None
    if outdoor:
        built = prepare("standard", LAYER_CELLS)
        if not built[0].open:
            built = None
    if built is None:
        # A closed model's air is still whichever way the wind is taken: no wind blows
        # through its walls (outdoors, a source shut in a closed building).
        built = prepare("standard" if scheme == "standard" or fitted else "isotropic", 0, windy)
        if built[0].open:
            if scheme == "isotropic":
                raise ValueError("the isotropic stencil has no matched layer for an open model: set "
                                 "simulation.wave.scheme to auto or standard")
            built = prepare("standard", LAYER_CELLS)
    solver, grid, fits, probes, courant, extremes, still = built
    still_inside = windy and not solver.open  # the wind left out of a closed model
    stilled = moving or still_inside  # the grid's speeds are the still air's

    def speed_at(point):
        return weather.still_at(point) if stilled else float(weather.at(point)[0])

    h = grid.cell
    at_source = 1.0
    if weather is not None:
        at_source = min(speed_at(source) / (courant * h * grid.rate), 1.0)
    signal, delay = source_signal(grid.rate, h, courant, at_source)
    samples = int(math.ceil(length_s * grid.rate))
    whole = max(1, round(grid.rate / sample_rate))  # steps per output sample
    lead = -(-delay // whole) * whole
    steps = lead + samples
    drive = np.zeros(steps, np.float32)
    drive[lead - delay: lead - delay + len(signal)] = signal[: steps - lead + delay]
    found = np.array([solver.probe_found(p) and solver.probe_in_sight(p) for p in range(len(probes))])
    found = off_the_source(found, probes, source, h, solver.stencil).tolist()
    missing = [n for i, n in enumerate(names) if not found[PROBES * i]]
    if missing:
        air, volume = solver.air_cells * h ** 3, getattr(room, "volume_m3", None)
        if volume and air < 0.1 * volume:
            who = f"source {label!r}" if label else "the source"
            raise ValueError(f"{who} is closed in: the wave grid has {air:.2g} m³ of air round it, in a room of "
                             f"{volume:.3g} m³")
        raise ValueError(f"listener {missing[0]!r} is not in the air of the wave grid")
    band_name = f"{centres[0]} Hz" if bands == 1 else f"{centres[0]}-{centres[bands - 1]} Hz"
    where = ""
    if extremes is not None:
        where = f", sound at {extremes[0]:.0f} to {extremes[1]:.0f} m/s"
    if still_inside:
        where += ", the wind left out: the air inside the closed model is still"
    # The fastest wind round the model's solids (m/s), where it was made to flow round them.
    fastest = solver.wind_fastest * h * grid.rate if solver.wind_round else None
    if solver.flowing:
        where += ", the air moving"
    if fastest is not None:
        where += f", round the model's solids at up to {fastest:.0f} m/s"
    say(f"{label}: wave solver, {band_name} ({solver.air_cells:,} cells of {100 * h:.0f} cm, {steps:,} steps "
        f"on the GPU{where})")
    who = f"source {label!r}" if label else "the source"
    unstable = "set simulation.wave.wind to effective"
    # Long runs report every tenth, and moving air runs in the pieces it is watched over
    # (WATCH_S): once the source's pulse has ended, the pressure squared over the air,
    # averaged over a piece, may not rise past WATCH_RISE times its lowest so far.
    tenths = 10 if solver.air_cells * steps > LONG_RUN else 1
    watch = max(1, round(WATCH_S * grid.rate))
    chunks = max(tenths, steps // watch) if solver.flowing else tenths
    with HipWaveSolver(solver) as stepper:
        if chunks == 1:
            recorded = stepper.run(drive)
        else:
            recorded = np.empty((len(probes), steps))
            bounds = np.linspace(0, steps, chunks + 1).astype(np.int64)
            ended = lead - delay + len(signal)  # the first step after the source's pulse
            lowest, largest = math.inf, 0.0
            squares, since = 0.0, 0  # the watched piece so far: its squares' sum, its first step
            for c in range(chunks):
                if c and c * tenths // chunks > (c - 1) * tenths // chunks:
                    say(f"{label}: wave solver, {band_name}, {100 * (c * tenths // chunks) // tenths} %")
                recorded[:, bounds[c]: bounds[c + 1]] = stepper.run(drive[bounds[c]: bounds[c + 1]])
                if not solver.flowing:
                    continue
                squares += stepper.mean_square * float(bounds[c + 1] - bounds[c])
                if not math.isfinite(squares):
                    recorded[:, bounds[c + 1]:] = math.nan  # refused below
                    break
                if bounds[c + 1] - since < watch:
                    continue
                level = squares / float(bounds[c + 1] - since)
                start, squares, since = since, 0.0, int(bounds[c + 1])
                largest = max(largest, level)
                if start >= ended:
                    if level > WATCH_RISE * lowest and level > WATCH_FLOOR * largest:
                        raise ValueError(f"{who}: the wave solver's sound grew in the moving air instead of dying "
                                         f"away ({10 * math.log10(level / lowest):.1f} dB over its lowest level, "
                                         f"{(since - lead) / grid.rate:.1f} s after emission): {unstable}")
                    lowest = min(lowest, level)
    if not np.all(np.isfinite(recorded)):
        raise ValueError(f"{who}: the wave solver's run did not stay finite" + (f": {unstable}" if solver.flowing else ""))
    info = {"wave_cells": float(solver.air_cells), "wave_cell_m": h, "wave_steps": float(steps),
            "wave_stencil": float(solver.stencil), "wave_gpu": 1.0}
    if solver.open:
        info["wave_open"] = 1.0
    if extremes is not None:
        info["wave_speed_min_m_s"], info["wave_speed_max_m_s"] = extremes
    if solver.flowing:
        info["wave_moving_air"] = 1.0
    if fastest is not None:
        info["wave_wind_fastest_m_s"] = fastest
    if still_inside:
        info["wave_still_inside"] = 1.0
    if fitted:
        info["wave_fitted_walls"] = 1.0
        info["wave_fit_s"] = sum(fit_seconds)  # on every grid made
    slowest = speed_of_sound if extremes is None else extremes[0]
    per_wavelength = slowest / (band_top(bands - 1, total) * h)
    info["wave_slow_percent"] = 100.0 * slow_by(solver.stencil, per_wavelength, still / (courant * h * grid.rate))
    used = np.unique(face_material)
    wanted = np.minimum(rows[used], MAX_ABSORPTION)
    info["wave_absorption_miss"] = float(np.max(np.abs(np.array([fits[m].achieved for m in used]) - wanted)))
    responses = []
    for i, name in enumerate(names):
        part = slice(PROBES * i, PROBES * (i + 1))
        local = courant
        if weather is not None:
            local *= speed_at(listeners[i]) / (courant * h * grid.rate)
        signals = first_order(recorded[part], found[part], name, local, cutoff(courant))
        responses.append(WaveResponse(grid.rate, signals, bands, h, total, lead / grid.rate))
    return responses, info
