"""The receivers' post-processing on the device: from the stepper's probe output in device
memory to each listener's W, X, Y, Z and band energies, without the run leaving the GPU.

The reference reads a run on the host: ``_first_order`` (engine/wave.py:757-802) for W, X,
Y, Z, and ``WaveResponse.energy`` (engine/wave.py:411-454) for the band energies per time
slot that its room parameters are read from. Here:
- the cell differences and their trapezoidal integral are a HIP kernel of this library
  (``as_first_order_device``, csrc/responses.hip), equal to the host reading bit for bit;
- the spectral steps (the trapezoid's gain correction, the zero-phase band filters) are
  FFTs on the device (torch.fft, rocFFT), equal to the host's NumPy FFTs to rounding;
- ``run_on_device`` steps a stepper into a device buffer and reads it there.

Tensors are torch tensors on the GPU; the library's kernel runs on torch's current stream.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from acoustic_spaces_amd import analysis, native
from acoustic_spaces_amd.wave import COURANT, LOW_PASS_CUTOFF, PROBES


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the device post-processing needs a GPU")
    return torch


def first_order(d_probes, found, courant, cut: float = LOW_PASS_CUTOFF, correct: bool = True):
    """W, X, Y, Z (listeners, steps, 4) on the device from the probe output ``d_probes``
    (13 listeners rows of steps, float64, on the GPU), ``found`` (13 per listener) and each
    listener's Courant number ``courant``; with ``correct``, the trapezoid's gain divided out
    up to ``cut`` of the sample rate (as the host reading does); without, the kernel's
    integrals alone."""
    torch = _torch()
    lib = native.load()
    rows, steps = int(d_probes.shape[0]), int(d_probes.shape[1])
    if rows % PROBES:
        raise ValueError(f"the probes come in groups of {PROBES}, not {rows} rows")
    listeners = rows // PROBES
    dev = d_probes.device
    probes = d_probes.to(torch.float64).contiguous()
    f = torch.as_tensor(np.asarray(found, bool).astype(np.uint8).reshape(-1), device=dev)
    c = torch.as_tensor(np.broadcast_to(np.asarray(courant, float), (listeners,)).copy(), device=dev)
    if len(f) != rows:
        raise ValueError("one found flag per probe")
    out = torch.empty((listeners, 4, steps), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    code = lib.as_first_order_device(ctypes.c_void_p(probes.data_ptr()), steps, listeners,
                                     ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(c.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    native.raise_for(code, "as_first_order_device failed")
    if correct:
        n_fft = 1 << int(math.ceil(math.log2(2 * steps + 1)))
        half_turn = np.pi * np.fft.rfftfreq(n_fft)
        gain = np.ones(len(half_turn))
        inside = (half_turn > 0) & (half_turn <= np.pi * cut)
        gain[inside] = np.tan(half_turn[inside]) / half_turn[inside]
        spectrum = torch.fft.rfft(out[:, 1:, :], n=n_fft, dim=-1) * torch.as_tensor(gain, device=dev)
        out[:, 1:, :] = torch.fft.irfft(spectrum, n=n_fft, dim=-1)[..., :steps]
    return out.transpose(1, 2)


def band_energy(signals, rate: float, bands: int, time_step: float, slots: int, lead_s: float = 0.0,
                delay_s: float = 0.0, band_gain=None):
    """Energy per band and time slot (listeners, bands, slots) of each listener's W, as
    ``WaveResponse.energy`` counts it (engine/wave.py:411-454), on the device: ``signals``
    (listeners, steps, 4) at ``rate`` from ``lead_s`` before emission; every band's
    zero-phase filtered pressure squared over ENERGY_TO_IR times the filter's share of the
    spectrum (times ``band_gain``, the feed's energy gain per band), into the slot of its
    time from emission."""
    torch = _torch()
    dev = signals.device
    listeners, steps = int(signals.shape[0]), int(signals.shape[1])
    pad = int(round(analysis.PAD_S * rate))
    n = steps + pad
    n_fft = analysis.fft_size(n, rate)
    gains = analysis.band_filter_gains(n_fft, rate)[:bands]
    one_sided = np.full(gains.shape[1], 2.0)
    one_sided[0] = one_sided[-1] = 1.0
    amp2 = np.ones(bands) if band_gain is None else np.asarray(band_gain, float)
    per_energy = amp2 / (analysis.ENERGY_TO_IR * (gains ** 2 @ one_sided) / n_fft)
    pressure = torch.nn.functional.pad(signals[:, :, 0], (pad, 0))
    spectrum = torch.fft.rfft(pressure, n=n_fft, dim=-1)
    banded = torch.fft.irfft(spectrum[:, None, :] * torch.as_tensor(gains, device=dev), n=n_fft, dim=-1)[..., :n]
    lead = int(round(lead_s * rate))
    t = (np.arange(n) - pad - lead) / rate + delay_s
    slot = np.maximum(t / time_step, 0).astype(np.int64)
    keep = np.flatnonzero(slot < slots)
    power = banded[..., torch.as_tensor(keep, device=dev)] ** 2 * torch.as_tensor(per_energy, device=dev)[None, :, None]
    out = torch.zeros((listeners, bands, slots), dtype=torch.float64, device=dev)
    out.index_add_(2, torch.as_tensor(slot[keep], device=dev), power)
    return out


def run_on_device(stepper, drive):
    """The stepper's probe output (probes, steps) for ``drive``, in a device buffer (a torch
    tensor), its steps enqueued on the stepper's own stream and waited for."""
    torch = _torch()
    dev = torch.device("cuda")
    d_drive = torch.as_tensor(np.ascontiguousarray(drive, np.float32), device=dev)
    out = torch.empty((max(stepper.probes, 1), len(drive)), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    stepper.run_device(d_drive.data_ptr(), len(drive), out.data_ptr())
    torch.cuda.synchronize()
    return out[: stepper.probes]


def courant_of(stencil: int) -> float:
    from acoustic_spaces_amd.wave import ISOTROPIC_COURANT

    return COURANT if stencil == 7 else ISOTROPIC_COURANT
