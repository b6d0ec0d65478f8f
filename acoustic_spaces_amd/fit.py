"""Fitting wall admittance to measured responses, by the adjoint's gradient.

The model: each wall m of a room (walls on cell faces) has the specific admittance

    Y_m(s) = beta_m + G_m (s / w_c) / (1 + s / w_c),

a constant and one first-order (relaxation) branch with its corner at w_c, the reference's
branch Y = 1 / (M s + R + S / s) with M = 0, R = 1 / G, S = w_c / G (engine/impedance.py:
168-211). Centred in time (fdtd.hpp:391-411) the branch has m = 0, r = 1 / G, q = w_c T / G,
so a = (1 + w_c T) / (2 G): its gain 1 / (2 a) = G gamma with gamma = 1 / (1 + w_c T), and
its `before` coefficient 2 w_c T / (1 + w_c T) and `now` 0 do not depend on (beta, G). Every
special cell's wall term is then g = sum_m W_m (beta_m + G_m gamma), W_m the weight of its
cut links of wall m (its branch state's weight), and the layout's coef_b = 1 - g,
coef_inv = 1 / (1 + g) (add_special, fdtd.hpp:3499-3517). The 2 x walls unknowns
theta = (beta, G) enter the scheme only there, so dJ/dtheta follows from the adjoint's
dJ/dg and dJ/dgain (HipWaveSolver.gradient):

    dJ/dbeta_m = sum_k dJ/dg_k W_km,   dJ/dG_m = gamma (sum_k dJ/dg_k W_km + sum_{j of m} dJ/dgain_j).

`WaveSolve` wraps the solve as a torch.autograd.Function; `decay_loss` is the loss on
band-filtered Schroeder decay curves and the first 50 ms of pressure.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

from acoustic_spaces_amd.gpu import HipWaveSolver
from acoustic_spaces_amd.solver import WaveSolver
from acoustic_spaces_amd.wave import high_pass


class WallModel:
    """The admittance model above on a scene's room (its mesh's materials are the walls)."""

    def __init__(self, scene, corner_hz: float = 177.0, every: int = 256, precision: int = 32,
                 steps: int | None = None, keep_history=None):
        self.scene = scene
        self.walls = scene.mesh.materials
        T = 1.0 / scene.rate
        self.wcT = 2.0 * math.pi * corner_hz * T
        self.gamma = 1.0 / (1.0 + self.wcT)
        self.before = 2.0 * self.wcT / (1.0 + self.wcT)
        self.corner_hz = corner_hz
        self.every = every
        self.keep_history = keep_history  # see HipWaveSolver: None, the library's choice
        self.precision = precision
        self.steps = steps or scene.steps
        # The scene's pulse high-passed as the reference's source signal is (engine/wave.py:
        # 281-327, "so it leaves no air in the room"): a closed room would otherwise hold the
        # pulse's net volume as a constant pressure as large as the sound, on which the
        # float scheme's rounding (relative to |p|) sits some 55 dB under the early sound,
        # at the decay curves' floor.
        self.drive = np.ascontiguousarray(high_pass(scene.drive[: self.steps], scene.rate), np.float32)
        # The layout's structure (which cells are special, their states) holds for every theta;
        # lay it out once with placeholder values, one branch per wall.
        placeholder = 0.1
        laid = dataclasses.replace(
            scene, beta=np.full(self.walls, placeholder),
            branches=np.array([(0.0, 1.0 / placeholder, self.wcT / placeholder)] * self.walls),
            branch_material=np.arange(self.walls, dtype=np.int64))
        solver = WaveSolver.from_scene(laid)
        # A wall cell next to the grid's edge adds the edge's admittance to its wall term,
        # which the model leaves out: a room closed by its walls, or a matched layer round
        # the grid (whose cells are all the air next to the edge).
        if solver.open and not solver.layer_cells:
            raise ValueError("the fit needs a room closed by its walls or a matched layer round the grid")
        self.base = solver.layout()
        self.probes = solver.probes
        L = self.base
        ss = np.asarray(L["state_start"], np.int64)
        self.state_owner = np.repeat(np.arange(len(L["special"])), np.diff(ss))
        self.state_wall = laid.branch_material[np.asarray(L["state_branch"], np.int64)]
        self.state_weight = np.asarray(L["state_weight"], np.float64)
        K = len(L["special"])
        # W (specials x walls): each state stands for its wall's cut links at its cell.
        self.W = np.zeros((K, self.walls))
        np.add.at(self.W, (self.state_owner, self.state_wall), self.state_weight)

    def coefficients(self, theta):
        """The layout for theta = (beta (walls), G (walls))."""
        theta = np.asarray(theta, np.float64)
        beta, G = theta[: self.walls], theta[self.walls:]
        if np.any(beta < 0) or np.any(G <= 0):
            raise ValueError("beta >= 0 and G > 0")
        g = self.W @ (beta + G * self.gamma)
        L = dict(self.base)
        L["coef_b"] = (1.0 - g).astype(np.float32)
        L["coef_inv"] = (1.0 / (1.0 + g)).astype(np.float32)
        L["branch_now"] = np.zeros(self.walls, np.float32)
        L["branch_before"] = np.full(self.walls, self.before, np.float32)
        L["branch_gain"] = (G * self.gamma).astype(np.float32)
        if self.precision == 64:
            # The double scheme takes its coefficients unrounded: a smooth function of theta.
            L["coef_b64"] = 1.0 - g
            L["coef_inv64"] = 1.0 / (1.0 + g)
            L["branch_gain64"] = G * self.gamma
        return L

    def forward(self, theta) -> np.ndarray:
        """The probes' signals (probes, steps) for theta, from rest."""
        stepper = HipWaveSolver(self.coefficients(theta), precision=self.precision)
        try:
            return stepper.run(self.drive)
        finally:
            stepper.close()

    def record(self, theta):
        """The forward half of :meth:`gradient`: (y, the stepper holding the run's
        checkpoints, for :meth:`backward`)."""
        stepper = HipWaveSolver(self.coefficients(theta), precision=self.precision, no_faces=True,
                                keep_history=self.keep_history)
        try:
            y = stepper.record(self.drive, every=self.every)
        except Exception:
            stepper.close()
            raise
        return y, stepper

    def backward(self, stepper, dl_dy):
        """dJ/dtheta for dJ/dy (probes, steps) from a :meth:`record` stepper (closed after)."""
        try:
            sens_g, sens_gain = stepper.adjoint(dl_dy)
        finally:
            stepper.close()
        per_wall = sens_g @ self.W
        gains = np.bincount(self.state_wall, sens_gain, minlength=self.walls)
        return np.concatenate([per_wall, self.gamma * (per_wall + gains)])

    def gradient(self, theta, dl_dy):
        """(y, dJ/dtheta) for dJ/dy (probes, steps), by the adjoint with checkpoints every
        `every` steps."""
        y, stepper = self.record(theta)
        return y, self.backward(stepper, dl_dy)

    def absorption(self, theta, bands=(125.0, 250.0)) -> np.ndarray:
        """Normal-incidence absorption 1 - |(1 - Y) / (1 + Y)|^2 of each wall at the bands'
        centres (walls, bands)."""
        theta = np.asarray(theta, np.float64)
        beta, G = theta[: self.walls], theta[self.walls:]
        s = 2j * math.pi * np.asarray(bands)[None, :]
        wc = 2.0 * math.pi * self.corner_hz
        Y = beta[:, None] + G[:, None] * (s / wc) / (1.0 + s / wc)
        R = (1.0 - Y) / (1.0 + Y)
        return 1.0 - np.abs(R) ** 2

    def theta_for(self, alpha_low, alpha_high, bands=(125.0, 250.0)):
        """theta whose absorption is alpha_low at bands[0] and alpha_high at bands[1] for
        each wall (Newton on the two equations), for synthetic 'true' walls."""
        theta = np.concatenate([np.full(self.walls, 0.05), np.full(self.walls, 0.05)])
        target = np.stack([alpha_low, alpha_high], axis=1)
        for _ in range(60):
            a = self.absorption(theta, bands)
            r = (a - target).reshape(-1)
            if np.max(np.abs(r)) < 1e-12:
                break
            J = np.zeros((2 * self.walls, 2 * self.walls))
            for p in range(2 * self.walls):
                d = np.zeros(2 * self.walls)
                d[p] = 1e-7
                J[:, p] = ((self.absorption(theta + d, bands) - self.absorption(theta - d, bands)) / 2e-7).reshape(-1)
            step = np.linalg.solve(J, r)
            theta = np.maximum(theta - step, 1e-6)
        return theta


def _torch():
    import torch

    return torch


def make_function():
    """WaveSolve: y = solve(theta) as a torch.autograd.Function (theta float64 on the CPU)."""
    torch = _torch()

    class WaveSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model):
            ctx.model = model
            ctx.save_for_backward(theta)
            ctx.stepper = None
            if ctx.needs_input_grad[0]:
                # The forward run keeps its checkpoints: the backward pass only recomputes
                # each segment and steps the adjoint back through it.
                y, ctx.stepper = model.record(theta.detach().cpu().numpy())
            else:
                y = model.forward(theta.detach().cpu().numpy())
            return torch.from_numpy(y)

        @staticmethod
        def backward(ctx, grad_y):
            (theta,) = ctx.saved_tensors
            stepper, ctx.stepper = ctx.stepper, None
            if stepper is None:
                _, stepper = ctx.model.record(theta.detach().cpu().numpy())
            grad = ctx.model.backward(stepper, grad_y.detach().cpu().numpy())
            return torch.from_numpy(grad).to(theta.dtype), None

    return WaveSolve


def band_gains(n: int, rate: float, centre: float, order: int = 3):
    """Zero-phase octave-band magnitude over an n-point real FFT: Butterworth band-pass of
    `order`, edges centre / sqrt 2 and centre sqrt 2 (the reference's band_filter_gains,
    parameters.py:1062-1080, is the same family)."""
    f = np.fft.rfftfreq(n, 1.0 / rate)
    lo, hi = centre / math.sqrt(2.0), centre * math.sqrt(2.0)
    with np.errstate(divide="ignore"):
        hp = 1.0 / np.sqrt(1.0 + (lo / np.maximum(f, 1e-12)) ** (2 * order))
    lp = 1.0 / np.sqrt(1.0 + (f / hi) ** (2 * order))
    return hp * lp


def source_high_pass(n_fft: int, rate: float) -> np.ndarray:
    """The frequency response, on the rFFT grid of n_fft, of the reference's source
    high-pass taken twice (an impulse less a unit-area Hann window of 0.12 s, engine/wave.py:
    281-327; acoustic_spaces_amd.wave.high_pass): zero at DC, so a signal's constant part,
    the air a pulse leaves in a closed room, does not pass."""
    from acoustic_spaces_amd.wave import _hann_high_pass_taps

    taps = _hann_high_pass_taps(rate)
    return np.fft.rfft(np.convolve(taps, taps), n_fft)


def decay_loss(y, target, rate: float, bands=(125.0, 250.0), early_s: float = 0.05, floor_db: float = -40.0,
               early_weight: float = 1e4):
    """L2 on the band-filtered Schroeder decay curves (dB, down to floor_db of the target's
    start, averaged over the samples kept) plus early_weight times the L2 on the first
    early_s seconds of pressure (relative to the target's energy there): y and target
    (probes, steps) torch tensors. The decay curves fix the walls' absorption together; the
    early reflections tell the walls apart. early_weight 1e4 is 10^(SNR / 10) for targets at
    40 dB SNR: a pressure misfit at the noise level then weighs about as much as 1 dB^2 of
    decay curve (the early pressure is the maximum-likelihood term under white noise; on
    the 6 x 4 x 3 m shoebox a weight of 100 left the fit's minimum 0.02 from the true
    absorption, 1e4 within 0.007).

    The model's curve carries the target's noise: the target's band power per sample,
    read late in the record, integrated backwards and added, so that the curves compare the
    decay, not the noise floor the target has and the model lacks; they are compared while
    the target's remaining sound is at least its remaining noise.

    Both terms read the signals high-passed as the reference's source signal is
    (source_high_pass), through filters zero-padded to twice the record (no wrap-around):
    a Gaussian pulse leaves air in a closed room, a constant pressure as large as the
    sound, which in float walks with the rounding (about 0.7 % of the signal over 0.5 s on
    the 6 x 4 x 3 m shoebox); an unpadded band filter wraps its step at the record's end
    into the bands near the decay's floor, where the curves' logarithm magnifies it."""
    torch = _torch()
    n = y.shape[-1]
    n_fft = 1 << int(math.ceil(math.log2(2 * n)))
    high = source_high_pass(n_fft, rate)

    def filtered(x, gain):
        spectrum = torch.fft.rfft(x, n=n_fft, dim=-1) * torch.from_numpy(gain)
        return torch.fft.irfft(spectrum, n=n_fft, dim=-1)[..., :n]

    def backward_integral(power):
        return torch.flip(torch.cumsum(torch.flip(power, dims=[-1]), dim=-1), dims=[-1])

    def level(energy):
        return 10.0 * torch.log10(energy / energy[..., :1] + 1e-30)

    remaining = torch.arange(n, 0, -1, dtype=y.dtype)  # samples from each to the end
    total = y.new_zeros(())
    for centre in bands:
        gain = high * band_gains(n_fft, rate, centre)
        tb = filtered(target, gain)
        # The target's noise in the band, a power per sample read from 75 to 95 % of the
        # record (where the sound has decayed under it, clear of the filter's taper at the
        # end): added, as its backward integral, to the model's curve, so that both level
        # off alike; compared where the target's remaining sound is at least its remaining
        # noise (its energy twice the noise's) and above floor_db.
        noise = torch.mean(tb[..., (3 * n) // 4: (19 * n) // 20] ** 2, dim=-1, keepdim=True).detach()
        left_t = backward_integral(tb * tb)
        ct = level(left_t)
        yb = filtered(y, gain)
        cy = level(backward_integral(yb * yb) + noise * remaining)
        mask = ((left_t > 2.0 * noise * remaining) & (ct > floor_db)).to(y.dtype).detach()
        total = total + torch.sum(mask * (cy - ct) ** 2) / torch.sum(mask)
    k = int(early_s * rate)
    yh, th = filtered(y, high), filtered(target, high)
    scale = torch.sum(th[..., :k] ** 2) + 1e-30
    total = total + early_weight * torch.sum((yh[..., :k] - th[..., :k]) ** 2) / scale
    return total
