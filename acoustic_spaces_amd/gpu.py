"""The wave solver's stepper on MI355X: the drop-in for the reference's GpuWaveSolver.

The reference picks its stepper at one line of wave_responses,
``stepper = GpuWaveSolver(solver) if on_gpu else solver`` (engine/wave.py:1178); the
stepper reads ``solver.layout()`` once and offers ``run(signal) -> (probes, steps)`` and
``probes`` (engine/gpu.py:445-612). ``HipWaveSolver`` offers the same, over the HIP
kernels of libasfdtd.so, for a layout from the reference's ``_native.WaveSolver`` or from
this package's own set-up (acoustic_spaces_amd.solver.WaveSolver).
"""

from __future__ import annotations

import ctypes

import numpy as np

from acoustic_spaces_amd import native


class HipWaveSolver:
    """Steps the grid of ``solver`` (anything with ``layout()``, fresh or reset, or a layout
    dict) on the GPU: :meth:`run` does what the solver's ``run`` does, from the GPU's own
    state. ``precision`` 32 is the reference's float scheme, bit for bit; 64 is the same
    scheme in double. ``tile_z``, ``row_align`` and ``graph_steps`` set the launch shape (0: the
    library's defaults; graph_steps < 0 launches every step from the host); ``split`` launches
    each kind of cell on its own instead of one launch per step; ``face_blocks`` steps the
    face specials by blocks of their own (True) or in their tiles (False; None: the
    library's choice); ``pairs`` takes two steps by two launches where that applies (True,
    or 8 or 16 for the planes a box marches). ``keep_history`` (record and adjoint): True keeps
    every step's history of the special cells in the record, so that the adjoint recomputes
    nothing; False keeps the checkpoints only, each segment recomputed from its own; None,
    the library's choice (the history where it fits 16 GiB and a quarter of free memory).
    ``special_order`` and ``special_kinds`` are as_tuning's (the general specials' order in
    their blocks, and whether they read their kinds' tables: 2, or their own slabs: 1)."""

    def __init__(self, solver, precision: int = 32, tile_z: int = 0, row_align: int = 0,
                 graph_steps: int = 0, split: bool = False, tile_order: int = 0,
                 no_faces: bool = False, face_blocks=None, pairs: bool = False, keep_history=None,
                 special_order: int = 0, special_kinds: int = 0) -> None:
        lib = native.load()
        layout = solver if isinstance(solver, dict) else solver.layout()
        struct, keep = native.pack_layout(layout)
        tuning = native.Tuning(0, int(tile_z), int(row_align), int(graph_steps), 1 if split else 0,
                               int(tile_order), 1 if no_faces else 0,
                               0 if face_blocks is None else (1 if face_blocks else -1),
                               int(pairs) if not isinstance(pairs, bool) and int(pairs) > 1 else (1 if pairs else 0),
                               0 if keep_history is None else (1 if keep_history else -1), int(special_order),
                               int(special_kinds))
        h = lib.as_fdtd_create(ctypes.byref(struct), int(precision), ctypes.byref(tuning))
        if not h:
            message = lib.as_fdtd_create_error().decode()
            if "hip" in message.lower() or "device" in message.lower():
                raise RuntimeError(message)
            raise ValueError(message)
        self._h = h
        self._lib = lib
        self._probes = int(struct.probes)
        self._specials = int(struct.specials)
        self._states = int(struct.states)
        self.shape = (int(struct.nx), int(struct.ny), int(struct.nz))
        self.precision = int(precision)
        del keep  # the library holds its own copies

    @property
    def probes(self) -> int:
        return self._probes

    def _check(self, code: int):
        if code != native.AS_OK:
            native.raise_for(code, self._lib.as_fdtd_last_error(self._h).decode())

    def run(self, signal: np.ndarray) -> np.ndarray:
        """Advances one step per sample of ``signal`` (added at the source) and returns
        the pressure at every probe before each step (probes, steps)."""
        signal = np.ascontiguousarray(signal, dtype=np.float32).reshape(-1)
        out = np.empty((self._probes, len(signal)), np.float64)
        if len(signal):
            self._check(self._lib.as_fdtd_run(self._h, signal.ctypes.data, len(signal), out.ctypes.data))
        return out

    def run_device(self, signal_ptr: int, steps: int, out_ptr: int) -> None:
        """Enqueues ``steps`` steps on the solver's stream, on device pointers (the drive,
        float32[steps]; the probes' output, float64[probes, steps]); does not wait."""
        self._check(self._lib.as_fdtd_run_device(self._h, ctypes.c_void_p(signal_ptr), int(steps),
                                                 ctypes.c_void_p(out_ptr)))

    def gradient(self, signal, dl_dout, every: int = 256):
        """The adjoint run (see as_fdtd_gradient in include/asfdtd.h): from rest, the probes'
        signals for ``signal`` and the gradient of a loss J given dJ/dy (probes, steps) with
        respect to each special cell's wall term g and each branch state's gain. Returns
        (y (probes, steps), dJ/dg (specials,), dJ/dgain (states,)). The solver is left at rest;
        it must be made with ``no_faces=True``."""
        signal = np.ascontiguousarray(signal, dtype=np.float32).reshape(-1)
        steps = len(signal)
        g = np.ascontiguousarray(dl_dout, dtype=np.float64).reshape(self._probes, steps)
        y = np.empty((self._probes, steps), np.float64)
        sens_g = np.zeros(self._specials, np.float64)
        sens_gain = np.zeros(self._states, np.float64)
        self._check(self._lib.as_fdtd_gradient(self._h, signal.ctypes.data, steps, int(every), g.ctypes.data,
                                               y.ctypes.data, sens_g.ctypes.data, sens_gain.ctypes.data))
        return y, sens_g, sens_gain

    def record(self, signal, every: int = 256) -> np.ndarray:
        """The forward half of :meth:`gradient`: from rest, the probes' signals for
        ``signal`` (probes, steps), with the checkpoints kept for :meth:`adjoint`."""
        signal = np.ascontiguousarray(signal, dtype=np.float32).reshape(-1)
        self._recorded = len(signal)
        y = np.empty((self._probes, len(signal)), np.float64)
        self._check(self._lib.as_fdtd_record(self._h, signal.ctypes.data, len(signal), int(every), y.ctypes.data))
        return y

    def adjoint(self, dl_dout):
        """The backward half of :meth:`gradient`, after :meth:`record`: (dJ/dg (specials,),
        dJ/dgain (states,)) for dJ/dy (probes, recorded steps). The solver is left at rest."""
        steps = getattr(self, "_recorded", 0)
        if not steps:
            raise ValueError("no forward run recorded: call record() first")
        g = np.ascontiguousarray(dl_dout, dtype=np.float64).reshape(self._probes, steps)
        sens_g = np.zeros(self._specials, np.float64)
        sens_gain = np.zeros(self._states, np.float64)
        self._recorded = 0
        self._check(self._lib.as_fdtd_adjoint(self._h, g.ctypes.data, sens_g.ctypes.data, sens_gain.ctypes.data))
        return sens_g, sens_gain

    @property
    def mean_square(self) -> float:
        """In moving air, the pressure squared over the grid's air and the steps of the last
        run, on average; 0 in still air (the binding's mean_square, native_module.cpp:1606-1608)."""
        out = ctypes.c_double()
        self._check(self._lib.as_fdtd_mean_square(self._h, ctypes.byref(out)))
        return float(out.value)

    def reset(self) -> None:
        """Zeroes the fields and states (WaveSolver::reset, fdtd.hpp:906-915)."""
        self._check(self._lib.as_fdtd_reset(self._h))

    def field(self) -> np.ndarray:
        """The current pressure (the next step's probes read it), float64, x fastest."""
        out = np.empty(int(np.prod(self.shape)), np.float64)
        self._check(self._lib.as_fdtd_field(self._h, out.ctypes.data))
        return out

    @property
    def stream(self) -> int:
        """The HIP stream the steps run on (for events)."""
        return int(self._lib.as_fdtd_stream(self._h) or 0)

    def info(self) -> dict:
        out = np.zeros(11, np.uint64)
        self._check(self._lib.as_fdtd_info(self._h, out.ctypes.data))
        keys = ("interior_cells", "specials", "states", "layer_cells", "probes", "tiles",
                "row_pitch", "steps_done", "pair_planes", "pair_cells", "special_kinds")
        return {k: int(v) for k, v in zip(keys, out)}

    def close(self) -> None:
        """Frees the handle's device memory and graphs now (the garbage collector's __del__
        is a fallback, not the way to release a handle)."""
        h = getattr(self, "_h", None)
        if h:
            self._lib.as_fdtd_destroy(h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    @classmethod
    def run_once(cls, solver, signal, **kwargs) -> np.ndarray:
        """One run from rest on a handle of its own, closed before returning."""
        with cls(solver, **kwargs) as s:
            return s.run(signal)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
