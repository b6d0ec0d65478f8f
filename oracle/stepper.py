"""Test infrastructure: the oracle's stepping (oracle/step.c), loaded through ctypes.

``OracleSolver(layout)`` steps a layout dict (the arrays WaveSolver::layout hands out,
reference fdtd.hpp:727-782) on one CPU thread, bit for bit as the reference's run()
(fdtd.hpp:786-904) in single precision, or the same scheme in double with precision=64.
Only tests, smoke() and bench.py's cpu_baseline leg use it.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"

P = ctypes.c_void_p
U64 = ctypes.c_uint64


class _Layout(ctypes.Structure):
    _fields_ = [
        ("nx", U64), ("ny", U64), ("nz", U64), ("stencil", U64),
        ("weights", P), ("centre", ctypes.c_float), ("total", ctypes.c_float),
        ("scale", P), ("layer_scale", P), ("types", P),
        ("specials", U64), ("special", P), ("link_start", P), ("link_direction", P),
        ("link_weight", P), ("coef_a", P), ("coef_b", P), ("coef_inv", P), ("source", P),
        ("state_start", P), ("states", U64), ("state_branch", P), ("state_weight", P),
        ("branch_now", P), ("branch_before", P), ("branch_gain", P),
        ("layer_cells", U64), ("layer_cell", P), ("keep", P * 6), ("push", P * 6),
        ("probes", U64), ("probe_start", P), ("probe_cells", P), ("probe_weights", P),
        ("flow_mask", P), ("flow_x", P), ("flow_y", P), ("wind", P * 3), ("loop", P * 3),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise FileNotFoundError(f"{LIBRARY} is not built (make -C oracle oracle)")
        lib = ctypes.CDLL(str(LIBRARY))
        lib.oracle_create.restype = P
        lib.oracle_create.argtypes = [ctypes.POINTER(_Layout), ctypes.c_int]
        lib.oracle_destroy.argtypes = [P]
        lib.oracle_reset.argtypes = [P]
        lib.oracle_run.argtypes = [P, P, ctypes.c_size_t, P]
        lib.oracle_field.argtypes = [P, P]
        _lib = lib
    return _lib


_AXES = ("x", "y", "z")


class OracleSolver:
    """Steps a layout on the CPU as the reference does; ``run(signal)`` returns the probes'
    pressures before each step, float64 (probes, steps), and keeps the state across calls."""

    def __init__(self, layout: dict, precision: int = 32):
        lib = _load()
        L = {k: np.ascontiguousarray(v) for k, v in layout.items()}
        typed = {
            "weights": np.float32, "scale": np.float32, "layer_scale": np.float32,
            "types": np.uint8, "special": np.int64, "link_start": np.uint64,
            "link_direction": np.uint8, "link_weight": np.float32, "coef_a": np.float32,
            "coef_b": np.float32, "coef_inv": np.float32, "source": np.float32,
            "state_start": np.uint64, "state_branch": np.int32, "state_weight": np.float32,
            "branch_now": np.float32, "branch_before": np.float32, "branch_gain": np.float32,
            "layer_cells": np.int64, "probe_start": np.uint64, "probe_cells": np.uint64,
            "probe_weights": np.float64,
        }
        for name, dt in typed.items():
            L[name] = np.ascontiguousarray(L[name], dtype=dt)
        for k in ("keep", "push"):
            for ax in _AXES:
                for face in ("", "face_"):
                    key = f"{k}_{face}{ax}"
                    L[key] = np.ascontiguousarray(L.get(key, np.zeros(0)), dtype=np.float32)
        self._arrays = L
        shape = [int(v) for v in L["shape"]]
        s = _Layout()
        s.nx, s.ny, s.nz = shape
        s.stencil = int(L["stencil"][0])
        s.weights = L["weights"].ctypes.data
        s.centre = float(L["centre"][0])
        s.total = float(L["total"][0])
        s.scale = L["scale"].ctypes.data if int(L["scaled"][0]) else None
        s.layer_scale = L["layer_scale"].ctypes.data if len(L["layer_scale"]) else None
        s.types = L["types"].ctypes.data
        s.specials = len(L["special"])
        for name in ("special", "link_start", "link_direction", "link_weight", "coef_a", "coef_b",
                     "coef_inv", "source", "state_start", "state_branch", "state_weight",
                     "branch_now", "branch_before", "branch_gain", "probe_start", "probe_cells",
                     "probe_weights"):
            setattr(s, name, L[name].ctypes.data)
        s.states = len(L["state_branch"])
        s.layer_cells = len(L["layer_cells"])
        s.layer_cell = L["layer_cells"].ctypes.data
        for a, ax in enumerate(_AXES):
            s.keep[2 * a] = L[f"keep_{ax}"].ctypes.data
            s.keep[2 * a + 1] = L[f"keep_face_{ax}"].ctypes.data
            s.push[2 * a] = L[f"push_{ax}"].ctypes.data
            s.push[2 * a + 1] = L[f"push_face_{ax}"].ctypes.data
        s.probes = len(L["probe_start"]) - 1
        # Moving air: the masks and the plain shifts, or the faces' winds and loops.
        mask = L.get("flow_mask", np.zeros(0))
        if len(mask):
            L["flow_mask"] = np.ascontiguousarray(mask, np.uint8)
            s.flow_mask = L["flow_mask"].ctypes.data
            if len(L.get("flow_x", ())):
                for k in ("flow_x", "flow_y"):
                    L[k] = np.ascontiguousarray(L[k], np.float32)
                s.flow_x, s.flow_y = L["flow_x"].ctypes.data, L["flow_y"].ctypes.data
            else:
                for a, ax in enumerate(_AXES):
                    for k in ("wind", "loop"):
                        L[f"{k}_{ax}"] = np.ascontiguousarray(L[f"{k}_{ax}"], np.float32)
                        getattr(s, k)[a] = L[f"{k}_{ax}"].ctypes.data
        self._struct = s
        self.probes = int(s.probes)
        self.shape = tuple(shape)
        self._h = lib.oracle_create(ctypes.byref(s), int(precision))

    def run(self, signal) -> np.ndarray:
        sig = np.ascontiguousarray(signal, np.float32)
        out = np.zeros((self.probes, len(sig)), np.float64)
        _load().oracle_run(self._h, sig.ctypes.data, len(sig), out.ctypes.data)
        return out

    def reset(self):
        _load().oracle_reset(self._h)

    def field(self) -> np.ndarray:
        out = np.zeros(int(np.prod(self.shape)), np.float64)
        _load().oracle_field(self._h, out.ctypes.data)
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_destroy(h)
            self._h = None
