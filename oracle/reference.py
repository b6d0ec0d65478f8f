"""Test infrastructure: the reference's own wave solver, through oracle/_ref/libasref.so.

The library is built by ``make -C oracle ref`` from the reference's header-only C++
(cpp/include/acoustic_spaces/fdtd.hpp, read in place) and this repository's wrapper
ref_harness.cpp. It is the checker of the product's layout producer and of the oracle's
stepping, and the CPU baseline that bench.py times; the product never loads it.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "_ref" / "libasref.so"

_KINDS = {0: np.uint8, 1: np.int32, 2: np.int64, 3: np.uint64, 4: np.float32, 5: np.float64}
_lib = None


def available() -> bool:
    return LIBRARY.exists()


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise FileNotFoundError(f"{LIBRARY} is not built (make -C oracle ref)")
    lib = ctypes.CDLL(str(LIBRARY))
    P = ctypes.c_void_p
    S = ctypes.c_size_t
    D = ctypes.c_double
    lib.asref_create.restype = P
    lib.asref_create.argtypes = [P, S, P, S, P, P, S, P, P, S, P, D, S, S, S, P, P, S, D, S, D, S,
                                 S, P, P, P, S, ctypes.c_int, ctypes.c_char_p, S]
    lib.asref_layout_count.restype = ctypes.c_int
    lib.asref_layout_count.argtypes = [P]
    lib.asref_layout_entry.restype = ctypes.c_int
    lib.asref_layout_entry.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(S),
                                       ctypes.POINTER(P)]
    lib.asref_run.argtypes = [P, P, S, P]
    lib.asref_reset.argtypes = [P]
    lib.asref_destroy.argtypes = [P]
    lib.asref_info.argtypes = [P, P]
    lib.asref_open_sides.argtypes = [P, P]
    lib.asref_flow.argtypes = [P, P]
    lib.asref_probe_found.restype = ctypes.c_int
    lib.asref_probe_found.argtypes = [P, S]
    lib.asref_probe_in_sight.restype = ctypes.c_int
    lib.asref_probe_in_sight.argtypes = [P, S]
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class ReferenceSolver:
    """The reference's ``WaveSolver`` (fdtd.hpp:339-3601) built from a scene's arrays, with
    the Python contract of its binding: ``layout()``, ``run(signal)``, ``reset()`` and the
    counts (cpp/bindings/native_module.cpp:1561-1707)."""

    def __init__(self, vertices, faces, face_material, beta, branches, branch_material, origin,
                 cell, shape, source, probes, edge_admittance=1.0, layer=0, layer_damping=0.0,
                 threads=1, stencil=7, speed=None, probes_per_listener=1, fitted_walls=False,
                 flow_x=None, flow_y=None):
        lib = _load()
        self._keep = [
            np.ascontiguousarray(vertices, np.float64),
            np.ascontiguousarray(faces, np.int64),
            np.ascontiguousarray(face_material, np.int64),
            np.ascontiguousarray(beta, np.float64),
            np.ascontiguousarray(np.asarray(branches, np.float64).reshape(-1, 3)),
            np.ascontiguousarray(branch_material, np.int64),
            np.ascontiguousarray(origin, np.float64),
            np.ascontiguousarray(source, np.float64),
            np.ascontiguousarray(np.asarray(probes, np.float64).reshape(-1, 3)),
        ]
        v, f, fm, b, br, bm, o, s, p = self._keep
        sp = None if speed is None else np.ascontiguousarray(speed, np.float32)
        fx = None if flow_x is None else np.ascontiguousarray(flow_x, np.float32)
        fy = None if flow_y is None else np.ascontiguousarray(flow_y, np.float32)
        self._keep += [sp, fx, fy]
        err = ctypes.create_string_buffer(4096)
        nx, ny, nz = (int(x) for x in shape)
        h = lib.asref_create(_ptr(v), len(v), _ptr(f), len(f), _ptr(fm), _ptr(b), len(b), _ptr(br),
                             _ptr(bm), len(bm), _ptr(o), float(cell), nx, ny, nz, _ptr(s), _ptr(p),
                             len(p), float(edge_admittance), int(layer), float(layer_damping),
                             int(threads), int(stencil), _ptr(sp), _ptr(fx), _ptr(fy), int(probes_per_listener),
                             1 if fitted_walls else 0, err, len(err))
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        self.shape = (nx, ny, nz)
        info = np.zeros(8, np.uint64)
        lib.asref_info(h, _ptr(info))
        (self.air_cells, self.boundary_cells, self.branch_states, self.layer_cells, self.probes,
         self.threads) = (int(x) for x in info[:6])
        self.open = bool(info[6])
        self.source_on_face = bool(info[7])
        sides = np.zeros(6, np.int32)
        lib.asref_open_sides(h, _ptr(sides))
        self.open_sides = tuple(bool(x) for x in sides)

    @classmethod
    def from_scene(cls, scene, threads=1, **kw):
        a = scene.solver_args()
        return cls(threads=threads, **a, **kw)

    def layout(self) -> dict:
        lib = _load()
        out = {}
        name = ctypes.c_char_p()
        kind = ctypes.c_int()
        count = ctypes.c_size_t()
        data = ctypes.c_void_p()
        for k in range(lib.asref_layout_count(self._h)):
            lib.asref_layout_entry(self._h, k, ctypes.byref(name), ctypes.byref(kind),
                                   ctypes.byref(count), ctypes.byref(data))
            dtype = np.dtype(_KINDS[kind.value])
            n = count.value
            if n:
                buf = (ctypes.c_char * (n * dtype.itemsize)).from_address(data.value)
                arr = np.frombuffer(buf, dtype=dtype).copy()
            else:
                arr = np.zeros(0, dtype)
            out[name.value.decode()] = arr
        return out

    def run(self, signal) -> np.ndarray:
        sig = np.ascontiguousarray(signal, np.float32)
        out = np.zeros((self.probes, len(sig)), np.float64)
        _load().asref_run(self._h, _ptr(sig), len(sig), _ptr(out))
        return out

    def reset(self):
        _load().asref_reset(self._h)

    def flow(self) -> dict:
        """Moving air: flowing, wind_round, wind_fastest, wind_sweeps and the last run's
        mean_square, as the binding's properties give them."""
        out = np.zeros(5, np.float64)
        _load().asref_flow(self._h, _ptr(out))
        return {"flowing": bool(out[0]), "wind_round": bool(out[1]), "wind_fastest": float(out[2]),
                "wind_sweeps": int(out[3]), "mean_square": float(out[4])}

    _CUT_STATS = ("cubes", "regions", "ties", "farthest", "conduit_links", "conduit_scaled", "far_volume",
                  "diagonal_area", "dropped_area", "edge_area", "air_volume", "conduit_least",
                  "conduit_wall_moved")
    _FITTED_COUNTS = ("raised", "most_raised", "sliver_links", "sliver_both", "sliver_thick",
                      "layer_whole_links", "plain_rows")

    def cut_rows(self, materials: int) -> dict | None:
        """Fitted walls: the fit's rows as the solver keeps them (cut_cells.hpp:1249-1265):
        cells, air volumes, open areas (n x 18), wall areas (n x (materials + 1)), slivers
        (n x 6), the fit's counts and the solver's counts of the sliver rule; None without
        fitted walls."""
        lib = _load()
        P = ctypes.c_void_p
        lib.asref_cut_rows.restype = ctypes.c_long
        lib.asref_cut_rows.argtypes = [P, P, P, P, P, P, P]
        lib.asref_fitted_counts.argtypes = [P, P]
        n = lib.asref_cut_rows(self._h, None, None, None, None, None, None)
        if n < 0:
            return None
        out = {"cells": np.zeros(n, np.int64), "air_volume": np.zeros(n), "open_area": np.zeros((n, 18)),
               "wall_area": np.zeros((n, materials + 1)), "sliver": np.zeros((n, 6))}
        stats = np.zeros(len(self._CUT_STATS))
        lib.asref_cut_rows(self._h, _ptr(out["cells"]), _ptr(out["air_volume"]), _ptr(out["open_area"]),
                           _ptr(out["wall_area"]), _ptr(out["sliver"]), _ptr(stats))
        counts = np.zeros(len(self._FITTED_COUNTS))
        lib.asref_fitted_counts(self._h, _ptr(counts))
        out["stats"] = dict(zip(self._CUT_STATS, stats.tolist()))
        out["counts"] = dict(zip(self._FITTED_COUNTS, counts.tolist()))
        return out

    def probe_found(self, p) -> bool:
        return bool(_load().asref_probe_found(self._h, p))

    def probe_in_sight(self, p) -> bool:
        return bool(_load().asref_probe_in_sight(self._h, p))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.asref_destroy(h)
            self._h = None


def host_cores() -> int:
    return len(os.sched_getaffinity(0))
