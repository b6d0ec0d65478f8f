"""The wave solver's set-up: the reference's ``_native.WaveSolver`` constructor and layout.

``WaveSolver(bvh, face_material, beta, branches, branch_material, origin, cell, nx, ny, nz,
source, probes, ...)`` takes the reference binding's arguments (cpp/bindings/
native_module.cpp:1238-1285, 1561-1580) and lays the grid out on the host, in C++
(csrc/setup.cpp, through libasfdtd.so): walls on cell faces, both stencils, a speed of sound
per cell, the matched layer, admittance branches. ``layout()`` hands out the same arrays as
the reference's (WaveSolver::layout, fdtd.hpp:727-782), to be stepped by
``acoustic_spaces_amd.gpu.HipWaveSolver``. ``run`` steps it on the GPU, so that the object is
also the reference solver's drop-in where it is stepped directly.

``on_device=True`` runs the set-up's two passes over the whole grid, the cut links and the
air's flood, on the GPU (csrc/setup_device.hip); the layout and every refusal are the host
set-up's bit for bit.

``fitted_walls=True`` lays the grid out on fitted walls (csrc/fitted.cpp): each cell a face
meets is cut exactly and holds its own air volume, open areas and wall areas
(set_up_fitted, fdtd.hpp:970-1264), for closed meshes that lie off the grid's planes.

Not here: sources or probes that lie on a face (the reference's leans): refused.
"""

from __future__ import annotations

import ctypes

import numpy as np

from acoustic_spaces_amd import native

P = ctypes.c_void_p
S = ctypes.c_size_t
D = ctypes.c_double

_setup_lib = None
SETUP_ON_DEVICE = 1  # AS_SETUP_ON_DEVICE: the cut links and the air's flood on the GPU
SETUP_FITTED = 2  # AS_SETUP_FITTED: fitted walls, each cut cell's air, open faces and walls exact


def _lib():
    global _setup_lib
    if _setup_lib is None:
        lib = native.load()
        lib.as_setup_create.restype = P
        lib.as_setup_create.argtypes = [P, S, P, S, P, P, S, P, P, S, P, D, S, S, S, P, P, S, D, S, D,
                                        S, P, S, ctypes.c_int]
        lib.as_setup_create_ex.restype = P
        base = list(lib.as_setup_create.argtypes)
        # ... speed, then flow_x and flow_y, then probes_per_listener, threads, and the flags.
        lib.as_setup_create_ex.argtypes = [*base[:-2], P, P, *base[-2:], ctypes.c_uint]
        lib.as_setup_create_error.restype = ctypes.c_char_p
        lib.as_setup_create_error.argtypes = []
        lib.as_setup_layout.restype = ctypes.c_int
        lib.as_setup_layout.argtypes = [P, ctypes.POINTER(native.Layout)]
        lib.as_setup_info.restype = ctypes.c_int
        lib.as_setup_info.argtypes = [P, P]
        lib.as_setup_wind.restype = ctypes.c_int
        lib.as_setup_wind.argtypes = [P, P, P, P]
        lib.as_setup_probe.restype = ctypes.c_int
        lib.as_setup_probe.argtypes = [P, S, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        lib.as_setup_destroy.restype = None
        lib.as_setup_destroy.argtypes = [P]
        lib.as_setup_cut_cells.restype = ctypes.c_int
        lib.as_setup_cut_cells.argtypes = [P, P, P, P, P, P, P, P, P, P]
        _setup_lib = lib
    return _setup_lib


class Bvh:
    """The mesh a WaveSolver is laid out over: ``Bvh(vertices (V, 3), faces (F, 3))``, as the
    reference's ``_native.Bvh`` is built (native_module.cpp:1422-1434)."""

    def __init__(self, vertices, faces, split: str = "sah"):
        if split not in ("sah", "median"):
            raise ValueError("Bvh: split must be 'sah' or 'median'")
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, dtype=np.int64).reshape(-1, 3)
        if len(self.faces) == 0:
            raise ValueError("Bvh: the mesh has no faces")
        if self.faces.min() < 0 or self.faces.max() >= len(self.vertices):
            raise ValueError("Bvh: face refers to a missing vertex")

    def face_count(self) -> int:
        return len(self.faces)


def _array(ptr, count, dtype):
    dtype = np.dtype(dtype)
    if count == 0 or not ptr:
        return np.zeros(0, dtype)
    buf = (ctypes.c_char * (count * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).copy()


class WaveSolver:
    """The grid of the reference's WaveSolver, laid out here (see the module's notes)."""

    def __init__(self, bvh, face_material, beta, branches, branch_material, origin, cell, nx, ny,
                 nz, source, probes, edge_admittance=1.0, layer=0, layer_damping=0.0, threads=0,
                 stencil=7, speed=None, flow_x=None, flow_y=None, probes_per_listener=1,
                 fitted_walls=False, vertices=None, faces=None, on_device=False):
        if (flow_x is None) != (flow_y is None):
            raise ValueError("the air's motion needs both of its components")
        lib = _lib()
        fm = np.ascontiguousarray(face_material, np.int64)
        if len(fm) != bvh.face_count():
            raise ValueError("WaveSolver: one material per face")
        b = np.ascontiguousarray(beta, np.float64)
        br = np.ascontiguousarray(np.asarray(branches, np.float64).reshape(-1, 3))
        bm = np.ascontiguousarray(branch_material, np.int64)
        if len(br) != len(bm):
            raise ValueError("WaveSolver: m, r and q of each branch, and its material")
        o = np.ascontiguousarray(origin, np.float64).reshape(3)
        s = np.ascontiguousarray(source, np.float64).reshape(3)
        p = np.ascontiguousarray(np.asarray(probes, np.float64).reshape(-1, 3))
        n = int(nx) * int(ny) * int(nz)
        sp = None
        if speed is not None:
            sp = np.ascontiguousarray(speed, np.float32)
            if len(sp) != n:
                raise ValueError("WaveSolver: one speed per cell of the grid")
        fx = fy = None
        if flow_x is not None:
            fx = np.ascontiguousarray(flow_x, np.float32).reshape(-1)
            fy = np.ascontiguousarray(flow_y, np.float32).reshape(-1)
            if len(fx) != n or len(fy) != n:
                raise ValueError("WaveSolver: the air's motion, one shift per cell of the grid")
        self._keep = (bvh, fm, b, br, bm, o, s, p, sp, fx, fy)
        ptr = lambda a: None if a is None or len(a) == 0 else a.ctypes.data  # noqa: E731
        h = lib.as_setup_create_ex(
            bvh.vertices.ctypes.data, len(bvh.vertices), bvh.faces.ctypes.data, len(bvh.faces),
            fm.ctypes.data, b.ctypes.data, len(b), ptr(br), ptr(bm), len(bm), o.ctypes.data,
            float(cell), int(nx), int(ny), int(nz), s.ctypes.data, ptr(p), len(p),
            float(edge_admittance), int(layer), float(layer_damping), int(stencil), ptr(sp), ptr(fx), ptr(fy),
            int(probes_per_listener), int(threads),
            (SETUP_ON_DEVICE if on_device else 0) | (SETUP_FITTED if fitted_walls else 0))
        if not h:
            raise ValueError(lib.as_setup_create_error().decode())
        self._h = h
        self._shape = (int(nx), int(ny), int(nz))
        self._stencil = int(stencil)
        info = np.zeros(8, np.uint64)
        lib.as_setup_info(h, info.ctypes.data)
        self.air_cells = int(info[0])
        self.boundary_cells = int(info[1])
        self.branch_states = int(info[2])
        self.layer_cells = int(info[3])
        self._probes = int(info[4])
        self.open = bool(info[5])
        self.open_sides = tuple(bool((int(info[6]) >> k) & 1) for k in range(6))
        self.source_on_face = bool(info[7])
        self.flowing = fx is not None
        # Moving air's set-up (the binding's wind_round, wind_fastest, wind_sweeps).
        rnd, fastest, sweeps = ctypes.c_int(), ctypes.c_double(), ctypes.c_uint64()
        lib.as_setup_wind(h, ctypes.byref(rnd), ctypes.byref(fastest), ctypes.byref(sweeps))
        self.wind_round = bool(rnd.value)
        self.wind_fastest = float(fastest.value)
        self.wind_sweeps = int(sweeps.value)
        self._stepper = None
        self._layout = None
        self._fitted = bool(fitted_walls)
        self._materials = len(b)

    @classmethod
    def from_scene(cls, scene, threads=0, **kw):
        """The solver of an ``acoustic_spaces_amd.scenes.Scene``."""
        a = scene.solver_args()
        nx, ny, nz = a.pop("shape")
        bvh = Bvh(a.pop("vertices"), a.pop("faces"))
        return cls(bvh, a.pop("face_material"), a.pop("beta"), a.pop("branches"),
                   a.pop("branch_material"), a.pop("origin"), a.pop("cell"), nx, ny, nz,
                   a.pop("source"), a.pop("probes"), threads=threads, **a, **kw)

    @property
    def probes(self) -> int:
        return self._probes

    @property
    def stencil(self) -> int:
        return self._stencil

    _CUT_STATS = ("cubes", "regions", "ties", "farthest", "conduit_links", "conduit_scaled", "far_volume",
                  "diagonal_area", "dropped_area", "edge_area", "air_volume", "conduit_least",
                  "conduit_wall_moved", "faces", "planes", "leaves", "squares", "quantum_bits")

    @property
    def fitted_walls(self) -> bool:
        """Whether the model's faces were fitted to the grid (the binding's fitted_walls)."""
        return self._fitted

    def cut_cells(self) -> dict:
        """With fitted walls, the fit's rows as the reference's ``cut_cells()`` hands them out
        (native_module.cpp:1614-1632): cell, air_volume (cells^3), open_area (n, 18),
        wall_area (n, materials + 1; the last column the grid's edge), leak_area (zeros),
        sliver (n, 6), and the solver's special cells with the volume each is stepped with;
        empty otherwise."""
        if not self._fitted:
            return {}
        rows = ctypes.c_uint64()
        ptrs = [ctypes.c_void_p() for _ in range(7)]
        stats = np.zeros(len(self._CUT_STATS))
        if _lib().as_setup_cut_cells(self._h, ctypes.byref(rows), *(ctypes.byref(p) for p in ptrs),
                                     stats.ctypes.data) != 0:
            raise RuntimeError("as_setup_cut_cells failed")
        n = rows.value
        cell, vol, open_, wall, sliver, svol, kind = (p.value for p in ptrs)
        lay = self.layout()
        k = len(lay["special"])
        w = self._materials + 1
        return {
            "cell": _array(cell, n, np.int64),
            "air_volume": _array(vol, n, np.float64),
            "open_area": _array(open_, 18 * n, np.float64).reshape(n, 18),
            "wall_area": _array(wall, w * n, np.float64).reshape(n, w),
            "leak_area": np.zeros(n),
            "sliver": _array(sliver, 6 * n, np.float64).reshape(n, 6),
            "special": lay["special"],
            "volume": _array(svol, k, np.float64),
        }

    @property
    def cut_stats(self) -> dict:
        """With fitted walls, the fit's counts (the binding's cut_stats, native_module.cpp:
        1631-1655, those this fit keeps); empty otherwise."""
        if not self._fitted:
            return {}
        stats = np.zeros(len(self._CUT_STATS))
        _lib().as_setup_cut_cells(self._h, None, None, None, None, None, None, None, None, stats.ctypes.data)
        return dict(zip(self._CUT_STATS, stats.tolist()))

    def cut_air(self) -> np.ndarray:
        """With fitted walls, each cell's kind in the fit (0 solid, 1 air, 2 matched layer)."""
        if not self._fitted:
            return np.zeros(0, np.uint8)
        kind = ctypes.c_void_p()
        _lib().as_setup_cut_cells(self._h, None, None, None, None, None, None, None, ctypes.byref(kind), None)
        return _array(kind.value, self._shape[0] * self._shape[1] * self._shape[2], np.uint8)

    def probe_found(self, p: int) -> bool:
        found = ctypes.c_int()
        if _lib().as_setup_probe(self._h, int(p), ctypes.byref(found), None) != 0:
            raise IndexError(p)
        return bool(found.value)

    def probe_in_sight(self, p: int) -> bool:
        sight = ctypes.c_int()
        if _lib().as_setup_probe(self._h, int(p), None, ctypes.byref(sight)) != 0:
            raise IndexError(p)
        return bool(sight.value)

    def layout(self) -> dict:
        """The layout's arrays by the reference's names and types (fdtd.hpp:727-782)."""
        if self._layout is not None:
            return {k: v.copy() for k, v in self._layout.items()}
        L = native.Layout()
        if _lib().as_setup_layout(self._h, ctypes.byref(L)) != 0:
            raise RuntimeError("as_setup_layout failed")
        n = L.nx * L.ny * L.nz
        u64, f32 = np.uint64, np.float32
        out = {
            "shape": np.array([L.nx, L.ny, L.nz], u64),
            "stencil": np.array([L.stencil], u64),
            "weights": _array(L.weights, L.n_weights, f32),
            "centre": np.array([L.centre], f32),
            "total": np.array([L.total], f32),
            "scaled": np.array([1 if L.scale else 0], u64),
            "scale": _array(L.scale, n if L.scale else 0, f32),
            "layer_scale": _array(L.layer_scale, L.layer_cells if L.layer_scale else 0, f32),
            "types": _array(L.types, n, np.uint8),
            "special": _array(L.special, L.specials, np.int64),
            "link_start": _array(L.link_start, L.specials + 1, u64),
            "link_direction": _array(L.link_direction, L.links, np.uint8),
            "link_weight": _array(L.link_weight, L.links, f32),
            "coef_a": _array(L.coef_a, L.specials, f32),
            "coef_b": _array(L.coef_b, L.specials, f32),
            "coef_inv": _array(L.coef_inv, L.specials, f32),
            "source": _array(L.source, L.specials, f32),
            "state_start": _array(L.state_start, L.specials + 1, u64),
            "state_branch": _array(L.state_branch, L.states, np.int32),
            "state_weight": _array(L.state_weight, L.states, f32),
            "branch_now": _array(L.branch_now, L.branches, f32),
            "branch_before": _array(L.branch_before, L.branches, f32),
            "branch_gain": _array(L.branch_gain, L.branches, f32),
            "layer_cells": _array(L.layer_cell, L.layer_cells, np.int64),
            "flow_x": _array(L.flow_x, n if L.flow_x else 0, f32),
            "flow_y": _array(L.flow_y, n if L.flow_y else 0, f32),
            "flow_mask": _array(L.flow_mask, n if L.flow_mask else 0, np.uint8),
        }
        for d, ax in enumerate("xyz"):
            out[f"wind_{ax}"] = _array(L.wind[d], n if L.wind[d] else 0, f32)
            out[f"loop_{ax}"] = _array(L.loop[d], n if L.loop[d] else 0, f32)
        lengths = (L.nx + 1, L.ny + 1, L.nz + 1)
        for a, ax in enumerate("xyz"):
            for j, face in enumerate(("", "face_")):
                m = lengths[a] if L.layer_cells else 0
                out[f"keep_{face}{ax}"] = _array(L.keep[2 * a + j], m, f32)
                out[f"push_{face}{ax}"] = _array(L.push[2 * a + j], m, f32)
        out["probe_start"] = _array(L.probe_start, L.probes + 1, u64)
        out["probe_cells"] = _array(L.probe_cells, L.taps, u64)
        out["probe_weights"] = _array(L.probe_weights, L.taps, np.float64)
        self._layout = out
        return {k: v.copy() for k, v in out.items()}

    # Stepping, on the GPU (the reference's run/reset, fdtd.hpp:786-915).
    def _gpu(self):
        if self._stepper is None:
            from acoustic_spaces_amd.gpu import HipWaveSolver

            self._stepper = HipWaveSolver(self.layout())
        return self._stepper

    def run(self, signal) -> np.ndarray:
        return self._gpu().run(signal)

    def reset(self) -> None:
        if self._stepper is not None:
            self._stepper.reset()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _setup_lib is not None:
            _setup_lib.as_setup_destroy(h)
            self._h = None
