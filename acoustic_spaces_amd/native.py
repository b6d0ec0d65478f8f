"""ctypes binding of libasfdtd.so (include/asfdtd.h).

The library is built in-tree by ``acoustic_spaces_amd.build`` (or __graft_entry__.build());
loading it when it is missing raises, loudly: there is no other path to the kernels.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIBRARY = Path(__file__).resolve().parent / "libasfdtd.so"

P = ctypes.c_void_p
U64 = ctypes.c_uint64
SIZE = ctypes.c_size_t

AS_OK, AS_EINVAL, AS_EDEVICE, AS_ERANGE = 0, -1, -2, -3


class Layout(ctypes.Structure):
    """as_layout (include/asfdtd.h)."""

    _fields_ = [
        ("nx", U64), ("ny", U64), ("nz", U64), ("stencil", U64),
        ("weights", P), ("n_weights", U64), ("centre", ctypes.c_float), ("total", ctypes.c_float),
        ("scale", P), ("layer_scale", P), ("types", P),
        ("specials", U64), ("special", P), ("link_start", P), ("links", U64),
        ("link_direction", P), ("link_weight", P),
        ("coef_a", P), ("coef_b", P), ("coef_inv", P), ("source", P),
        ("state_start", P), ("states", U64), ("state_branch", P), ("state_weight", P),
        ("branches", U64), ("branch_now", P), ("branch_before", P), ("branch_gain", P),
        ("layer_cells", U64), ("layer_cell", P), ("keep", P * 6), ("push", P * 6),
        ("probes", U64), ("probe_start", P), ("taps", U64), ("probe_cells", P),
        ("probe_weights", P), ("coef_b64", P), ("coef_inv64", P), ("branch_gain64", P),
        ("flow_x", P), ("flow_y", P), ("flow_mask", P), ("wind", P * 3), ("loop", P * 3),
    ]


class Tuning(ctypes.Structure):
    """as_tuning (include/asfdtd.h)."""

    _fields_ = [("tile_y", ctypes.c_int), ("tile_z", ctypes.c_int), ("row_align", ctypes.c_int),
                ("graph_steps", ctypes.c_int), ("split", ctypes.c_int),
                ("tile_order", ctypes.c_int), ("no_faces", ctypes.c_int),
                ("face_blocks", ctypes.c_int), ("pairs", ctypes.c_int), ("keep_history", ctypes.c_int),
                ("special_order", ctypes.c_int), ("special_kinds", ctypes.c_int)]


_lib = None

#: Every symbol include/asfdtd.h declares.
SYMBOLS = (
    "as_fdtd_create", "as_fdtd_create_error", "as_fdtd_run", "as_fdtd_run_device", "as_fdtd_run_static",
    "as_fdtd_reset", "as_fdtd_field", "as_fdtd_stream", "as_fdtd_info", "as_fdtd_last_error",
    "as_fdtd_destroy", "as_fdtd_gradient", "as_fdtd_record", "as_fdtd_adjoint", "as_fdtd_mean_square",
)

#: The device post-processing's symbols (acoustic_spaces_amd.device_responses binds them).
RESPONSE_SYMBOLS = ("as_first_order_device",)

#: The set-up's symbols (acoustic_spaces_amd.solver binds them).
SETUP_SYMBOLS = ("as_setup_create", "as_setup_create_ex", "as_setup_create_error", "as_setup_layout", "as_setup_info",
                 "as_setup_wind", "as_setup_probe", "as_setup_destroy")


def load():
    """The library, its functions typed; raises when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise RuntimeError(
            f"the HIP wave-solver library {LIBRARY} is not built: run "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or python -m acoustic_spaces_amd.build)"
        )
    # One HIP runtime a process: torch carries its own libamdhip64. Imported first, it is the
    # one the library binds to; loaded after the library's (/opt/rocm), a second runtime would
    # claim the device beside the first, and torch's then finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(LIBRARY))
    lib.as_fdtd_create.restype = P
    lib.as_fdtd_create.argtypes = [ctypes.POINTER(Layout), ctypes.c_int, ctypes.POINTER(Tuning)]
    lib.as_fdtd_create_error.restype = ctypes.c_char_p
    lib.as_fdtd_create_error.argtypes = []
    lib.as_fdtd_run.restype = ctypes.c_int
    lib.as_fdtd_run.argtypes = [P, P, SIZE, P]
    lib.as_fdtd_run_device.restype = ctypes.c_int
    lib.as_fdtd_run_device.argtypes = [P, P, SIZE, P]
    lib.as_fdtd_run_static.restype = ctypes.c_int
    lib.as_fdtd_run_static.argtypes = [P, P, SIZE, P]
    lib.as_fdtd_reset.restype = ctypes.c_int
    lib.as_fdtd_reset.argtypes = [P]
    lib.as_fdtd_mean_square.restype = ctypes.c_int
    lib.as_fdtd_mean_square.argtypes = [P, P]
    lib.as_fdtd_field.restype = ctypes.c_int
    lib.as_fdtd_field.argtypes = [P, P]
    lib.as_fdtd_stream.restype = P
    lib.as_fdtd_stream.argtypes = [P]
    lib.as_fdtd_info.restype = ctypes.c_int
    lib.as_fdtd_info.argtypes = [P, P]
    lib.as_fdtd_last_error.restype = ctypes.c_char_p
    lib.as_fdtd_last_error.argtypes = [P]
    lib.as_fdtd_destroy.restype = None
    lib.as_fdtd_destroy.argtypes = [P]
    lib.as_fdtd_gradient.restype = ctypes.c_int
    lib.as_fdtd_gradient.argtypes = [P, P, SIZE, SIZE, P, P, P, P]
    lib.as_fdtd_record.restype = ctypes.c_int
    lib.as_fdtd_record.argtypes = [P, P, SIZE, SIZE, P]
    lib.as_fdtd_adjoint.restype = ctypes.c_int
    lib.as_fdtd_adjoint.argtypes = [P, P, P, P]
    lib.as_first_order_device.restype = ctypes.c_int
    lib.as_first_order_device.argtypes = [P, SIZE, SIZE, P, P, P, P]
    _lib = lib
    return lib


def raise_for(code: int, message: str):
    """The reference's exceptions for the library's codes: ValueError for a refused
    argument or range (engine/gpu.py:440-441, 495-504), RuntimeError for the device."""
    if code == AS_OK:
        return
    if code in (AS_EINVAL, AS_ERANGE):
        raise ValueError(message)
    raise RuntimeError(message)


_AXES = ("x", "y", "z")

_TYPES = {
    "weights": np.float32, "scale": np.float32, "layer_scale": np.float32, "types": np.uint8,
    "special": np.int64, "link_start": np.uint64, "link_direction": np.uint8,
    "link_weight": np.float32, "coef_a": np.float32, "coef_b": np.float32,
    "coef_inv": np.float32, "source": np.float32, "state_start": np.uint64,
    "state_branch": np.int32, "state_weight": np.float32, "branch_now": np.float32,
    "branch_before": np.float32, "branch_gain": np.float32, "layer_cells": np.int64,
    "probe_start": np.uint64, "probe_cells": np.uint64, "probe_weights": np.float64,
}


def pack_layout(layout: dict):
    """An as_layout over the arrays of a layout dict (WaveSolver::layout's names,
    reference fdtd.hpp:727-782); returns (struct, arrays kept alive)."""
    A = {}
    for name, dt in _TYPES.items():
        A[name] = np.ascontiguousarray(layout.get(name, np.zeros(0, dt)), dtype=dt)
    for k in ("keep", "push"):
        for ax in _AXES:
            for face in ("", "face_"):
                key = f"{k}_{face}{ax}"
                A[key] = np.ascontiguousarray(layout.get(key, np.zeros(0)), dtype=np.float32)
    shape = [int(v) for v in layout["shape"]]
    n = shape[0] * shape[1] * shape[2]
    if len(A["types"]) != n:
        raise ValueError("the layout's cell types do not cover its grid")
    s = Layout()
    s.nx, s.ny, s.nz = shape
    s.stencil = int(np.asarray(layout["stencil"]).reshape(-1)[0])
    s.weights = A["weights"].ctypes.data
    s.n_weights = len(A["weights"])
    s.centre = float(np.asarray(layout["centre"]).reshape(-1)[0])
    s.total = float(np.asarray(layout["total"]).reshape(-1)[0])
    scaled = int(np.asarray(layout.get("scaled", [0])).reshape(-1)[0])
    if scaled and len(A["scale"]) != n:
        raise ValueError("the layout's speeds do not cover its grid")
    s.scale = A["scale"].ctypes.data if scaled else None
    s.layer_scale = A["layer_scale"].ctypes.data if len(A["layer_scale"]) else None
    s.types = A["types"].ctypes.data
    s.specials = len(A["special"])
    s.links = len(A["link_direction"])
    s.states = len(A["state_branch"])
    s.branches = len(A["branch_now"])
    s.layer_cells = len(A["layer_cells"])
    s.probes = len(A["probe_start"]) - 1 if len(A["probe_start"]) else 0
    s.taps = len(A["probe_cells"])
    if len(A["link_start"]) != s.specials + 1 or len(A["state_start"]) != s.specials + 1:
        raise ValueError("the layout's link and state ranges do not match its special cells")
    # Every array against the count it is indexed by, so that the library never reads past
    # the end of a host array.
    counts = {"coef_a": s.specials, "coef_b": s.specials, "coef_inv": s.specials, "source": s.specials,
              "link_weight": s.links, "state_weight": s.states, "branch_before": s.branches,
              "branch_gain": s.branches, "probe_weights": s.taps}
    for name, count in counts.items():
        if len(A[name]) != count:
            raise ValueError(f"the layout's {name} has {len(A[name])} values, not {count}")
    if s.specials and (int(A["link_start"][-1]) != s.links or int(A["state_start"][-1]) != s.states):
        raise ValueError("the layout's link and state ranges do not end at its links and states")
    if len(A["weights"]) != (6 if s.stencil == 7 else 18):
        raise ValueError(f"the layout's weights have {len(A['weights'])} values for a {s.stencil}-point stencil")
    for name in ("special", "link_start", "link_direction", "link_weight", "coef_a", "coef_b",
                 "coef_inv", "source", "state_start", "state_branch", "state_weight",
                 "branch_now", "branch_before", "branch_gain", "probe_start", "probe_cells",
                 "probe_weights"):
        setattr(s, name, A[name].ctypes.data if len(A[name]) else None)
    s.layer_cell = A["layer_cells"].ctypes.data if len(A["layer_cells"]) else None
    for a, ax in enumerate(_AXES):
        for j, face in enumerate(("", "face_")):
            keep, push = A[f"keep_{face}{ax}"], A[f"push_{face}{ax}"]
            if s.layer_cells and (len(keep) != shape[a] + 1 or len(push) != shape[a] + 1):
                raise ValueError("the matched layer's coefficients do not cover the grid")
            s.keep[2 * a + j] = keep.ctypes.data if len(keep) else None
            s.push[2 * a + j] = push.ctypes.data if len(push) else None
    # Moving air: plain (flow_x, flow_y and the row masks) or round (each face's wind and the
    # loops, the mask 1 or 2), each over the whole grid.
    A["flow_mask"] = np.ascontiguousarray(layout.get("flow_mask", np.zeros(0, np.uint8)), dtype=np.uint8)
    for name in ("flow_x", "flow_y", *(f"{k}_{ax}" for k in ("wind", "loop") for ax in _AXES)):
        A[name] = np.ascontiguousarray(layout.get(name, np.zeros(0, np.float32)), dtype=np.float32)
    flowing = len(A["flow_mask"]) > 0
    plain = [A[k] for k in ("flow_x", "flow_y")]
    round_ = [A[f"{k}_{ax}"] for k in ("wind", "loop") for ax in _AXES]
    if flowing:
        if len(A["flow_mask"]) != n:
            raise ValueError("the moving air's mask does not cover the grid")
        s.flow_mask = A["flow_mask"].ctypes.data
        if all(len(a) == n for a in plain) and all(len(a) == 0 for a in round_):
            s.flow_x, s.flow_y = A["flow_x"].ctypes.data, A["flow_y"].ctypes.data
        elif all(len(a) == n for a in round_) and all(len(a) == 0 for a in plain):
            for d, ax in enumerate(_AXES):
                s.wind[d] = A[f"wind_{ax}"].ctypes.data
                s.loop[d] = A[f"loop_{ax}"].ctypes.data
        else:
            raise ValueError("moving air takes the shifts (flow_x, flow_y) or the wind and loops over the "
                             "whole grid, not both")
    elif any(len(a) for a in plain + round_):
        raise ValueError("moving air without its mask (flow_mask)")
    # Optional double coefficients (precision 64): coef_b64, coef_inv64, branch_gain64.
    for name, count in (("coef_b64", s.specials), ("coef_inv64", s.specials), ("branch_gain64", s.branches)):
        if name in layout:
            arr = np.ascontiguousarray(layout[name], dtype=np.float64)
            if len(arr) != count:
                raise ValueError(f"{name} has {len(arr)} values, not {count}")
            A[name] = arr
            setattr(s, name, arr.ctypes.data if len(arr) else None)
    return s, A
