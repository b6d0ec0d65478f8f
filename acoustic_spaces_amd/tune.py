"""The stepper's autotuner: tile depth, steps per graph, where the face specials are
stepped, the steps a launch takes and the general specials' order, chosen by timing them.

For a layout it times each candidate launch shape (tile_z planes per tile, graph_steps
steps per captured graph, face specials by blocks of their own or in their tiles, one step
a launch or two steps by two launches with boxes of 8 or 16 planes, `pairs`) over a short
run on the GPU, with hipEvents on the stepper's stream, and keeps the fastest. Results
are cached by the layout's signature (its grid, cell counts and stencil) and the library's
build (a hash of libasfdtd.so) in a JSON file, as the reference keeps its own device timings
(engine/tuning.py:76-180); the reference tunes only the CPU/GPU choice, this tunes the
kernels' shape.
"""

from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path

import numpy as np

# In order of preference: the first is the library's default on a large grid. A candidate
# within TIE of the fastest counts as fast as it, so that the choice holds from run to run.
# Tiles of 16 planes re-read less of the planes above and below a tile; they pay off where
# tiles are mostly interior (outdoors, 1.5 %), and lose in a room with walls all round.
ONE_STEP = tuple({"tile_z": tz, "graph_steps": gs, "face_blocks": fb, "pairs": 0}
                 for tz in (8, 4, 16) for gs in (64, 128, 32) for fb in (True, False))
# Two steps by two launches: the first steps the cells whose neighbours are all interior
# twice, in boxes marching 8 or 16 planes in registers (float, 7 points, one speed of sound).
TWO_STEPS = tuple({"tile_z": 8, "graph_steps": 64, "face_blocks": fb, "pairs": planes}
                  for planes in (8, 16) for fb in (True, False))
# The general specials' order (the library's default, 3: by the tile holding them), against
# by their counts of links and states.
ORDERS = tuple({"tile_z": 8, "graph_steps": 64, "face_blocks": True, "pairs": 0, "special_order": order}
               for order in (1,))
CANDIDATES = ONE_STEP + TWO_STEPS + ORDERS
TIE = 0.01


def applicable(layout: dict, candidates, precision: int = 32) -> tuple:
    """The candidates the layout can take: two steps by two launches only for float grids
    of 7 points with one speed of sound; the specials' order only where general specials
    (links of their own weights or diagonal, as fitted walls make) are many."""
    pairs_ok = precision == 32 and int(layout["stencil"][0]) == 7 and not int(layout["scaled"][0])
    weights = np.asarray(layout["link_weight"])
    own = int(np.count_nonzero((np.asarray(layout["link_direction"]) >= 6) | (weights != weights.max(initial=0))))
    return tuple(c for c in candidates
                 if (pairs_ok or not c.get("pairs")) and (own >= 1000 or not c.get("special_order")))


def cache_path() -> Path:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return Path(base) / "acoustic_spaces_amd" / "tuning.json"


def signature(layout: dict) -> str:
    shape = "x".join(str(int(v)) for v in layout["shape"])
    air = int(np.count_nonzero(np.asarray(layout["types"])))
    return (f"{shape}|air{air}|sp{len(layout['special'])}|st{len(layout['state_branch'])}|"
            f"ly{len(layout['layer_cells'])}|s{int(layout['stencil'][0])}|sc{int(layout['scaled'][0])}")


def library_tag() -> str:
    """The first 16 hex digits of the SHA-256 of the built library: a rebuilt library, whose
    kernels may rank the shapes otherwise, gets timings of its own."""
    from acoustic_spaces_amd import native

    digest = hashlib.sha256()
    with open(native.LIBRARY, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            digest.update(chunk)
    return digest.hexdigest()[:16]


def cache_key(layout: dict, precision: int = 32, live: bool = False, candidates=CANDIDATES) -> str:
    """The cache's key: the layout's signature, the precision, live or zero fields, the
    library's build and the candidates' keys."""
    keys = ",".join(sorted({k for c in candidates for k in c}))
    return f"{signature(layout)}|p{precision}{'|live' if live else ''}|lib{library_tag()}|{keys}"


def time_shape(layout: dict, steps: int = 256, warmup: int = 64, precision: int = 32, drive=None,
               **shape) -> float:
    """Microseconds per step of the layout under one launch shape, from rest, driven by
    ``drive`` (its first warmup + steps samples) or by zeros. (On zero fields the same
    steps run a few per cent faster than on live ones; compare shapes on one drive.)"""
    from acoustic_spaces_amd import hip
    from acoustic_spaces_amd.gpu import HipWaveSolver

    signal = np.zeros(warmup + steps, np.float32)
    if drive is not None:
        d = np.asarray(drive, np.float32).reshape(-1)[: warmup + steps]
        signal[: len(d)] = d
    stepper = HipWaveSolver(layout, precision=precision, **shape)
    try:
        drive = hip.DeviceArray(signal)
        out = hip.DeviceArray(np.zeros((max(stepper.probes, 1), warmup + steps), np.float64))
        stepper.run_device(drive.ptr, warmup, out.ptr)
        hip.synchronize()
        start, stop = hip.Event(), hip.Event()
        start.record(stepper.stream)
        stepper.run_device(drive.ptr + 4 * warmup, steps, out.ptr + 8 * max(stepper.probes, 1) * warmup)
        stop.record(stepper.stream)
        hip.synchronize()
        return start.elapsed_ms(stop) * 1e3 / steps
    finally:
        stepper.close()


def autotune(layout: dict, steps: int = 256, candidates=CANDIDATES, use_cache: bool = True,
             precision: int = 32, drive=None) -> dict:
    """The fastest launch shape for the layout (the first candidate within TIE of the
    fastest): the candidate's keys (tile_z, graph_steps, face_blocks, pairs, and where it
    was timed special_order), "us_per_step" and
    "timings"; from the cache when it holds this layout's
    signature. With ``drive`` each candidate is timed on live fields (after 256 steps of it),
    which can rank shapes otherwise than zero fields do."""
    candidates = applicable(layout, candidates, precision)
    key = cache_key(layout, precision, drive is not None, candidates)
    path = cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]
    timings = []
    for shape in candidates:
        us = time_shape(layout, steps=max(steps, shape["graph_steps"]), warmup=64 if drive is None else 256,
                        precision=precision, drive=drive, **shape)
        timings.append({**shape, "us_per_step": round(us, 3)})
    fastest = min(t["us_per_step"] for t in timings)
    best = next(t for t in timings if t["us_per_step"] <= fastest * (1.0 + TIE))
    result = {**best, "timings": timings}
    if use_cache:
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            cache[key] = result
            path.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return result
