"""Many rooms in one launch a step: the batch's layouts merged into one grid (or a few).

Rooms with one cell size and stencil (one time step) are stacked along z into one grid,
each padded in x and y to the batch's largest room. Every room keeps its own outermost
layer of cells, which is never air (fdtd.hpp:340-342), so no update of one room reads a
cell of another: the merged grid steps each room exactly as the room alone, to the bit.
One fused launch a step then carries every room's interior tiles, special cells and
probes, and the stepper's graphs replay it; the tiles list holds only tiles with air, so
the padding costs memory, not bandwidth.

The batch shares one drive (one sample rate); each room's source cells carry their own
weights. Rooms with a matched layer are refused (the layer's coefficients follow each
room's own coordinates), as are rooms with a speed per cell unless every room has one.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from acoustic_spaces_amd.gpu import HipWaveSolver


@dataclasses.dataclass
class Placement:
    """Where room r lies in the merged grid: z offset, its shape, its probes' rows."""

    z0: int
    shape: tuple
    probes: slice


def merge_layouts(layouts):
    """One layout for a batch of layouts (see the module's notes); returns it and each
    room's Placement."""
    if not layouts:
        raise ValueError("a batch needs at least one room")
    first = layouts[0]
    stencil = int(first["stencil"][0])
    for L in layouts:
        if int(L["stencil"][0]) != stencil or not np.array_equal(L["weights"], first["weights"]):
            raise ValueError("a batch's rooms share one stencil")
        if len(L["layer_cells"]):
            raise ValueError("a batch's rooms have no matched layer")
        if any(len(L.get(k, ())) for k in ("flow_x", "flow_y", "flow_mask")):
            raise ValueError("a batch's rooms are in still air")
    scaled = [int(L["scaled"][0]) for L in layouts]
    if any(scaled) and not all(scaled):
        raise ValueError("either every room of a batch has a speed per cell or none has")
    shapes = [tuple(int(v) for v in L["shape"]) for L in layouts]
    NX = max(s[0] for s in shapes)
    NY = max(s[1] for s in shapes)
    NZ = sum(s[2] for s in shapes)
    plane = NX * NY
    types = np.zeros(plane * NZ, np.uint8)
    scale = np.ones(plane * NZ, np.float32) if all(scaled) else np.zeros(0, np.float32)
    parts = {k: [] for k in ("special", "link_direction", "link_weight", "coef_a", "coef_b", "coef_inv",
                             "source", "state_branch", "state_weight", "branch_now", "branch_before",
                             "branch_gain", "probe_cells", "probe_weights")}
    link_start, state_start, probe_start = [np.zeros(1, np.uint64)], [np.zeros(1, np.uint64)], [np.zeros(1, np.uint64)]
    links = states = taps = branches = probes = 0
    placements = []
    z0 = 0
    for L, (nx, ny, nz) in zip(layouts, shapes):
        idx = np.arange(nx * ny * nz, dtype=np.int64)
        x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        remap = x + NX * (y + NY * (z + z0))
        types[remap] = L["types"]
        if len(scale):
            scale[remap] = L["scale"]
        sp = np.asarray(L["special"], np.int64)
        parts["special"].append(remap[sp] if len(sp) else sp)
        for k in ("link_direction", "link_weight", "coef_a", "coef_b", "coef_inv", "source", "state_weight",
                  "branch_now", "branch_before", "branch_gain", "probe_weights"):
            parts[k].append(np.asarray(L[k]))
        parts["state_branch"].append(np.asarray(L["state_branch"], np.int32) + branches)
        pc = np.asarray(L["probe_cells"], np.int64)
        parts["probe_cells"].append(remap[pc].astype(np.uint64) if len(pc) else pc.astype(np.uint64))
        link_start.append(np.asarray(L["link_start"], np.uint64)[1:] + links)
        state_start.append(np.asarray(L["state_start"], np.uint64)[1:] + states)
        probe_start.append(np.asarray(L["probe_start"], np.uint64)[1:] + taps)
        p = len(L["probe_start"]) - 1
        placements.append(Placement(z0, (nx, ny, nz), slice(probes, probes + p)))
        links += len(L["link_direction"])
        states += len(L["state_branch"])
        taps += len(L["probe_cells"])
        branches += len(L["branch_now"])
        probes += p
        z0 += nz
    merged = {
        "shape": np.array([NX, NY, NZ], np.uint64),
        "stencil": np.array([stencil], np.uint64),
        "weights": np.asarray(first["weights"]).copy(),
        "centre": np.asarray(first["centre"]).copy(),
        "total": np.asarray(first["total"]).copy(),
        "scaled": np.array([1 if len(scale) else 0], np.uint64),
        "scale": scale,
        "layer_scale": np.zeros(0, np.float32),
        "types": types,
        "link_start": np.concatenate(link_start),
        "state_start": np.concatenate(state_start),
        "probe_start": np.concatenate(probe_start),
        "layer_cells": np.zeros(0, np.int64),
    }
    dtypes = {"special": np.int64, "link_direction": np.uint8, "link_weight": np.float32, "coef_a": np.float32,
              "coef_b": np.float32, "coef_inv": np.float32, "source": np.float32, "state_branch": np.int32,
              "state_weight": np.float32, "branch_now": np.float32, "branch_before": np.float32,
              "branch_gain": np.float32, "probe_cells": np.uint64, "probe_weights": np.float64}
    for k, dt in dtypes.items():
        merged[k] = np.concatenate([np.asarray(a, dt) for a in parts[k]]) if parts[k] else np.zeros(0, dt)
    return merged, placements


# Bytes a merged grid's cell costs on the device: two pressure fields and its class.
BYTES_PER_CELL = {32: 9, 64: 17}


def plan_groups(layouts, group_bytes=None, precision: int = 32):
    """The rooms (indices into ``layouts``) in groups to merge, each group's grid within
    ``group_bytes`` of device memory (None: one group). Rooms go largest footprint (x y)
    first, so that a group's rooms pad little to its largest; a room larger than the
    budget is a group of its own. The MI355X's 256 MB Infinity Cache keeps a grid of a
    fraction of it resident from one step to the next."""
    order = sorted(range(len(layouts)), key=lambda i: -int(layouts[i]["shape"][0]) * int(layouts[i]["shape"][1]))
    if group_bytes is None:
        return [order]
    per = BYTES_PER_CELL[precision]
    groups, current, nx, ny, nz = [], [], 0, 0, 0
    for i in order:
        sx, sy, sz = (int(v) for v in layouts[i]["shape"])
        gx, gy, gz = max(nx, sx), max(ny, sy), nz + sz
        if current and gx * gy * gz * per > group_bytes:
            groups.append(current)
            current, gx, gy, gz = [], sx, sy, sz
        current.append(i)
        nx, ny, nz = gx, gy, gz
    if current:
        groups.append(current)
    return groups


class BatchedWaveSolver:
    """Steps a batch of rooms (their solvers, or layouts) on the GPU: the rooms merged into
    one grid, or into groups of grids of at most ``group_bytes`` each (see plan_groups),
    one launch a step for each grid. ``run(signal)`` returns each room's probes (probes,
    steps), in the order given."""

    #: The grouping measured fastest on the 64 rooms at 5 cm (groups of 128 to 384 MB and one
    #: merged grid timed): grids of at most 192 MB, which the 256 MB Infinity Cache holds.
    DEFAULT_GROUP_BYTES = 192 << 20

    def __init__(self, rooms, precision: int = 32, group_bytes=None, **tuning):
        layouts = [r if isinstance(r, dict) else r.layout() for r in rooms]
        self.groups = []
        self.cells = self.air_cells = 0
        for idx in plan_groups(layouts, group_bytes, precision):
            merged, placements = merge_layouts([layouts[i] for i in idx])
            self.cells += int(np.prod(merged["shape"]))
            self.air_cells += int(np.count_nonzero(merged["types"]))
            self.groups.append((idx, placements, HipWaveSolver(merged, precision=precision, **tuning)))
        self.rooms = len(layouts)

    @property
    def steppers(self) -> list:
        return [g[2] for g in self.groups]

    @property
    def probes(self) -> int:
        return sum(s.probes for s in self.steppers)

    def run(self, signal) -> list:
        out = [None] * self.rooms
        for idx, placements, stepper in self.groups:
            y = stepper.run(signal)
            for i, p in zip(idx, placements):
                out[i] = y[p.probes]
        return out

    def reset(self) -> None:
        for s in self.steppers:
            s.reset()

    def close(self) -> None:
        for s in self.steppers:
            s.close()
