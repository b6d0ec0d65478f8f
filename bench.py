"""Throughput of the wave solver's time step on one MI355X.

A step is one time step of the FDTD scheme over the whole grid of the hall workload: the
30 x 20 x 12 m hall with a ceiling of six tilted planes and side walls splayed 5 degrees, at
10 cm cells (305 x 257 x 125 cells, about 7.57M air cells), fitted walls (each cell a face
cuts holds its own air volume, open areas and wall areas, diagonal links where the planes
are tilted), a constant admittance and three first-order admittance branches (each rising
or falling, fitted to the material's drawn absorption) per material, one source, 16
receivers; 2 s of response is 11,882 steps. The grid is laid
out by this package's fit (csrc/fitted.cpp, csrc/setup.cpp) and stepped by the HIP
kernels (csrc/fdtd_step.hip), in the launch shape the autotuner finds fastest on the
scene's drive (part of the set-up); the drive and the probes' output stay in device memory
over the timed region. ``--workload hall-faces`` times the same hall with its walls on cell
faces.

Printed: one JSON line with Mcell-updates/s (air cells x steps / s), the step kernel's
roofline (algorithmic bytes per step over its time per step, against HBM's 8 TB/s) and
the reference's own CPU solver (built from its headers into oracle/_ref) timed on this
host's cores on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
PER_AIR_CELL = 12  # bytes: read the current and previous pressure, write the next (float)
PER_LAYER_CELL = 56  # bytes: the current pressure, three parts and three faces read and written, the next
PER_STATE = 12  # bytes: a branch velocity read twice-over and written (float)


def algorithmic_bytes(info: dict, air_cells: int, layer_cells: int, states: int) -> int:
    """Bytes one step must move at least: 12 per air cell outside the matched layer, 56 per
    layer cell (each face velocity shared by the two cells it lies between is one value),
    12 per branch state (coefficient and index reads not counted)."""
    return PER_AIR_CELL * (air_cells - layer_cells) + PER_LAYER_CELL * layer_cells + PER_STATE * states


def octave_outputs(y: np.ndarray, rate: float) -> dict:
    """Each receiver's 63, 125 and 250 Hz octave-band strength G (dB) and T30 (s), of its
    signal high-passed as the reference's source signal is (the Gaussian drive carries air
    into the closed room, which would never decay)."""
    from acoustic_spaces_amd.analysis import octave_results
    from acoustic_spaces_amd.wave import high_pass

    rows = [octave_results(high_pass(r, rate), rate, bands=3) for r in y]
    return {"bands_hz": rows[0]["bands_hz"], "g_db": [r["g_db"] for r in rows], "t30_s": [r["t30_s"] for r in rows]}


def compare_outputs(a: dict, b: dict, ya: np.ndarray, yb: np.ndarray) -> dict:
    """The largest differences between two runs' receivers: relative L2 of the signals,
    G in dB and T30 in s."""
    rel = [float(np.linalg.norm(p - q) / np.linalg.norm(q)) for p, q in zip(ya, yb)]
    dg = np.abs(np.array(a["g_db"]) - np.array(b["g_db"]))
    dt = np.abs(np.array(a["t30_s"]) - np.array(b["t30_s"]))
    return {"max_rel_l2": max(rel), "max_abs_g_db": float(np.nanmax(dg)), "max_abs_t30_s": float(np.nanmax(dt))}


def insertion_loss(scene, layout) -> dict:
    """The barrier's insertion loss at each receiver in the 63 and 125 Hz octave bands: the
    band energy over the bare ground against that with the barrier (dB), from the whole
    drive in float; and its largest difference from the same in double."""
    import dataclasses

    from acoustic_spaces_amd import scenes as sc
    from acoustic_spaces_amd.analysis import band_energy
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    bare = WaveSolver.from_scene(sc.outdoor(barrier=False)).layout()
    slots = int(np.ceil(len(scene.drive) / scene.rate / 1e-3))

    def levels(y):
        return np.array([10 * np.log10(band_energy(r, scene.rate, 2, 1e-3, slots).sum(axis=1)) for r in y])

    out = {}
    for precision in (32, 64):
        with_barrier = levels(HipWaveSolver.run_once(layout, scene.drive, precision=precision))
        without = levels(HipWaveSolver.run_once(bare, scene.drive, precision=precision))
        out[precision] = without - with_barrier
    return {"bands_hz": [63, 125], "il_db_f32": np.round(out[32], 2).tolist(),
            "max_abs_il_db_f32_vs_f64": float(np.abs(out[32] - out[64]).max()),
            "receivers_m_behind_barrier": np.round(scene.probes[:, 0] - (scene.probes[:, 0].min() - 10.0), 1).tolist()}


def cpu_baseline(scene, target_s: float = 12.0, compare=None, others=()) -> dict:
    """The reference's CPU solver (oracle/_ref) on all host cores, on a bounded run of the
    same workload; the oracle's one-thread port if the reference's build is absent. With
    ``compare`` (the GPU's float probes over the scene's whole drive), the reference's own
    run of the whole drive too, and the differences of its signals, band levels and T30;
    the same for each (scene, GPU probes) of ``others`` (the hall's second source)."""
    cores = len(os.sched_getaffinity(0))
    from oracle import reference

    fitted = bool(getattr(scene, "fitted_walls", False))
    if reference.available():
        solver = reference.ReferenceSolver.from_scene(scene, threads=cores, fitted_walls=fitted)
        kind, used = "reference", solver.threads
        air = solver.air_cells
    else:
        from acoustic_spaces_amd.solver import WaveSolver
        from oracle.stepper import OracleSolver

        mine = WaveSolver.from_scene(scene, fitted_walls=fitted)
        solver = OracleSolver(mine.layout())
        kind, used = "port", 1
        air = mine.air_cells
    drive = scene.drive
    t = time.perf_counter()
    solver.run(drive[:8])
    first = max(time.perf_counter() - t, 1e-6) / 8
    steps = int(min(len(drive) - 8, max(16, target_s / first)))
    t = time.perf_counter()
    solver.run(drive[8:8 + steps])
    elapsed = time.perf_counter() - t
    out = {
        "value": round(air * steps / elapsed / 1e6, 1),
        "unit": "Mcell-updates/s",
        "cores": int(used),
        "kind": kind,
        "sample": f"{steps} steps of the same grid ({air:,} air cells) after 8, "
                  f"{elapsed:.1f} s, threads={used}",
    }
    if compare is not None:
        solver.reset()
        y = solver.run(drive[: compare.shape[1]])
        out["gpu_f32_vs_this"] = compare_outputs(octave_outputs(compare, scene.rate), octave_outputs(y, scene.rate),
                                                 compare, y)
    for k, (other, gpu) in enumerate(others):
        if kind != "reference":
            break
        ref = reference.ReferenceSolver.from_scene(other, threads=cores, fitted_walls=fitted)
        y = ref.run(other.drive[: gpu.shape[1]])
        out[f"gpu_f32_vs_this_source{k + 2}"] = compare_outputs(octave_outputs(gpu, other.rate),
                                                                 octave_outputs(y, other.rate), gpu, y)
    return out


def calibrate() -> int:
    """hipMemcpy device to device (the runtime's copy kernel) of 1 GiB, 96 MiB and 32 MiB,
    20 times each after a warm-up: GB/s counted as bytes read plus bytes written."""
    import ctypes

    from acoustic_spaces_amd import hip

    rt = hip.runtime()
    rows = []
    for mib in (1024, 96, 32):
        n = mib << 20
        a = hip.DeviceArray(np.zeros(n // 4, np.float32))
        b = hip.DeviceArray(np.zeros(n // 4, np.float32))
        rt.hipMemcpy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), n, 3)
        hip.synchronize()
        start, stop = hip.Event(), hip.Event()
        start.record(0)
        for _ in range(20):
            rt.hipMemcpy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), n, 3)
        stop.record(0)
        ms = start.elapsed_ms(stop) / 20
        rows.append({"bytes_copied": n, "us_per_copy": round(ms * 1e3, 1),
                     "GBps_read_plus_write": round(2 * n / (ms * 1e-3) / 1e9, 1)})
        a.free()
        b.free()
    print(json.dumps({"calibration": "hipMemcpy device to device", "runs": rows}), flush=True)
    return 0


def time_steps(stepper, drive: np.ndarray, warmup: int, steps: int, lead: int = 0):
    """A lead-in of ``lead`` steps from rest (so that the pulse has spread and the timed
    steps meet live fields whatever ``warmup`` is), the warm-up, then exactly ``steps``
    steps between synchronisations, timed with hipEvents on the stepper's own stream; the
    drive and the probes' output in device memory. Returns (milliseconds, the probes'
    output of all three, (probes, lead + warmup + steps))."""
    from acoustic_spaces_amd import hip

    need = lead + warmup + steps
    probes = max(stepper.probes, 1)
    d_drive = hip.DeviceArray(np.ascontiguousarray(drive[:need], np.float32))
    d_out = hip.DeviceArray(np.zeros((probes, need), np.float64))
    stream = stepper.stream
    if lead:
        stepper.run_device(d_drive.ptr, lead, d_out.ptr)
    if warmup:
        stepper.run_device(d_drive.ptr + 4 * lead, warmup, d_out.ptr + 8 * probes * lead)
    hip.synchronize()
    start, stop = hip.Event(), hip.Event()
    start.record(stream)
    stepper.run_device(d_drive.ptr + 4 * (lead + warmup), steps, d_out.ptr + 8 * probes * (lead + warmup))
    stop.record(stream)
    hip.synchronize()
    ms = start.elapsed_ms(stop)
    # Each call wrote its own (probes, its steps) block, in the order of the calls.
    flat = d_out.to_host().reshape(-1)
    blocks, at = [], 0
    for n in (lead, warmup, steps):
        blocks.append(flat[at: at + probes * n].reshape(probes, n))
        at += probes * n
    d_drive.free()
    d_out.free()
    return ms, np.concatenate(blocks, axis=1)


def graph_length(tuned: int, steps: int) -> int:
    """The steps a graph replays for a timed run of ``steps``: the tuned length where the
    run holds at least one graph, else the largest divisor of ``steps`` not above it, so
    that a short timed run still replays graphs, as a long one does."""
    if tuned <= 0 or steps >= tuned:
        return tuned
    return max(d for d in range(1, steps + 1) if steps % d == 0 and d <= tuned)


def launch_mode(graph_steps: int, steps: int) -> str:
    if graph_steps <= 0:
        return f"{steps} launches from the host, one a step"
    whole, rest = divmod(steps, graph_steps)
    tail = f" and {rest} launches from the host" if rest else ""
    return f"{whole} replays of a captured graph of {graph_steps} steps{tail}"


def measured_traffic(name: str):
    """HBM bytes per step from the rocprofv3 FETCH_SIZE and WRITE_SIZE runs summarised in
    profiles/<name>.json, each counter times the scale a copy of known size gives it
    (FETCH_SIZE 2, WRITE_SIZE 1), or None."""
    path = os.path.join(ROOT, "profiles", f"{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return int(t["fetch_kib"] * 1024 * t["fetch_scale"] + t["write_kib"] * 1024 * t["write_scale"])


def line_for(metric, value, K, W, ms, workload, config, roofline, dtype="f32"):
    return {
        "metric": metric,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": workload, **config},
        "roofline": roofline,
    }


def roofline_of(step_bytes, per_step_s, traffic, kernel):
    achieved = step_bytes / per_step_s / 1e9
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "bytes_per_step": step_bytes,
        "kernel": kernel,
    }


def plain(scene, name, args, traffic_name=None):
    """One plain solve of ``scene``: set-up, then K timed steps."""
    from acoustic_spaces_amd import scenes as sc
    from acoustic_spaces_amd import tune
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    K, W, lead = args.steps, args.warmup, args.lead_in
    need = lead + W + K
    drive = scene.drive if need <= len(scene.drive) else sc.gaussian_drive(scene.rate, need)
    # The set-up on the GPU (its passes over the whole grid on the device), then the host
    # set-up alone for comparison; the two layouts must be the same. The runtime is brought
    # up first, as it is in a process that steps.
    import torch

    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    fitted = bool(getattr(scene, "fitted_walls", False))
    host_threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    if fitted:
        # The fit runs on the host's threads.
        t = time.perf_counter()
        solver = WaveSolver.from_scene(scene, fitted_walls=True)
        layout = solver.layout()
        setup = {"host": round(time.perf_counter() - t, 3), "host_threads": host_threads,
                 "host_is": "the fit of the mesh to the grid (csrc/fitted.cpp) and the special cells' coefficients"}
    else:
        t = time.perf_counter()
        solver = WaveSolver.from_scene(scene, on_device=True)
        layout = solver.layout()
        setup_s = time.perf_counter() - t
        t = time.perf_counter()
        host_layout = WaveSolver.from_scene(scene).layout()
        host_setup_s = time.perf_counter() - t
        if sorted(host_layout) != sorted(layout) or not all(np.array_equal(layout[k], host_layout[k]) for k in layout):
            raise SystemExit("the device set-up's layout differs from the host's")
        del host_layout
        setup = {"device": round(setup_s, 3), "host": round(host_setup_s, 3),
                 "device_is": "every pass over the whole grid on the GPU (cut links, flood, layer, wall cells); the source's and probes' readings on the host",
                 "host_threads": host_threads}
    if args.tile_z or args.graph_steps or args.face_blocks >= 0 or args.special_order or args.special_kinds:
        launch = {"tile_z": args.tile_z, "graph_steps": args.graph_steps,
                  "face_blocks": None if args.face_blocks < 0 else bool(args.face_blocks), "pairs": 0,
                  "special_order": args.special_order, "special_kinds": args.special_kinds}
        launch_from = "flags (0 or null: the library's default)"
    else:
        # Set-up, outside the timed region: the launch shape timed on the scene's drive.
        shape = tune.autotune(layout, steps=512, drive=drive)
        launch = {k: shape.get(k, 0) for k in ("tile_z", "graph_steps", "face_blocks", "pairs", "special_order")}
        launch_from = "the autotuner, each candidate timed on the scene's drive (live fields)"
    # A timed run shorter than the tuned graph replays graphs of a length that divides it
    # (0 is the library's default, 64).
    launch["graph_steps"] = graph_length(launch["graph_steps"] or 64, K)
    stepper = HipWaveSolver(layout, split=args.split, **launch)
    info = stepper.info()
    ms, out = time_steps(stepper, drive, W, K, lead=lead)
    if not np.all(np.isfinite(out)):
        raise SystemExit("non-finite probe output")
    air = solver.air_cells
    per_step_s = ms / 1e3 / K
    step_bytes = algorithmic_bytes(info, air, solver.layer_cells, solver.branch_states)
    config = {
        "air_cells": air, "layer_cells": solver.layer_cells, "branch_states": solver.branch_states,
        "special_cells": info["specials"], "special_kinds": info["special_kinds"], "grid": list(scene.grid.shape),
        "setup_s": setup,
        "launch": launch, "launch_from": launch_from,
        "timed_launches": launch_mode(launch["graph_steps"], K) + (
            f", two steps by two launches (boxes of {launch['pairs'] if launch['pairs'] > 1 else 8} planes)"
            if launch.get("pairs") else ", one step a launch"),
        "lead_in_steps": lead,
        "lead_in": "steps of the drive from rest before the warm-up, so that the timed steps "
                   "meet the spreading pulse (live fields), not fields at rest",
        "precision": ("float: the scheme bit-identical to the reference's CPU scheme on one layout; this "
                      "package's fitted layout has the reference's cell kinds, specials, links and states, its "
                      "coefficients within a float of the reference's" if fitted else
                      "float (bit-identical to the reference's CPU scheme)"),
    }
    traffic = measured_traffic(traffic_name) if traffic_name else None
    stepper.close()
    if args.outputs:
        # The receivers' outputs over the whole drive, in float and in double.
        y32 = HipWaveSolver.run_once(layout, scene.drive, precision=32)
        y64 = HipWaveSolver.run_once(layout, scene.drive, precision=64)
        o32, o64 = octave_outputs(y32, scene.rate), octave_outputs(y64, scene.rate)
        config["outputs"] = {"steps": len(scene.drive), "f32": o32,
                             "f32_vs_f64": compare_outputs(o32, o64, y32, y64)}
        args.signals = y32
    if args.insertion_loss:
        config["insertion_loss"] = insertion_loss(scene, layout)
    return line_for(f"FDTD Mcell-updates/s ({name}, plain solver)", air / per_step_s / 1e6, K, W, ms, scene.name,
                    config, roofline_of(step_bytes, per_step_s, traffic,
                                        f"pair_kernel + {step_kernel_name(info)} (two steps by two launches)"
                                        if launch.get("pairs") else
                                        f"{step_kernel_name(info)} (one fused launch per step)"))


def step_kernel_name(info: dict) -> str:
    """The fused step kernel a float grid of 7 points with one speed of sound launches: its
    last argument says whether the general specials read their kinds' tables."""
    return f"step_kernel<float,false,false,{'true' if info.get('special_kinds') else 'false'}>"


def source_outputs(scene, args) -> dict:
    """The receivers' outputs of ``scene`` over its whole drive, in float and in double (the
    float probes kept in args.others for the CPU leg's comparison)."""
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    layout = WaveSolver.from_scene(scene, fitted_walls=bool(getattr(scene, "fitted_walls", False))).layout()
    y32 = HipWaveSolver.run_once(layout, scene.drive, precision=32)
    y64 = HipWaveSolver.run_once(layout, scene.drive, precision=64)
    o32, o64 = octave_outputs(y32, scene.rate), octave_outputs(y64, scene.rate)
    args.others.append((scene, y32))
    return {"steps": len(scene.drive), "f32": o32, "f32_vs_f64": compare_outputs(o32, o64, y32, y64)}


def rooms(args):
    """64 rooms: batched solves (the rooms merged into one grid, one launch a step, its
    launch shape from the autotuner; and merged into groups of grids sized to stay in the
    256 MB Infinity Cache, one launch a step each) against 64 plain solves one after
    another, the same steps each. The value is the fastest batched layout's."""
    from acoustic_spaces_amd import scenes, tune
    from acoustic_spaces_amd.batch import BatchedWaveSolver, merge_layouts
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    K, W = args.steps, args.warmup
    batch = scenes.rooms_batch(count=64)
    t = time.perf_counter()
    layouts = [WaveSolver.from_scene(r).layout() for r in batch]
    setup_s = time.perf_counter() - t
    drive = scenes.gaussian_drive(batch[0].rate, W + K)
    merged, placements = merge_layouts(layouts)
    shape = tune.autotune(merged, steps=128, drive=drive)
    launch = {k: shape.get(k, 0) for k in ("tile_z", "graph_steps", "face_blocks", "pairs")}
    stepper = HipWaveSolver(merged, **launch)
    info = stepper.info()
    ms_one, out = time_steps(stepper, drive, W, K)
    stepper.close()
    if not np.all(np.isfinite(out)):
        raise SystemExit("non-finite probe output")
    merged_shape = [int(v) for v in merged["shape"]]
    del merged
    grouped, batched_out = {}, {}
    for mb in (384, 256, 192, 128):
        b = BatchedWaveSolver(layouts, group_bytes=mb << 20, **launch)
        total, rooms_out = 0.0, [None] * len(layouts)
        for idx, placements, s in b.groups:
            ms, y = time_steps(s, drive, W, K)
            total += ms
            for i, p in zip(idx, placements):
                rooms_out[i] = y[p.probes]
        grouped[mb] = (total, len(b.groups))
        batched_out[mb] = rooms_out
        b.close()
    ms_seq, plain_out = 0.0, []
    for L in layouts:
        s = HipWaveSolver(L)
        ms, y = time_steps(s, drive, W, K)
        ms_seq += ms
        plain_out.append(y)
        s.close()
    # Each room's receivers, batched in every group size, against its plain solve.
    same = {f"{mb} MB": all(np.array_equal(a, b) for a, b in zip(batched_out[mb], plain_out)) for mb in batched_out}
    # The study's result: each band's T30 and G over the rooms' receivers.
    bands = [octave_outputs(y, batch[0].rate) for y in plain_out]
    study = {}
    for q, key in ((0, "t30_s"), (1, "g_db")):
        for b_i, hz in enumerate(bands[0]["bands_hz"]):
            v = np.array([row[b_i] for o in bands for row in o[key]], float)
            v = v[np.isfinite(v)]
            study[f"{key}_{hz}hz"] = {"receivers": int(v.size), "p10": round(float(np.percentile(v, 10)), 4),
                                      "median": round(float(np.median(v)), 4),
                                      "p90": round(float(np.percentile(v, 90)), 4)}
    best_mb = min(grouped, key=lambda k: grouped[k][0])
    ms_batch = min(ms_one, grouped[best_mb][0])
    air = int(sum(np.count_nonzero(L["types"]) for L in layouts))
    states = int(sum(len(L["state_branch"]) for L in layouts))
    per_step_s = ms_batch / 1e3 / K
    step_bytes = PER_AIR_CELL * air + PER_STATE * states
    config = {
        "rooms": len(batch), "air_cells": air, "merged_grid": merged_shape,
        "setup_s": round(setup_s, 2), "tuned": {k: shape.get(k, 0) for k in ("tile_z", "graph_steps", "face_blocks", "pairs", "us_per_step")},
        "sequential_ms": round(ms_seq, 2), "one_grid_ms": round(ms_one, 2),
        "grouped_ms": {f"{mb} MB": {"ms": round(v[0], 2), "groups": v[1]} for mb, v in grouped.items()},
        "batched": "one grid" if ms_one <= grouped[best_mb][0] else f"groups of at most {best_mb} MB",
        "batched_ms": round(ms_batch, 2),
        "batched_over_sequential_speedup": round(ms_seq / ms_batch, 3),
        "sequential_mcell_updates_per_s": round(air * K / (ms_seq / 1e3) / 1e6, 1),
        "special_cells": info["specials"],
        "batched_equals_plain_bitwise": same,
        "drive": f"Gaussian (sigma 0.6 ms), {W + K} steps; G is against this drive's, not a calibrated source's",
        "study": study,
    }
    return line_for("FDTD Mcell-updates/s (64 rooms batched; vs 64 sequential plain solves)",
                    air / per_step_s / 1e6, K, W, ms_batch,
                    "64 shoeboxes U(3,12)xU(3,10)xU(2.5,5) m at 5 cm, Monte Carlo wall absorption, "
                    "1 source and 4 receivers each, merged into batched grids", config,
                    roofline_of(step_bytes, per_step_s, None, f"{step_kernel_name(info)} over all rooms"))


def adjoint(args):
    """Absorption fitting on the 6 x 4 x 3 m shoebox at 5 cm (5,941 steps, 6 receivers):
    a wall admittance model with 12 unknowns (6 walls x (beta, G), one relaxation branch
    each), a seeded true model, targets from the forward solve plus white noise at 40 dB
    SNR. Timed (gradient_over_forward): one forward solve, and a gradient's two halves, the
    forward run keeping checkpoints every 256 steps (and the history of the special cells,
    the default) then the backward pass; the value is the whole gradient over the forward
    solve. Checked: the float and double adjoint's gradients of the decay-curve loss
    against central differences (h = 1e-3) of the double solve. Then 50 L-BFGS iterations
    on float solves from a flat start, and the recovered absorption at 125 and 250 Hz."""
    import torch

    from acoustic_spaces_amd import scenes
    from acoustic_spaces_amd.fit import WallModel, decay_loss, make_function
    from acoustic_spaces_amd.gpu import HipWaveSolver

    rng = np.random.default_rng(44)
    scene = scenes.shoebox(receivers=6, seed=44)
    m32 = WallModel(scene, precision=32, every=256)
    m64 = WallModel(scene, precision=64, every=256)
    truth = np.concatenate([rng.uniform(0.02, 0.15, 6), rng.uniform(0.05, 0.3, 6)])
    clean = m32.forward(truth)
    noise = rng.normal(size=clean.shape) * np.sqrt(np.mean(clean ** 2, axis=1, keepdims=True)) * 10 ** (-40 / 20)
    target = torch.from_numpy(clean + noise)
    theta0 = np.concatenate([np.full(6, 0.08), np.full(6, 0.15)])
    # Timing: a forward solve against a gradient (its record and backward pass), in both of
    # the record's modes.
    cost = gradient_over_forward()
    # The loss's gradient against central differences of the double solve.
    WaveSolve = make_function()

    def loss_grad(model, theta):
        th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
        loss = decay_loss(WaveSolve.apply(th, model), target, scene.rate)
        loss.backward()
        return loss.item(), th.grad.numpy()

    _, grad32 = loss_grad(m32, theta0)
    _, grad64 = loss_grad(m64, theta0)
    h = 1e-3
    fd = np.zeros(12)
    for p in range(12):
        e = np.zeros(12)
        e[p] = h
        up = decay_loss(torch.from_numpy(m64.forward(theta0 + e)), target, scene.rate).item()
        dn = decay_loss(torch.from_numpy(m64.forward(theta0 - e)), target, scene.rate).item()
        fd[p] = (up - dn) / (2 * h)
    rel32 = float(np.linalg.norm(grad32 - fd) / np.linalg.norm(fd))
    rel64 = float(np.linalg.norm(grad64 - fd) / np.linalg.norm(fd))
    # L-BFGS on log theta, 50 iterations, each loss and gradient from a float solve and its
    # adjoint.
    phi = torch.tensor(np.log(theta0), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([phi], lr=1.0, max_iter=50, history_size=20, line_search_fn="strong_wolfe",
                            tolerance_grad=1e-12, tolerance_change=1e-14)
    evals = [0]

    def closure():
        opt.zero_grad()
        loss = decay_loss(WaveSolve.apply(torch.exp(phi), m32), target, scene.rate)
        loss.backward()
        evals[0] += 1
        return loss

    t = time.perf_counter()
    opt.step(closure)
    fit_s = time.perf_counter() - t
    found = np.exp(phi.detach().numpy())
    a_true, a_found = m32.absorption(truth), m32.absorption(found)
    return {
        "metric": "adjoint gradient time / forward solve time (shoebox, 12 wall parameters)",
        "value": cost["value"],
        "unit": "ratio",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": int(m32.steps),
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "shoebox 6x4x3 m at 5 cm, 6 receivers, 0.5 s; walls Y = beta + G s/(s + w_c), "
                        "w_c = 2 pi 177 Hz; targets with white noise at 40 dB SNR",
            "gradient_cost": cost,
            "grad_f32_vs_f64_rel": float(np.linalg.norm(grad32 - grad64) / np.linalg.norm(grad64)),
            "grad_vs_fd_rel_f32": rel32, "grad_vs_fd_rel_f64": rel64, "fd_h": h,
            "lbfgs_iterations": 50, "lbfgs_evaluations": evals[0], "lbfgs_solves": "float", "fit_s": round(fit_s, 2),
            "absorption_true": np.round(a_true, 4).tolist(),
            "absorption_found": np.round(a_found, 4).tolist(),
            "absorption_max_error": round(float(np.abs(a_found - a_true).max()), 5),
        },
    }


def batched_over_sequential() -> dict:
    """The batching workload's ratio: 64 rooms at 5 cm (each 0.5 s, 5,941 steps, from rest)
    as 64 plain solves one after another against the batched solver (the rooms merged into
    grids of at most 192 MB, one launch a step each), both on the library's default launch
    shape and timed with hipEvents over every step; plus every room's receivers batched
    against its plain solve, bit for bit."""
    from acoustic_spaces_amd import scenes
    from acoustic_spaces_amd.batch import BatchedWaveSolver
    from acoustic_spaces_amd.gpu import HipWaveSolver
    from acoustic_spaces_amd.solver import WaveSolver

    batch = scenes.rooms_batch(count=64)
    t = time.perf_counter()
    layouts = [WaveSolver.from_scene(r).layout() for r in batch]
    setup_s = time.perf_counter() - t
    drive = batch[0].drive
    steps = len(drive)
    ms_seq, plain_out = 0.0, []
    for L in layouts:
        with HipWaveSolver(L) as s:
            ms, y = time_steps(s, drive, 0, steps)
        ms_seq += ms
        plain_out.append(y)
    b = BatchedWaveSolver(layouts, group_bytes=BatchedWaveSolver.DEFAULT_GROUP_BYTES)
    ms_batch, same = 0.0, True
    for idx, placements, s in b.groups:
        ms, y = time_steps(s, drive, 0, steps)
        ms_batch += ms
        for i, p in zip(idx, placements):
            same = same and np.array_equal(y[p.probes], plain_out[i])
    groups = len(b.groups)
    b.close()
    air = int(sum(np.count_nonzero(L["types"]) for L in layouts))
    return {
        "value": round(ms_seq / ms_batch, 3),
        "unit": "x (64 sequential plain solves' time over the batched solve's)",
        "sequential_ms": round(ms_seq, 2), "batched_ms": round(ms_batch, 2), "groups": groups,
        "steps_each": steps, "rooms": len(batch), "air_cells": air,
        "batched_mcell_updates_per_s": round(air * steps / (ms_batch / 1e3) / 1e6, 1),
        "batched_equals_plain_bitwise": bool(same), "setup_s": round(setup_s, 2),
        "workload": "64 shoeboxes U(3,12)xU(3,10)xU(2.5,5) m at 5 cm, Monte Carlo wall absorption, "
                    "1 source and 4 receivers each, 0.5 s",
    }


def gradient_over_forward() -> dict:
    """The differentiation workload's cost: on the 6 x 4 x 3 m shoebox at 5 cm (5,941
    steps, 6 receivers, 12 wall parameters), one forward solve against the gradient, as the
    autograd function runs it: the forward run keeping checkpoints every 256 steps, then
    the backward pass (each segment recomputed from its checkpoint, the adjoint stepped back
    through it). Host clocks around the synchronous calls, after one warm gradient."""
    import torch

    from acoustic_spaces_amd import scenes
    from acoustic_spaces_amd.fit import WallModel, decay_loss
    from acoustic_spaces_amd.gpu import HipWaveSolver

    rng = np.random.default_rng(44)
    scene = scenes.shoebox(receivers=6, seed=44)
    m32 = WallModel(scene, precision=32, every=256)
    truth = np.concatenate([rng.uniform(0.02, 0.15, 6), rng.uniform(0.05, 0.3, 6)])
    theta0 = np.concatenate([np.full(6, 0.08), np.full(6, 0.15)])
    target = torch.from_numpy(m32.forward(truth))
    with HipWaveSolver(m32.coefficients(theta0)) as stepper:
        stepper.run(m32.drive[:256])
        stepper.reset()
        t = time.perf_counter()
        y0 = stepper.run(m32.drive)
        forward_s = time.perf_counter() - t
    yt = torch.tensor(y0, requires_grad=True)
    decay_loss(yt, target, scene.rate).backward()
    dl = yt.grad.numpy()
    modes = {}
    for name, keep in (("history_kept", True), ("recomputed", False)):
        m = WallModel(scene, precision=32, every=256, keep_history=keep)
        m.gradient(theta0, dl)  # warm
        # The record's handle is set up before the timer, as the forward solve's is.
        stepper = HipWaveSolver(m.coefficients(theta0), precision=32, no_faces=True, keep_history=keep)
        t = time.perf_counter()
        stepper.record(m.drive, every=m.every)
        record_s = time.perf_counter() - t
        t = time.perf_counter()
        grad = m.backward(stepper, dl)
        backward_s = time.perf_counter() - t
        modes[name] = {"whole": round((record_s + backward_s) / forward_s, 3),
                       "backward_only": round(backward_s / forward_s, 3),
                       "record_over_forward": round(record_s / forward_s, 3),
                       "record_s": round(record_s, 4), "backward_s": round(backward_s, 4), "gradient": grad}
    same = bool(np.array_equal(modes["history_kept"].pop("gradient"), modes["recomputed"].pop("gradient")))
    kept = modes["history_kept"]
    return {
        "value": kept["whole"],
        "unit": "x (the whole gradient's time over one forward solve's: the record, then the backward pass)",
        "backward_only": kept["backward_only"],
        "record_over_forward": kept["record_over_forward"],
        "mode": "checkpoints every 256 steps and each step's history of the special cells kept (the library's "
                "default where it fits: 1.0 GB here), so that the backward pass recomputes nothing; each handle "
                "set up before its timer (the forward solve's and the record's alike)",
        "recomputed": modes["recomputed"],
        "recomputed_mode": "checkpoints every 256 steps only, each segment recomputed from its own",
        "gradients_equal_bitwise": same,
        "forward_s": round(forward_s, 4), "checkpoint_every": 256, "steps": int(m32.steps), "parameters": 12,
        "workload": "shoebox 6x4x3 m at 5 cm, 6 receivers, 0.5 s, 12 wall parameters (6 walls x (beta, G))",
    }


def setup_times() -> int:
    """The set-up's wall times, host against device passes, on the hall at 10 and 5 cm (both
    stencils) and outdoors at 25 cm: one JSON line; the stages' times on stderr."""
    import torch

    from acoustic_spaces_amd import scenes
    from acoustic_spaces_amd.solver import WaveSolver

    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    os.environ["AS_FDTD_TIMING"] = "1"
    rows = []
    for name, scene, stencils in (("hall 10 cm", scenes.hall(cell=0.1)[0], (7, 19)),
                                  ("hall 5 cm", scenes.hall(cell=0.05)[0], (7, 19)),
                                  ("outdoor 25 cm", scenes.outdoor(), (7,))):
        for stencil in stencils:
            scene.stencil = stencil
            row = {"scene": name, "stencil": stencil, "grid": list(scene.grid.shape)}
            layouts = []
            for on_device in (False, True):
                print(f"== {name}, {stencil} points, on_device={on_device}", file=sys.stderr, flush=True)
                t = time.perf_counter()
                solver = WaveSolver.from_scene(scene, on_device=on_device)
                row["device_s" if on_device else "host_s"] = round(time.perf_counter() - t, 3)
                layouts.append(solver.layout())
                row["air_cells"] = solver.air_cells
            row["same_layout"] = all(np.array_equal(layouts[0][k], layouts[1][k]) for k in layouts[0])
            rows.append(row)
            del layouts
    print(json.dumps({"setup_times": rows, "host_threads": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())}),
          flush=True)
    return 0


def strict(value):
    """``value`` with every NaN or infinity (an unreadable T30, say) as None, so that the
    line is strict JSON."""
    if isinstance(value, float):
        return value if np.isfinite(value) else None
    if isinstance(value, dict):
        return {k: strict(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [strict(v) for v in value]
    return value


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=256)
    ap.add_argument("--lead-in", type=int, default=256,
                    help="steps from rest before the warm-up (set-up, untimed), so that the timed "
                         "steps meet live fields")
    ap.add_argument("--workload", choices=("hall", "hall-faces", "shoebox", "outdoor", "rooms", "adjoint"),
                    default="hall")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-outputs", action="store_true",
                    help="skip the receiver outputs (the hall's and shoebox's band levels and T30, "
                         "float against double; the outdoor barrier's insertion loss)")
    ap.add_argument("--no-ratios", action="store_true",
                    help="the hall line without the batched-over-sequential and gradient-over-forward "
                         "ratios (the 64 rooms and the shoebox adjoint)")
    ap.add_argument("--graph-steps", type=int, default=0)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--face-blocks", type=int, default=-1, choices=(-1, 0, 1),
                    help="face specials by blocks of their own (1) or in their tiles (0); with any "
                         "of --tile-z, --graph-steps, --face-blocks the plain workloads skip the "
                         "autotuner")
    ap.add_argument("--special-order", type=int, default=0, choices=(0, 1, 2, 3),
                    help="the general specials' order: 1 by their counts, 2 by cell, 3 by tile (0: the "
                         "library's, 3)")
    ap.add_argument("--special-kinds", type=int, default=0, choices=(0, 1, 2),
                    help="the general specials read their kinds' tables (2) or their own slabs (1); "
                         "0: the library's choice")
    ap.add_argument("--split", action="store_true",
                    help="launch the step's roles (probes, general specials, face specials, tiles, layer) "
                         "as kernels of their own, for a profile of each; the line then times that")
    ap.add_argument("--calibrate", action="store_true",
                    help="time the runtime's own device-to-device copy kernel instead (the bandwidth "
                         "reference and the rocprofv3 counter scale)")
    ap.add_argument("--setup-times", action="store_true",
                    help="time the set-up instead, host against device passes (hall at 10 and 5 cm, outdoors)")
    args = ap.parse_args(argv)
    args.outputs, args.signals, args.insertion_loss, args.others = False, None, False, []
    if args.calibrate:
        return calibrate()
    if args.setup_times:
        return setup_times()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the solver is built and measured at N=1 only "
              f"(--gpus {args.gpus} asked for); a multi-GPU run is not part of this build.",
              file=sys.stderr)
        return 2

    from acoustic_spaces_amd import scenes

    if args.workload in ("hall", "hall-faces"):
        fitted = args.workload == "hall"
        args.steps = args.steps or 11882
        sources = scenes.hall()
        for sc in sources:
            sc.fitted_walls = fitted
        scene = sources[0]
        args.outputs = not args.no_outputs
        line = plain(scene, "hall" if fitted else "hall, walls on cell faces", args,
                     traffic_name="step_kernel_hall_fitted_traffic" if fitted else "step_kernel_hall_traffic")
        if args.outputs:
            # The second source's run (the config's two runs), its outputs as the first's.
            line["config"]["outputs_source2"] = source_outputs(sources[1], args)
        walls = ("fitted walls (this package's fit: each cut cell's air volume, open areas towards its "
                 f"18 neighbours and wall areas; {line['config']['special_cells']:,} special cells)"
                 if fitted else "walls on cell faces")
        line["config"]["workload"] = (
            "hall 30x20x12 m, ceiling of 6 tilted planes, side walls splayed 5 deg, 10 cm cells "
            f"{scene.grid.shape[0]}x{scene.grid.shape[1]}x{scene.grid.shape[2]}, "
            f"{line['config']['air_cells']:,} air cells, {walls}, 3 first-order admittance branches "
            "per material (rising or falling, fitted to its drawn absorption), 2 sources (the timed steps from the first; the outputs of both), 16 receivers")
    elif args.workload == "shoebox":
        args.steps = args.steps or 5941
        scene = scenes.shoebox()
        args.outputs = not args.no_outputs
        line = plain(scene, "shoebox", args, traffic_name="step_kernel_shoebox_traffic")
    elif args.workload == "outdoor":
        args.steps = args.steps or 2377
        scene = scenes.outdoor()
        args.insertion_loss = not args.no_outputs
        line = plain(scene, "outdoor", args)
    elif args.workload == "rooms":
        args.steps = args.steps or 5941
        line = rooms(args)
        scene = None
    else:
        line = adjoint(args)
        scene = None
    if not args.no_cpu and scene is not None:
        line["cpu_baseline"] = cpu_baseline(scene, compare=args.signals, others=args.others)
    if args.workload in ("hall", "hall-faces") and not args.no_ratios:
        # The metric's other two parts, on their own workloads, in the same run.
        line["batched_over_sequential"] = batched_over_sequential()
        line["gradient_over_forward"] = gradient_over_forward()
    print(json.dumps(strict(line), allow_nan=False), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
