"""as_fdtd_run_static, the run that allocates nothing, through the C ABI: its probes equal
as_fdtd_run's bit for bit on two calls in a row (the second shorter), in float and double,
over runs long enough to take several pieces of its staging array; a layout whose one step
of probes outgrows the array is refused with a clear error."""

import ctypes

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _handle(lib, layout, precision):
    from acoustic_spaces_amd import native

    struct, keep = native.pack_layout(layout)
    h = lib.as_fdtd_create(ctypes.byref(struct), precision, None)
    assert h, lib.as_fdtd_create_error().decode()
    del keep
    return h


def _run(lib, fn, h, drive, probes):
    drive = np.ascontiguousarray(drive, np.float32)
    out = np.full((probes, len(drive)), np.nan)
    rc = fn(h, drive.ctypes.data, len(drive), out.ctypes.data)
    assert rc == 0, lib.as_fdtd_last_error(h).decode()
    return out


@pytest.mark.parametrize("precision", [32, 64])
def test_static_run_equals_run_on_two_calls(precision):
    from acoustic_spaces_amd import native, scenes
    from acoustic_spaces_amd.solver import WaveSolver

    lib = native.load()
    scene = cases.listeners13()  # 26 probes: pieces of 4,992 steps in the 1 MiB array
    layout = WaveSolver.from_scene(scene).layout()
    probes = len(layout["probe_start"]) - 1
    drive = scenes.gaussian_drive(scene.rate, 12000)
    first, second = drive[:9000], drive[9000:11500]
    a, b = _handle(lib, layout, precision), _handle(lib, layout, precision)
    try:
        print(f"precision {precision}: {probes} probes, calls of {len(first)} then {len(second)} steps", flush=True)
        for part in (first, second):
            want = _run(lib, lib.as_fdtd_run, a, part, probes)
            got = _run(lib, lib.as_fdtd_run_static, b, part, probes)
            assert np.array_equal(got, want)
            assert np.all(np.isfinite(got)) and np.any(got != 0)
    finally:
        lib.as_fdtd_destroy(a)
        lib.as_fdtd_destroy(b)


def test_static_run_refuses_probes_beyond_its_array():
    from acoustic_spaces_amd import native
    from acoustic_spaces_amd.solver import WaveSolver

    lib = native.load()
    layout = dict(WaveSolver.from_scene(cases.small_box()).layout())
    cell = int(np.flatnonzero(np.asarray(layout["types"]) == 1)[0])
    many = 131041  # one past what one step's doubles may take of 1 MiB
    layout["probe_start"] = np.arange(many + 1, dtype=np.uint64)
    layout["probe_cells"] = np.full(many, cell, np.uint64)
    layout["probe_weights"] = np.ones(many)
    h = _handle(lib, layout, 32)
    try:
        drive = np.zeros(4, np.float32)
        out = np.zeros((many, 4))
        rc = lib.as_fdtd_run_static(h, drive.ctypes.data, 4, out.ctypes.data)
        assert rc == native.AS_EINVAL
        assert "at most 131040 probes" in lib.as_fdtd_last_error(h).decode()
    finally:
        lib.as_fdtd_destroy(h)
