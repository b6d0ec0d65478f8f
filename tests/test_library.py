"""The C-ABI library loads and exports every symbol include/asfdtd.h declares (no device
call: this runs on the CPU)."""

import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "asfdtd.h").read_text()
    return sorted(set(re.findall(r"\b(as_(?:fdtd|setup|first)_[a-z_]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    from acoustic_spaces_amd import native

    assert sorted(native.SYMBOLS) == [s for s in _declared() if s.startswith("as_fdtd_")]


def test_library_exports_every_declared_symbol():
    import ctypes

    from acoustic_spaces_amd import native

    assert native.LIBRARY.exists(), "build first: __graft_entry__.build()"
    lib = ctypes.CDLL(str(native.LIBRARY))
    for name in _declared():
        assert hasattr(lib, name), name


def test_pack_layout_checks_each_array_against_its_count():
    """The binding refuses, before the library sees them, arrays shorter or longer than the
    count that indexes them (so the library never reads past a host array's end)."""
    import numpy as np
    import pytest

    import cases
    from acoustic_spaces_amd import native
    from acoustic_spaces_amd.solver import WaveSolver

    L = WaveSolver.from_scene(cases.small_box()).layout()
    native.pack_layout(L)
    for key in ("coef_a", "coef_b", "coef_inv", "source", "link_weight", "state_weight",
                "branch_before", "branch_gain", "probe_weights"):
        bad = dict(L)
        bad[key] = np.asarray(L[key])[:-1]
        with pytest.raises(ValueError, match=key):
            native.pack_layout(bad)
    bad = dict(L)
    bad["weights"] = np.asarray(L["weights"])[:4]
    with pytest.raises(ValueError, match="weights"):
        native.pack_layout(bad)
