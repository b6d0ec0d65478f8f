"""The C-ABI library loads and exports every symbol include/asfdtd.h declares (no device
call: this runs on the CPU)."""

import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "asfdtd.h").read_text()
    return sorted(set(re.findall(r"\b(as_(?:fdtd|setup)_[a-z_]+)\s*\(", text)))


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
