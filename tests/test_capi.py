"""The HIP library loads on the CPU host and exports every symbol declared in
include/cavalry.h (no device calls)."""
from cavalry_amd import _lib


def test_library_exports_header_symbols():
    lib = _lib.lib()
    names = _lib.header_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in lib.cav_version()


def test_ctypes_signatures_cover_header():
    declared = set(_lib.header_symbols())
    bound = set(_lib._SIGS)
    assert declared <= bound, declared - bound
