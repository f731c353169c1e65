"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import re
from pathlib import Path

from gtrap_amd import _native

HEADER = Path(__file__).resolve().parent.parent / "include" / "gtrap_hip.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(gtrap_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "gtrap_medfilt" in names and len(names) >= 4


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(_native.SIGNATURES) == declared()


def test_error_reporting_without_gpu():
    lib = _native.load()
    # argument validation runs on the host before any HIP call
    st = lib.gtrap_medfilt(None, None, 10, 1, 1, 0, 10, 1, 0, None)
    assert st == -1
    assert b"null" in lib.gtrap_last_error()
    assert lib.gtrap_version().startswith(b"gtrap_amd")


def test_host_validation_of_tlsfold_settings_and_sizes():
    """Launch-geometry settings and workspace sizes are checked and computed on
    the host, before any HIP call: out-of-range tile settings are refused with
    the allowed range in the message, the valid ones are accepted (ending on
    the built-in geometry), and the workspace grows with nmax and nblocks."""
    lib = _native.load()
    for bad in ((0, 1, 2, 0), (145, 1, 2, 0), (144, 3, 2, 0), (144, 1, 9, 0), (144, 1, 2, 49)):
        assert lib.gtrap_tlsfold_set_tiles(*bad) == -1, bad
        assert b"gtrap_tlsfold_set_tiles" in lib.gtrap_last_error()
    assert lib.gtrap_tlsfold_set_tiles(96, 2, 1, 24) == 0
    assert lib.gtrap_tlsfold_set_tiles(144, 2, 2, 0) == 0
    one = lib.gtrap_tlsfold_ws_bytes(1000, 1)
    assert one >= 256 + 1000 * 20 and (one - 256) % 256 == 0
    assert lib.gtrap_tlsfold_ws_bytes(1000, 8) == 256 + 8 * (one - 256)
    assert lib.gtrap_tlsfold_ws_bytes(70000, 1) > one
    # geometry arguments of the static entry point are refused before any
    # device work (no current device needed for that)
    st = lib.gtrap_tlsfold_static(None, None, None, None, None, None, None, None, None, 0, None,
                                  None, 1, 1, None, 10, 10, 0.0204, 0, 64, None)
    assert st == -1
