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
