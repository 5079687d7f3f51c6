"""CPU: the built C-ABI library loads and exports every function that
include/*.h declares, and the Python binding declares exactly those."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "fyzzxyzzy-sdk_amd" / "fyzz_amd" / "libfyzz_hip.so"


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(fyzz_hip_\w+)\s*\(", text))
    return names


def test_header_declares_functions():
    assert len(declared_functions()) >= 10


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail(f"{LIB} not built (run python fyzzxyzzy-sdk_amd/build.py)")
    lib = C.CDLL(str(LIB))
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_matches_header():
    from fyzz_amd import _lib
    assert set(_lib.SIGNATURES) == declared_functions()


def test_version_string_without_gpu():
    from fyzz_amd import lib
    assert lib().fyzz_hip_version().startswith(b"fyzz_hip gfx950")
