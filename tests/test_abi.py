"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/idmx.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'idmx.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\bint\s+(idmx_\w+)\s*\(', txt)))


def test_header_declares_entry_points():
    names = _declared()
    assert 'idmx_sc_run' in names and 'idmx_ctx_create' in names
    assert len(names) >= 25


def test_library_exports_all_declared_symbols():
    from idmx import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.idmx_abi_version() == 1


def test_bindings_match_header():
    from idmx import _lib
    assert set(_declared()) == set(_lib.EXPORTED)


def test_error_channel_without_gpu():
    """Argument errors are reported through idmx_last_error, no GPU needed."""
    from idmx import _lib
    L = _lib.lib()
    rc = L.idmx_sc_create(None, None, None)
    assert rc == -1
    assert 'idmx_sc_create' in _lib.last_error()
