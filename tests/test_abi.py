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


def test_product_package_never_reaches_the_oracle():
    """The oracle is test infrastructure: no module of the product package
    imports it, loads its library or names its folder."""
    pkg = os.path.join(ROOT, 'idea.deploy_amd')
    hits = []
    for d, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(('.py', '.h', '.hip', '.cpp')):
                txt = open(os.path.join(d, f), encoding='utf-8').read()
                if re.search(r'\bimport\s+oracle\b|\bfrom\s+oracle\b|libcpuref|cpuref\.', txt):
                    hits.append(os.path.join(d, f))
    assert not hits, hits


def test_collision_enum_values():
    """IDMX_BGK / _MRT / _BGK_SCF / _BGK_ISO keep the values the Python
    engine writes into idmx_sc_desc.collision (0, 1, 2, 3)."""
    txt = open(os.path.join(ROOT, 'include', 'idmx.h')).read()
    vals = dict((k, int(v)) for k, v in re.findall(r'(IDMX_(?:BGK|MRT|BGK_SCF|BGK_ISO))\s*=\s*(\d+)',
                                                    txt))
    assert vals == {'IDMX_BGK': 0, 'IDMX_MRT': 1, 'IDMX_BGK_SCF': 2, 'IDMX_BGK_ISO': 3}
