"""The C-ABI library loads and exports every symbol include/s1ifg.h declares
(no compute calls: no GPU needed)."""

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "s1ifg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(s1ifg_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    from sentinel_gpu_processor_amd import _native as N
    lib = N.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def test_ctypes_table_matches_header():
    from sentinel_gpu_processor_amd import _native as N
    assert sorted(N.SIGNATURES) == declared_symbols()


def test_abi_version_and_errors():
    import ctypes
    from sentinel_gpu_processor_amd import _native as N
    lib = N.load()
    assert lib.s1ifg_abi_version() == 1
    # argument validation happens before any device call
    rc = lib.s1ifg_ctx_set_stream(None, None)
    assert rc == -1
    assert b"null context" in lib.s1ifg_last_error()
    rc = lib.s1ifg_geom_query(None, None, 0)
    assert rc == -1


def test_struct_layouts():
    import ctypes
    from sentinel_gpu_processor_amd import _native as N
    # offsets that the C header implies on x86-64
    assert ctypes.sizeof(N.Orbit) == 32
    assert N.Burst.t_start.offset == 24
    assert N.Burst.n_fm.offset == 88
    assert N.Burst.fm.offset == 96
    assert ctypes.sizeof(N.Burst) == 112
    assert ctypes.sizeof(N.Grid) == 40
    assert N.Dem.lat0.offset == 24
    assert ctypes.sizeof(N.Dem) == 56
