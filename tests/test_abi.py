"""The C-ABI library loads and exports every symbol include/s1ifg.h declares
(no compute calls: no GPU needed)."""

import os
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_symbols():
    src = open(ROOT / "include" / "s1ifg.h").read()
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
    assert lib.s1ifg_stitch_subswath(None, None, 0, 0, None, 0, 0, 0, 0, None) == -1
    assert lib.s1ifg_merge_subswath(None, None, 0, 0, 0, None, 0) == -1
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


def test_struct_layouts_match_c_compiler(tmp_path):
    """Every ctypes structure has the size and field offsets that a C
    compiler gives the header's structures."""
    import ctypes
    import shutil
    import subprocess
    from sentinel_gpu_processor_amd import _native as N
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"s1ifg_orbit": N.Orbit, "s1ifg_burst": N.Burst, "s1ifg_grid": N.Grid,
               "s1ifg_dem": N.Dem, "s1ifg_pair_task": N.PairTask,
               "s1ifg_batch_opts": N.BatchOpts, "s1ifg_stitch_burst": N.StitchBurst}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "s1ifg.h"',
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _t in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, value = line.split()
        cls = structs[cname]
        got = ctypes.sizeof(cls) if field == "sizeof" else getattr(cls, field).offset
        assert got == int(value), (cname, field, got, value)


def test_ps_scan_offsets_matches_oracle():
    """The library's neighbour table (host code, no GPU) against the
    oracle's restatement of scan_array (phase_similarity.cu:295-366), and
    its argument checks."""
    from oracle import psps as P
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    import numpy as np
    for rdmin, rdmax in ((0, 5), (2, 5), (0, 1), (0, 0), (-3, 6), (3, 30), (0, 64), (10, 200)):
        np.testing.assert_array_equal(K.ps_scan_offsets(rdmin, rdmax),
                                       P.scan_offsets(max(rdmin, 0), rdmax))
    lib = N.load()
    assert lib.s1ifg_ps_scan_offsets(0, 5000, None, 0) == N.S1IFG_ERANGE
    assert lib.s1ifg_ps_scan_offsets(0, 5, None, 4) == N.S1IFG_EINVAL
    assert lib.s1ifg_ps_scan_offsets(0, 5, None, 0) == 8 + 12 + 16 + 32
