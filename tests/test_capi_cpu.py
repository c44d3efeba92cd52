"""The C-ABI library loads without a GPU and matches include/measure_ia_amd.h.

No compute call is made here: the library is loaded, every function the
header declares must be exported, the ctypes mirrors must have the header's
struct layouts (checked with a small C program built by gcc), and the
argument checks must reject a bad call before touching a device.
"""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from measure_ia_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "measure_ia_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void|const char \*)\s*(mia_\w+)\s*\(", text, re.M)))


def test_header_functions_are_bound():
    names = _declared_functions()
    assert names, "no functions parsed from the header"
    assert sorted(_capi.SIGNATURES) == names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("library not built (python -m measure_ia_amd.build)")
    return _capi.load()


def test_library_exports_every_symbol(lib):
    for name in _declared_functions():
        assert hasattr(lib, name), name
    assert lib.mia_abi_version() == 4


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_struct_layouts_match_header(tmp_path):
    structs = {"mia_sample": _capi.MiaSample, "mia_bins": _capi.MiaBins,
               "mia_counts": _capi.MiaCounts, "mia_tuning": _capi.MiaTuning,
               "mia_stats": _capi.MiaStats, "mia_lc_sample": _capi.MiaLcSample}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for s, cls in structs.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{s}.{fname} %zu\\n", offsetof({s}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for s, cls in structs.items():
        assert int(out[s]) == ctypes.sizeof(cls), s
        for fname, _ in cls._fields_:
            assert int(out[f"{s}.{fname}"]) == getattr(cls, fname).offset, f"{s}.{fname}"


def test_null_context_is_rejected(lib):
    rc = lib.mia_count(None, None, None, None, 0, None, None, None)
    assert rc == -1
    assert lib.mia_region_ids(None, None, 0, 1.0, 1, None, None) == -1
    assert lib.mia_count_lightcone(None, None, None, None, None, None, None) == -1
    assert lib.mia_prepare_sample(None, None, None, None, None, None) == -1
    assert lib.mia_count_prepared(None, 0, None, None, None, None, None, None) == -1
    lib.mia_release_prepared(None, None)  # NULL is ignored
    assert lib.mia_count_noalloc(None, None, None, None, None, None) == -1
    assert lib.mia_noalloc_workspace_bytes() == 256 << 20
