"""The C-ABI library: it builds, loads, exports every entry point declared in
include/*.h, and reports errors through status codes plus mfh_last_error
(no compute calls; these run without a GPU)."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(headers=None):
    names = set()
    for h in headers or glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mfh_\w+|mufflon_\w+|render_\w+|renderer_\w+|"
                             r"world_\w+|core_\w+|polygon_\w+|object_\w+|instance_\w+|scenario_\w+|scene_\w+|"
                             r"profiling_\w+|loader_\w+)"
                             r"\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    from mufflon_amd import hip
    if not os.path.exists(hip.library_path()):
        from mufflon_amd import build
        build.build_library()
    return hip.load()


def test_headers_declare_entry_points():
    names = _declared()
    assert "mfh_create" in names and "mfh_trace_closest" in names and "mfh_lbvh_build" in names


def test_every_declared_symbol_is_exported(lib):
    loader = os.path.join(ROOT, "include", "mff_loader.h")
    names = _declared([h for h in glob.glob(os.path.join(ROOT, "include", "*.h")) if h != loader])
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_scene_loader_exports_its_header():
    from mufflon_amd import core
    ldr = core.load_loader()
    names = _declared([os.path.join(ROOT, "include", "mff_loader.h")])
    assert "loader_load_json" in names and "loader_load_lod" in names
    missing = [n for n in names if not hasattr(ldr, n)]
    assert not missing, missing


def test_error_convention_without_compute(lib):
    h = C.c_void_p()
    rc = lib.mfh_create(987654, C.byref(h))
    assert rc != 0
    assert h.value is None
    assert len(lib.mfh_last_error()) > 0
    # a null context is rejected, not dereferenced
    assert lib.mfh_synchronize(None) == 1
    assert b"null context" in lib.mfh_last_error()
    out = C.c_int32()
    assert lib.mfh_lod_create(None, None, 0, None, 0, C.byref(out)) != 0
