"""The C-ABI library loads and exports every entry point the header declares
(no compute calls: runs without a GPU)."""

import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gnss_gpu_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    assert "gg_pf_device_weight" in names and "gg_wls_batch" in names
    assert len(names) > 40


def test_library_exports_every_declared_symbol():
    from gnss_gpu_amd import _lib

    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from gnss_gpu_amd import _lib

    assert sorted(_declared()) == sorted(_lib.EXPORTED)


def test_struct_layouts_match():
    from gnss_gpu_amd import _lib

    sa, dy, fi = C.c_int32(), C.c_int32(), C.c_int32()
    assert _lib.lib.gg_struct_sizes(C.byref(sa), C.byref(dy), C.byref(fi)) == 0
    assert sa.value == C.sizeof(_lib.StepArgs)
    assert dy.value == C.sizeof(_lib.Dyn)


def test_error_channel_without_gpu_work():
    from gnss_gpu_amd import _lib

    # invalid arguments are rejected before any device call
    h = C.c_void_p()
    rc = _lib.lib.gg_pf_create(0, None, None, 0, 1, 0, C.byref(h))
    assert rc != 0
    assert "n_filters" in _lib.last_error()
