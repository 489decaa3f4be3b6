"""The C-ABI library loads and exports every function include/illume_c.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "illume_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(ilm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["ilm_open", "ilm_close", "ilm_scene_upload", "ilm_scene_free", "ilm_render", "ilm_render_batch",
                     "ilm_submit", "ilm_poll", "ilm_wait", "ilm_result", "ilm_cancel", "ilm_stats_get",
                     "ilm_last_error", "ilm_debug_closest_hit", "ilm_debug_path_radiance", "ilm_desc_load_json"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    import illume_amd
    L = illume_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_errors_without_gpu():
    import illume_amd
    L = illume_amd.lib()
    assert L.ilm_version() == 1
    # null arguments are rejected with ILM_E_INVALID, not a crash
    assert L.ilm_open(0, None, None) == -1
    assert L.ilm_render(None, None, None, None, None, None) == -1
    assert L.ilm_debug_instance_matrices(None, 0, None) == -1


def test_oracle_library_loads():
    from oracle.pyoracle import lib
    L = lib()
    assert L.orc_threads_available() >= 1


def test_open_fails_cleanly_without_device():
    import illume_amd
    if os.path.exists("/dev/kfd"):
        return   # a GPU box: opening succeeds there
    L = illume_amd.lib()
    h = ctypes.c_void_p()
    rc = L.ilm_open(0, None, ctypes.byref(h))
    assert rc == -3   # ILM_E_DEVICE
    assert b"device" in L.ilm_global_error()


@pytest.mark.parametrize("field,value", [("sample_group", 3), ("sample_group", -1), ("tile_rows", -2),
                                         ("extend_mode", 4), ("refill_min", 65), ("bvh_layout", 7), ("chunk_passes", -1)])
def test_open_rejects_bad_options(field, value):
    """Option checks run before any device call, so they hold with or without a GPU."""
    import illume_amd
    from illume_amd._abi import Opts
    L = illume_amd.lib()
    o = Opts()
    setattr(o, field, value)
    h = ctypes.c_void_p()
    assert L.ilm_open(0, ctypes.byref(o), ctypes.byref(h)) == -1   # ILM_E_INVALID
    assert not h.value
