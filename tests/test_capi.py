"""The C-ABI library loads and exports every entry point include/mauvehip.h declares.

No compute is run (no GPU here); only symbol resolution and host-side helpers."""
import ctypes
import os
import re

import mauvehip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "mauvehip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mauvehip.LIB_PATH)
    names = declared()
    assert len(names) >= 26
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert sorted(mauvehip.EXPORTS) == names


def test_error_paths_without_device():
    L = mauvehip.load_library()
    h = ctypes.c_void_p()
    rc = L.mh_create(0, ctypes.byref(h)) if os.environ.get("HIP_VISIBLE_DEVICES", "") == "-1" else None
    assert rc in (None, -1, -2)
    assert L.mh_sml_get_header(None, None) == -1
    assert L.mh_synchronize(None) == -1
