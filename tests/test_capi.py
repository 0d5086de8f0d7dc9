"""The C-ABI library loads and exports every entry point include/mauvehip.h declares.

No compute is run (no GPU here); only symbol resolution and host-side helpers."""
import ctypes
import os
import re

import mauvehip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class SmlHeaderFile(ctypes.Structure):
    """SMLHeader as the reference declares it (LM/SortedMerList.h:48-63;
    boolean is unsigned char, sarID_t int16, UINT8_MAX 255, DESCRIPTION_SIZE
    2048), laid out by the C ABI of the platform (x86-64)."""
    _fields_ = [("version", ctypes.c_uint32), ("alphabet_bits", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("seed_length", ctypes.c_uint32), ("seed_weight", ctypes.c_uint32), ("length", ctypes.c_uint64),
                ("unique_mers", ctypes.c_uint32), ("word_size", ctypes.c_uint32),
                ("little_endian", ctypes.c_ubyte), ("id", ctypes.c_int16), ("circular", ctypes.c_ubyte),
                ("translation_table", ctypes.c_ubyte * 255), ("description", ctypes.c_char * 2048)]


SML_HEADER_BYTES = ctypes.sizeof(SmlHeaderFile)


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


def test_sslist_header_layout():
    """The offsets the .sslist writer and reader use (sml_file.cpp) are the
    ones the C ABI gives SMLHeader."""
    want = {"version": 0, "alphabet_bits": 4, "seed": 8, "seed_length": 16, "seed_weight": 20, "length": 24,
            "unique_mers": 32, "word_size": 36, "little_endian": 40, "id": 42, "circular": 44,
            "translation_table": 45, "description": 300}
    got = {f: getattr(SmlHeaderFile, f).offset for f in want}
    assert got == want
    assert SML_HEADER_BYTES == 2352
    src = open(os.path.join(ROOT, "mauve-py_amd", "csrc", "sml_file.cpp")).read()
    assert "HEADER_BYTES = 2352" in src


def test_sslist_null_arguments():
    L = mauvehip.load_library()
    h = ctypes.c_void_p()
    st = ctypes.c_int(-5)
    assert L.mh_sml_save(None, None, b"x") == -1
    assert L.mh_sml_load(None, b"x", ctypes.byref(h), ctypes.byref(st)) == -1
