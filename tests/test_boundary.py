"""The C/C++ boundary, with no GPU.

* Result formats: fiu_result_str / fiu_result_serialize / fiu_result_deserialize
  / fiu_write_results against the text and bytes the reference's own
  Result<int, float>::str and ::serialize write (include/fiuncho/utils/
  Result.h:37-66; the output file of app/main.cpp:152-158), from
  tests/golden/result_formats.json (written by the reference build).  The
  Python Result mirror is held to the same fixtures.
* GpuSearch (include/fiuncho_amd_search.h) compiles against the reference's
  headers: the Search interface it overrides (Search.h:19-46), Dataset,
  Distribution, Result.
* A C++ program links libfiuncho_amd.so and uses its host-only entry points
  through include/fiuncho_amd.h.
"""
import ctypes as C
import json
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import fiuncho_amd as fa
from fiuncho_amd import _lib

from .conftest import GOLDEN, ROOT

REF_INCLUDE = "/root/reference/include"


@pytest.fixture(scope="module")
def formats():
    with open(os.path.join(GOLDEN, "result_formats.json")) as f:
        return json.load(f)["results"]


def f32(hexs):
    return struct.unpack("<f", bytes.fromhex(hexs))[0]


def test_result_str_matches_reference(formats):
    L = fa.lib()
    for r in formats:
        combo = np.asarray(r["combo"], np.int32)
        mi = f32(r["mi_f32_hex"])
        n = L.fiu_result_str(combo.ctypes.data, len(combo), mi, None, 0)
        buf = C.create_string_buffer(n + 1)
        assert L.fiu_result_str(combo.ctypes.data, len(combo), mi, buf, n + 1) == n
        assert buf.value.decode() == r["str"], r
        # truncation keeps a terminated prefix and reports the full length
        small = C.create_string_buffer(4)
        assert L.fiu_result_str(combo.ctypes.data, len(combo), mi, small, 4) == n
        assert small.value.decode() == r["str"][:3]
        assert fa.Result([int(x) for x in combo], mi).str() == r["str"]


def test_result_serialize_matches_reference(formats):
    L = fa.lib()
    for r in formats:
        combo = np.asarray(r["combo"], np.int32)
        mi = f32(r["mi_f32_hex"])
        want = bytes.fromhex(r["serialize_hex"])
        assert L.fiu_result_serialize(combo.ctypes.data, len(combo), mi, None, 0) == len(want)
        buf = C.create_string_buffer(len(want))
        assert L.fiu_result_serialize(combo.ctypes.data, len(combo), mi, buf, len(want)) == len(want)
        assert buf.raw == want, r
        assert fa.Result([int(x) for x in combo], mi).serialize() == want
        # round trip
        back = np.zeros(16, np.int32)
        order, val, used = C.c_uint16(), C.c_float(), C.c_size_t()
        _lib.check(L.fiu_result_deserialize(buf, len(want), back.ctypes.data, 16, C.byref(order), C.byref(val),
                                            C.byref(used)))
        assert order.value == len(combo) and used.value == len(want)
        assert back[:len(combo)].tolist() == combo.tolist()
        assert struct.pack("<f", val.value) == struct.pack("<f", mi)


def test_result_deserialize_refuses_bad_records():
    L = fa.lib()
    combo = np.array([1, 2, 3], np.int32)
    rec = C.create_string_buffer(24)
    assert L.fiu_result_serialize(combo.ctypes.data, 3, 0.5, rec, 24) == 24
    back = np.zeros(8, np.int32)
    order, val = C.c_uint16(), C.c_float()
    assert L.fiu_result_deserialize(rec, 23, back.ctypes.data, 8, C.byref(order), C.byref(val), None) == \
        _lib.FIU_ERR_INVALID
    assert L.fiu_result_deserialize(rec, 24, back.ctypes.data, 2, C.byref(order), C.byref(val), None) == \
        _lib.FIU_ERR_INVALID
    assert L.fiu_result_deserialize(rec, 7, back.ctypes.data, 8, C.byref(order), C.byref(val), None) == \
        _lib.FIU_ERR_INVALID


def test_write_results_file(formats, tmp_path):
    L = fa.lib()
    rows = [r for r in formats if len(r["combo"]) == 3][:40]
    combos = np.asarray([r["combo"] for r in rows], np.int32)
    mi = np.asarray([f32(r["mi_f32_hex"]) for r in rows], np.float32)
    path = tmp_path / "out.txt"
    _lib.check(L.fiu_write_results(str(path).encode(), combos.ctypes.data, mi.ctypes.data, len(rows), 3))
    assert path.read_text() == "".join(r["str"] + "\n" for r in rows)
    assert L.fiu_write_results(str(tmp_path / "no" / "such" / "dir.txt").encode(), combos.ctypes.data,
                               mi.ctypes.data, len(rows), 3) == _lib.FIU_ERR_IO


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or shutil.which("g++") is None,
                    reason="the reference's headers are not present")
def test_gpusearch_header_compiles_against_the_reference(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "fiuncho_amd_search.h"\n'
                   "#include <memory>\n"
                   "static_assert(std::is_base_of<Search, GpuSearch>::value, \"a Search\");\n"
                   "std::unique_ptr<Search> make() { return std::make_unique<GpuSearch>(0); }\n")
    r = subprocess.run(["g++", "-std=gnu++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-DALIGN=64",
                        "-I" + REF_INCLUDE, "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_cxx_caller_host_entry_points(tmp_path):
    exe = tmp_path / "abi_caller"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "abi_caller.cpp"), "-L" + os.path.join(ROOT, "fiuncho_amd"),
                        "-lfiuncho_amd", "-Wl,-rpath," + os.path.join(ROOT, "fiuncho_amd"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), "host"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "abi_caller ok" in r.stdout, r.stdout + r.stderr
