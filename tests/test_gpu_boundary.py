"""GPU: the boundary used from C++.

* oracle/_ref/shim_check: GpuSearch (include/fiuncho_amd_search.h) compiled
  against the reference's headers, reading tests/data through the reference's
  own Dataset<uint64_t>::read and searching it as the reference's tests use
  ThreadedSearch (tests/threadedsearch.cpp:17-31), next to ThreadedSearch(1)
  on the same calls (orders 2 to 4, the whole scan and a (3, 1) share, the
  dataset uploaded per call and bound once): lists equal up to ties.  Built
  where the reference sources are (oracle/Makefile, target shim).
* tests/cxx/abi_caller.cpp: a plain C++ caller of the C ABI searching a
  synthetic dataset; its Result::str lines equal those of the same searches
  made from Python.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import fiuncho_amd as fa

from .conftest import DATA, ROOT

pytestmark = pytest.mark.gpu

SHIM = os.path.join(ROOT, "oracle", "_ref", "shim_check")


@pytest.mark.skipif(not os.path.exists(SHIM), reason="oracle/_ref/shim_check not built (needs the reference sources)")
def test_gpusearch_shim_against_threadedsearch():
    print("case: shim_check on tests/data", flush=True)
    r = subprocess.run([SHIM, os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "shim ok" in r.stdout, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert sum(1 for x in lines if x.startswith("order ")) == 6
    # the order-3 Result::str lines: the top triple of tests/data is (4, 8, 9)
    first = next(x for x in lines if not x.startswith("order "))
    assert first.startswith("4 8 9 ")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_cxx_caller_search(tmp_path):
    exe = tmp_path / "abi_caller"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "abi_caller.cpp"), "-L" + os.path.join(ROOT, "fiuncho_amd"),
                        "-lfiuncho_amd", "-Wl,-rpath," + os.path.join(ROOT, "fiuncho_amd"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    print("case: C++ caller search", flush=True)
    r = subprocess.run([str(exe), "gpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "abi_caller ok" in r.stdout, r.stdout + r.stderr
    got = [x for x in r.stdout.splitlines() if x != "abi_caller ok"]
    spec = fa.synth_spec(300, 1000, 900, seed=5, planted=(7, 250))
    d = fa.Dataset.synth(spec)
    want = []
    for order, outputs in ((2, 10), (3, 5)):
        want += [x.str() for x in fa.GpuSearch().run(d, order, fa.Distribution(300, order - 1), outputs)]
    d.close()
    assert got == want
