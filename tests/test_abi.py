"""The C ABI: every function include/minimc_amd.h declares is exported by
libminimc_amd.so, and the ctypes mirror has the C struct layouts (sizes
measured by compiling the header with gcc).  No GPU calls."""
import os
import re
import subprocess

import ctypes as C
import pytest

from minimc_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "minimc_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(mmc_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mmc_create", "mmc_run_fixed_source", "mmc_run_cycle",
                 "mmc_get_tallies", "mmc_eval_event", "mmc_destroy", "mmc_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from minimc_amd import tracker
    if not os.path.exists(tracker.LIB_PATH):
        pytest.fail("libminimc_amd.so not built")
    L = C.CDLL(tracker.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.mmc_abi_version() == 3


def test_struct_sizes_match_c(tmp_path):
    names = sorted(abi.STRUCTS)
    src = ['#include <stdio.h>', '#include "%s"' % HEADER, "int main(void){"]
    for n in names:
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (n, n))
    src.append("return 0;}")
    c = tmp_path / "sizes.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.split("\n"):
        if line.strip():
            n, size = line.split()
            assert C.sizeof(abi.STRUCTS[n]) == int(size), n
