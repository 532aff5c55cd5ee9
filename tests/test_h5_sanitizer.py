"""The pandas-HDF5 reader and writer (minimc_amd/csrc/mmc_pandas_h5.cpp)
under AddressSanitizer and UndefinedBehaviorSanitizer: the same source built
with g++ -fsanitize=address,undefined, loaded by a child Python process that
preloads the sanitizer runtime, writes and reads back tables of one to three
levels, and hits the error paths (missing file, wrong level count, a file
that is not pandas-HDF5).  Any sanitizer report fails the test.  CPU only."""
import os
import subprocess
import sys

import pytest

from minimc_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCENARIO = r"""
import os, sys, tempfile
import numpy as np
from minimc_amd import pandas_h5 as h5
d = tempfile.mkdtemp()
rng = np.random.default_rng(1)
for shape in ((7,), (5, 3), (4, 6, 2), (1, 1, 1)):
    axes = [np.sort(rng.random(n)) for n in shape]
    vals = rng.random(shape)
    p = os.path.join(d, "t%d.hdf5" % len(shape))
    h5.write(p, axes, vals)
    ax, v = h5.read(p, len(shape))
    assert all(np.array_equal(a, b) for a, b in zip(axes, ax)) and np.array_equal(v, vals)
for bad in ((os.path.join(d, "missing.hdf5"), 1), (os.path.join(d, "t2.hdf5"), 3)):
    try:
        h5.read(*bad)
        raise SystemExit("no error for %r" % (bad,))
    except RuntimeError as e:
        print("error as expected:", e)
junk = os.path.join(d, "junk.hdf5")
open(junk, "wb").write(b"not an hdf5 file" * 10)
try:
    h5.read(junk)
    raise SystemExit("no error for junk")
except RuntimeError as e:
    print("error as expected:", e)
print("scenario done")
"""


def test_h5_reader_writer_under_asan(tmp_path):
    inc = os.path.join(build.HDF5_ROOT, "include")
    libdir = os.path.join(build.HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        pytest.skip("no HDF5 C library")
    runtime = subprocess.run(["gcc", "-print-file-name=libasan.so"], check=True,
                             capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(runtime):
        pytest.skip("no libasan runtime for gcc")
    # The HDF5 library's folder also holds another compiler's sanitizer and
    # C++ runtimes: link and load libhdf5 from a folder of its own, so that
    # the instrumented library gets gcc's runtimes (those the child preloads).
    h5dir = tmp_path / "hdf5"
    h5dir.mkdir()
    real = os.path.realpath(os.path.join(libdir, "libhdf5.so"))
    soname = subprocess.run(["objdump", "-p", real], check=True, capture_output=True,
                            text=True).stdout.split("SONAME", 1)[1].split()[0]
    h5 = h5dir / soname
    h5.symlink_to(real)
    lib = str(tmp_path / "libmmc_h5_asan.so")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
                    "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-I", inc,
                    os.path.join(build.CSRC, "mmc_pandas_h5.cpp"), str(h5),
                    "-Wl,-rpath," + str(h5dir), "-o", lib], check=True)
    env = dict(os.environ, MMC_H5_LIB=lib, LD_PRELOAD=runtime, PYTHONPATH=ROOT,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([sys.executable, "-c", SCENARIO], env=env, capture_output=True,
                       text=True, timeout=600, cwd=ROOT)
    report = [l for l in p.stderr.splitlines() if "Sanitizer" in l or "runtime error" in l]
    assert p.returncode == 0 and "scenario done" in p.stdout, (p.stdout[-2000:], p.stderr[-3000:])
    assert not report, "\n".join(report[:20])
