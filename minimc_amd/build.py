"""Builds libminimc_amd.so (HIP, gfx950) in-tree with hipcc.

python -m minimc_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libminimc_amd.so")
SOURCES = ["mmc_tracker.hip", "mmc_mlp.hip", "mmc_tally.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Contraction stays off for the tracker so that its arithmetic matches the
# reference's x86-64 build operation for operation.  Packed FP32 is disabled
# for every kernel.
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-Wno-unused-result"]
PER_SOURCE = {"mmc_tracker.hip": ["-ffp-contract=off"],
              "mmc_mlp.hip": ["-ffp-contract=off"],
              "mmc_tally.hip": ["-ffp-contract=off"]}


def build(verbose=False, jobs=3):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        cmd = [HIPCC] + COMMON + PER_SOURCE.get(src, []) + [
            "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs,
                   check=True)
    os.replace(tmp, LIB)
    build_h5(verbose)
    return LIB


# The pandas-HDF5 reader/writer (host code) against the HDF5 C library of the
# image.  Without that library the reader is not built, and reading an .hdf5
# table raises.
HDF5_ROOT = os.environ.get("MMC_HDF5_ROOT", "/opt/conda")
H5_LIB = os.path.join(OUT_DIR, "libmmc_h5.so")


def build_h5(verbose=False):
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print("no HDF5 headers under %s: libmmc_h5.so not built" % HDF5_ROOT)
        return None
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", inc,
           os.path.join(CSRC, "mmc_pandas_h5.cpp"), "-L", libdir, "-lhdf5",
           "-Wl,-rpath," + libdir, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
