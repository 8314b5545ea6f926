"""Builds libptg.so (the HIP product library) in-tree for gfx950, and
libptg_h5.so (host-only HDF5 particle dump and beam import) where libhdf5 is
installed.

    python -m ptarmigan_amd.build           # or: python ptarmigan_amd/build.py

Every csrc/*.hip file is compiled to an object with hipcc (in parallel) and
linked into ptarmigan_amd/libptg.so next to tables/ptg_tables.bin, where the
library looks for its tables by default. csrc/h5io.cpp is compiled with g++
against the HDF5 C library found under HDF5_DIR (default /opt/conda); without
one the HDF5 output and import are unavailable and decks asking for them are
refused.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libptg.so")
ROOT = os.path.dirname(HERE)

H5LIB = os.path.join(HERE, "libptg_h5.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    # packed FP32 forms are kept out of generated code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # spills go to scratch, not to AGPRs: a kernel capped at 256 VGPRs for two
    # waves per SIMD (the LP event kernels) otherwise takes AGPRs on top and
    # drops to one wave per SIMD (13 % slower at a0 = 10)
    "-mllvm", "-amdgpu-spill-vgpr-to-agpr=0",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    shutil.move(tmp, LIB)
    build_h5()
    return LIB


def build_h5():
    """libptg_h5.so from csrc/h5io.cpp, linked to HDF5_DIR's libhdf5 (with its
    directory as the run path). Returns the library, or None without libhdf5."""
    inc, lib = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(lib, "libhdf5.so"))):
        return None
    tmp = H5LIB + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", "-I", inc, "-o", tmp,
           os.path.join(CSRC, "h5io.cpp"), "-L", lib, "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"h5io build failed:\n{r.stderr[-8000:]}")
    shutil.move(tmp, H5LIB)
    return H5LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
