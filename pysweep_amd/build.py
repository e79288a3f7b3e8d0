"""Build libpysweep_hip.so in-tree with hipcc for gfx950, and pre-compile the
shipped plugins' kernel modules into the code-object cache.

    python -m pysweep_amd.build            # library + shipped plugin kernels
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "_lib")
SRC = os.path.join(HERE, "csrc", "pysweep_hip.cpp")
OUT = os.path.join(LIBDIR, "libpysweep_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build_library():
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-I", os.path.join(ROOT, "include"), SRC, "-o", tmp,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


H5_SRC = os.path.join(HERE, "csrc", "h5io.c")
H5_OUT = os.path.join(LIBDIR, "libpysweep_h5.so")
H5_PREFIX = os.environ.get("PYSWEEP_HDF5_PREFIX", "/opt/conda")


def build_h5():
    """libpysweep_h5.so against the HDF5 C library when it is installed
    (HDF5 output in the reference's layout); None otherwise."""
    inc, lib = os.path.join(H5_PREFIX, "include"), os.path.join(H5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")) or not os.path.exists(os.path.join(lib, "libhdf5.so")):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = H5_OUT + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-I", inc, H5_SRC, "-o", tmp, "-L", lib, "-lhdf5",
           "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(tmp, H5_OUT)
    return H5_OUT


def precompile_plugins(verbose=False):
    """hiprtc-compile the kernel modules the tests, smoke() and bench.py use
    (no GPU needed) so the GPU host loads them from the cache."""
    from pysweep_amd import batch, kernels
    n = 0
    for spec in kernels.shipped_module_specs() + batch.shipped_class_specs():
        kernels.build_module_code(spec)
        n += 1
        if verbose:
            print("compiled", spec.describe(), flush=True)
    return n


def main():
    print("built", build_library(), flush=True)
    print("built", build_h5() or "no HDF5 library found: .npz output", flush=True)
    n = precompile_plugins(verbose="-v" in sys.argv)
    print(f"pre-compiled {n} kernel modules", flush=True)


if __name__ == "__main__":
    main()
