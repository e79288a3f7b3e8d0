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
    n = precompile_plugins(verbose="-v" in sys.argv)
    print(f"pre-compiled {n} kernel modules", flush=True)


if __name__ == "__main__":
    main()
