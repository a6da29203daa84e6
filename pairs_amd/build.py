"""Build the gfx950 shared library libpairs_hip.so in-tree with hipcc.

Usage: python -m pairs_amd.build   (or pairs_amd.build.build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpairs_hip.so")
ARCH = os.environ.get("PAIRS_HIP_ARCH", "gfx950")

SOURCES = ["md_kernels.hip", "dem_kernels.hip", "mlp_kernels.hip", "pairs_hip.cpp", "dem_capi.cpp", "mlp_capi.cpp",
           "io.cpp"]

# -ffp-contract=off: every multiply and add rounds separately, as in the
#   reference's generated CPU code (bit-exact parity of the FP64 paths).
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 generated from float code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-command-line-argument"]


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I" + os.path.join(HERE, "..", "include"), "-o", tmp]
    for s in srcs:
        cmd += (["-x", "hip", s] if s.endswith(".cpp") else [s])
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_host_program(cpp, exe, verbose=False):
    """Compile a host program emitted by the DSL's HIP target (pairs_amd.dsl,
    emit_cpp) against the library: g++, the C ABI headers of include/, and
    libpairs_hip.so found at run time next to the package (rpath relative to
    the executable's directory when it lies inside the tree)."""
    if not os.path.exists(LIB):
        raise RuntimeError("libpairs_hip.so is not built: run pairs_amd.build.build() first")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    exe = os.path.abspath(exe)
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    rel = os.path.relpath(HERE, os.path.dirname(exe))
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"), "-o", exe + ".tmp", cpp,
           "-L" + HERE, "-lpairs_hip", "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath-link," + os.path.join(rocm, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
