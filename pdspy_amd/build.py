"""Build the gfx950 shared library ``pdspy_amd/lib/libpdspy_amd.so`` with hipcc.

Run ``python -m pdspy_amd.build``.  Cross-compiles without a GPU.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libpdspy_amd.so")
ARCH = os.environ.get("PDSPY_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed-FP32 instructions stay off (no v_pk_* fp32 ops are generated)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # hardware FP64 atomics for the adjoint accumulators
    "-munsafe-fp-atomics",
    # FP32 divide and square root by the hardware approximations (<= 2.5 ulp)
    # instead of the correctly rounded expansions; near-tangent geometry is
    # re-evaluated in FP64 anyway
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose=False, extra=(), out=None):
    """Compile every source and link the library at ``out`` (default LIB).

    ``extra`` adds hipcc flags, e.g. ("-DRT_THERMAL_PROFILE=1",) for the
    phase-clock build."""
    lib = out or LIB
    libdir = os.path.dirname(os.path.abspath(lib))
    os.makedirs(libdir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    for src in sources():
        obj = os.path.join(libdir, os.path.basename(src) + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, *extra, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = lib + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    for o in objs:
        os.remove(o)
    return lib


if __name__ == "__main__":
    args = sys.argv[1:]
    out = None
    if "-o" in args:
        out = args[args.index("-o") + 1]
    extra = [a for a in args if a.startswith("-D")]
    print(build(verbose="-v" in args, extra=extra, out=out))
