"""Build recipe for liblp.so (HIP kernels for gfx950 + C++ host runtime).

Compiles in-tree with hipcc so the shared library travels with the source
tree.  Flags that the kernel's bit-exact arithmetic depends on:
  -ffp-contract=off                          no fused multiply-add contraction
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE-rounded fp32 / and sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                             no packed fp32 instructions
"""
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblp.so")

SOURCES = ["csrc/lp_kernel.hip", "csrc/lp_capi.cpp", "csrc/lp_bvh.cpp", "csrc/lp_model.cpp", "host/lp_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    "-ffp-contract=off", "-fno-fast-math",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(HERE, "host"),
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + cmd[-1])
    return r


def build(verbose=False, jobs=None):
    """Compile every source to an object, then link liblp.so."""
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(HERE, src)
        obj = os.path.join(build_dir, os.path.splitext(os.path.basename(src))[0] + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC] + COMMON + DEVICE + ["-c", "-o", obj, path]
        else:
            cmd = [HIPCC] + COMMON + ["-x", "c++", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", "-o", obj, path]
        if verbose:
            print(" ".join(shlex.quote(c) for c in cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), src))
        objs.append(obj)
    failed = []
    for p, src in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(src)
        elif verbose and err.strip():
            sys.stderr.write(err)
    if failed:
        raise RuntimeError("build failed: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + ["-lamdhip64"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
