"""Build libabie_hip.so (gfx950) in-tree with hipcc.

    python -m abie_amd.build          # or abie_amd.build.build()

Every .hip translation unit compiles to an object in parallel, then they
link into abie_amd/libabie_hip.so.  Flags:
  -ffp-contract=off   products and sums round separately unless fma() is
                      written out (bit-exact GR15 state updates);
  -packed-fp32-ops    disabled: no packed FP32 instructions are emitted.
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libabie_hip.so")
SOURCES = ["force.hip", "gr15.hip", "gr15_fp64_plain.hip", "gr15_fp64_massive.hip", "gr15_fp64_learned.hip",
           "gr15_exact_plain.hip", "gr15_mixed_massive.hip", "gr15_mixed_learned.hip", "mlp.hip", "noalloc.hip",
           "clock.hip", "runtime.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ABIE_HIP_ARCH", "gfx950")
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wno-unused-command-line-argument"]


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    t0 = time.time()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write("=== %s ===\n%s\n" % (src, text))
        raise RuntimeError("hipcc failed for " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built %s in %.1f s" % (LIB, time.time() - t0))
    return LIB


if __name__ == "__main__":
    build(verbose=True)
