"""Builds libenlightning_hip.so (gfx950) in-tree with hipcc.

Every HIP source is compiled with:
  -ffp-contract=off   the reference arithmetic has no FMA contraction; keeping
                      the exact operation order is what makes the kernels agree
                      with it to the last bit outside exp/log;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions on this hardware pool.
The stage kernels' sources also take -mllvm -disable-machine-licm: their
kernels loop over tiles (persistent packs) or WENO cells, and machine
loop-invariant code motion moved the floating-point constants and LDS
addresses of a whole stencil out of those loops into registers held across
them (60-87 VGPRs spilled by the split kernel's pack loop; the WENO list
kernel at 197 VGPRs against 162 without it).  It moves code only, so the
results are the same bits.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libenlightning_hip.so")
SOURCES = ["enl_rhs.hip", "enl_kernels.hip", "enl_abi.hip", "enl_driver.hip", "enl_batch.hip", "enl_batch32.hip", "enl_switch.hip",
           "enl_absorb.hip", "enl_source.hip", "enl_regrid.hip", "enl_batch_rects.hip", "enl_batch_io.hip"]
ARCH = os.environ.get("ENL_OFFLOAD_ARCH", "gfx950")
# sources whose kernels loop over tiles or cells (see the module docstring)
NO_MACHINE_LICM = {"enl_rhs.hip", "enl_batch.hip", "enl_batch32.hip"}


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I" + CSRC, "-I" + INCLUDE,
    ]


def build(verbose=False, jobs=4):
    """Compile every source to an object, then link the shared library."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        extra = ["-mllvm", "-disable-machine-licm"] if src in NO_MACHINE_LICM else []
        cmd = [cc] + flags() + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature: harmless
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((cmd, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for cmd, text in failed:
            sys.stderr.write(" ".join(cmd) + "\n" + text + "\n")
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
