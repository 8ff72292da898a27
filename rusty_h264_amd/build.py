"""Builds librh264_me.so for gfx950 with hipcc (in-tree, so it ships with the tree)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librh264_me.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    # no packed-FP32 instructions anywhere in the build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per source: extra flags.  The search kernels keep matrix-core results in ordinary
# VGPRs (the refinement reads them back with plain VALU ops, no accumulator moves).
SOURCES = {
    "rh264_me.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "rh264_mvnet.hip": [],
    "rh264_accel.hip": [],
    "rh264_bitstream.hip": [],
}


def build_library(verbose=False):
    objs, procs = [], []
    for src, extra in SOURCES.items():
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC] + FLAGS + extra + ["-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    for cmd, pr in procs:
        out, err = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out + err)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
    for o in objs:
        os.remove(o)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
