"""Build libnes_render.so (gfx950 HIP kernels + C-ABI) in-tree.

    python ngp-encode-server_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libnes_render.so")
SOURCES = ["common.hip", "scene.hip", "frameprep.hip", "overlay.hip", "render.hip", "train.hip",
           "serve.hip", "wire.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # never emit packed FP32 ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-command-line-argument",
    "-pthread",
]


# Per-file additions.  render.hip: keep MFMA results in VGPRs (the FP16 MLP
# otherwise copies every accumulator out of AGPRs, one instruction per
# register); train.hip keeps the default because its FP32 MFMA kernel needs
# the AGPRs for its accumulators.
FILE_FLAGS = {"render.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _filter(text):
    return "\n".join(l for l in text.splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l)


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    obj_dir = os.path.join(LIB_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    jobs = []
    for src in sources():
        obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
        flags = [f for f in HIPCC_FLAGS if f != "-shared"]
        cmd = [hipcc] + flags + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                           text=True)))
    objs, failed = [], False
    for obj, proc in jobs:
        out, _ = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(_filter(out) + "\n")
            failed = True
        objs.append(obj)
    if failed:
        raise RuntimeError("hipcc failed building libnes_render.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread"] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(_filter(res.stdout) + "\n")
        raise RuntimeError("linking libnes_render.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
