"""Build libbbgpu.so (the HIP kernels and the C-ABI) in-tree with hipcc for gfx950: every source compiled to an
object in parallel (under bbarolo2_amd/_build, not kept in the repository), then linked.

python -m bbarolo2_amd.build
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbbgpu.so")
SOURCES = ["bb_capi.hip", "bb_like.hip", "bb_galmod.hip", "bb_conv.hip", "bb_resid.hip", "bb_profile.hip", "bb_mc.hip",
           "bb_norm.hip", "bb_mlp.hip", "bb_setup.hip", "bb_noise.hip"]
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # pixel and channel indices must round as in the FP64 restatement: no contraction of a*b+c
    "-ffp-contract=off",
    # no packed FP32 instructions in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 1)) as ex:
        codes = list(ex.map(lambda c: _run(c, verbose), cmds))
    bad = [s for s, c in zip(SOURCES, codes) if c != 0]
    if bad:
        raise RuntimeError(f"hipcc failed on {bad} building {LIB}")
    tmp = LIB + ".tmp"
    if _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose) != 0:
        raise RuntimeError(f"hipcc failed linking {LIB}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
