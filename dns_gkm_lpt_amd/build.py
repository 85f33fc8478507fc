"""Build the gfx950 HIP library libdnsgkm.so in-tree with hipcc.

    python -m dns_gkm_lpt_amd.build [--resource-usage]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdnsgkm.so")
SOURCES = ["gkm_kernels.hip", "ens_persist.hip", "lpt_kernels.hip", "lpt_sort.hip", "adj_kernels.hip",
           "gkm_api.hip", "gkm_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) is kept off for every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# Per-source flags.  The sweep kernels keep their uniform values (kernel
# arguments, per-variable base addresses, FP64 constants) in scalar registers;
# with machine-level loop-invariant code motion on, hoisting the constants'
# materialisation out of the chunk loop overflowed the 102 SGPRs, and the
# spilled SGPRs were reloaded inside the loop with v_readlane (a VALU
# instruction each, ~100 per face in the compensated z sweep).  Without it the
# constants are rematerialised in the loop by scalar moves.
SOURCE_FLAGS = {"gkm_kernels.hip": ["-mllvm", "-disable-machine-licm"]}


def build(resource_usage=False, verbose=False, jobs=None):
    """Compile every source (in parallel, `jobs` at a time) and link the library."""
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if resource_usage:
            cmd.append("-Rpass-analysis=kernel-resource-usage")
        cmds.append(cmd)
        objs.append(obj)
    jobs = jobs or min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1)))
    running = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd))
        while len(running) >= jobs:
            _wait_first(running)
    while running:
        _wait_first(running)
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def _wait_first(running):
    proc = running.pop(0)
    if proc.wait() != 0:
        for other in running:
            other.wait()
        raise subprocess.CalledProcessError(proc.returncode, proc.args)


if __name__ == "__main__":
    build(resource_usage="--resource-usage" in sys.argv, verbose=True)
