"""Builds libyamc_ewald.so in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", f)
       for f in ("yamc_kernels.hip", "yamc_forces.hip", "yamc_rdf.hip", "yamc_ops.hip", "yamc_api.hip")]
OUT = os.path.join(HERE, "libyamc_ewald.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
# -disable-machine-licm: the chain kernel is one long move loop; machine LICM
# hoists every FP64 literal of the loop body (erfc, exp, trapezoid
# coefficients) in front of it, where they are live for the whole move, get
# spilled and are reloaded from scratch, with a memory wait, at each use
# (145 spilled VGPRs with it, 12 without; 25 % more trial moves/s).
# -amdgpu-sched-strategy=max-ilp: the chain kernel's occupancy is fixed at two
# waves per SIMD, so the scheduler aims at instruction-level parallelism
# instead (about 1 % more trial moves/s, same results).
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-disable-machine-licm", "-mllvm", "-amdgpu-sched-strategy=max-ilp"]


def build(verbose: bool = False) -> str:
    """Compiles every source to an object (in parallel) and links the
    library; the objects go to a temporary directory."""
    import concurrent.futures as cf
    import tempfile
    tmp = OUT + ".tmp"
    compile_flags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="yamc_build_") as td:
        objs = [os.path.join(td, os.path.basename(src) + ".o") for src in SRC]

        def cc(src, obj):
            return subprocess.run([HIPCC, *compile_flags, "-c", "-o", obj, src], capture_output=True, text=True)

        with cf.ThreadPoolExecutor(max_workers=min(len(SRC), os.cpu_count() or 1)) as ex:
            results = list(ex.map(cc, SRC, objs))
        for src, r in zip(SRC, results):
            if r.returncode != 0:
                sys.stderr.write(r.stderr)
                raise RuntimeError(f"hipcc failed compiling {os.path.basename(src)}")
            if verbose:
                sys.stderr.write(r.stderr)
        r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                           capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise RuntimeError("hipcc failed linking libyamc_ewald.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
