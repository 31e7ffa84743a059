"""Build libprtlb_hip.so in-tree for gfx950 with hipcc.

    python -m prtlb.build        (from prt-deeponet-gui_amd/)

Each .hip file is compiled on its own so that the transport kernels get
-ffp-contract=off (they must reproduce float32 NumPy bit for bit) while the
lattice-Boltzmann kernels keep contraction where their source allows it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OUT = os.path.join(HERE, "libprtlb_hip.so")
ARCH = os.environ.get("PRTLB_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result"]
SOURCES = {
    "prt_common.hip": [],
    "prt_lbm.hip": [],
    "prt_sparse.hip": [],
    "prt_adr.hip": ["-ffp-contract=off"],
    "prt_reduce.hip": ["-ffp-contract=off"],
    "prt_batch.hip": [],
    "prt_batch_adr.hip": ["-ffp-contract=off"],
    "prt_geom.hip": ["-ffp-contract=off"],
    "prt_check.hip": ["-ffp-contract=off"],
    "prt_handle.hip": [],
    "prt_features.hip": ["-ffp-contract=off"],
}


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    """Compile every source (in parallel, one hipcc per file) and link the
    library; returns its path."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src, extra in SOURCES.items():
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmds.append([cc, "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + extra)
        objs.append(obj)

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = OUT + ".tmp"
    cmd = [cc, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
