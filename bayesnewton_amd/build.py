"""Build recipe for libbn_hip.so (gfx950).  `python -m bayesnewton_amd.build`.

Compiles the HIP sources in-tree with hipcc; the library lands in
bayesnewton_amd/_lib/ so it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libbn_hip.so")
SOURCES = ["bn_markov.hip", "bn_sites.hip", "bn_spacetime.hip", "bn_warmstart.hip", "bn_sparse.hip"]
ARCH = os.environ.get("BN_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions stay disabled (-packed-fp32-ops): this code is
# FP64-dominated anyway and the packed forms are not trusted on this platform.
# Matrix-core accumulators in VGPRs (-amdgpu-mfma-vgpr-form): with the default AGPR
# form the register allocator copied the 64 accumulators of the 128 x 128 FP64 GEMM
# between the two register files on every K panel.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(verbose=True, extra=()):
    """Compile every source to an object (in parallel, one hipcc per source), then
    link the shared library.  Each translation unit carries its own device code."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OUT_DIR, exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"] + list(extra)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        jobs.append(([hipcc()] + flags + ["-c", "-o", obj, os.path.join(CSRC, src)], obj))
    if verbose:
        for cmd, _ in jobs:
            print(" ".join(cmd), flush=True)
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(lambda j: subprocess.run(j[0], capture_output=True, text=True), jobs))
    for (cmd, _), res in zip(jobs, results):
        err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
        if err.strip():
            print(err, file=sys.stderr, flush=True)
        if res.returncode != 0:
            raise subprocess.CalledProcessError(res.returncode, cmd)
    tmp = LIB + ".tmp"
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [o for _, o in jobs]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(extra=sys.argv[1:])
