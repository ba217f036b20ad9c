"""Build libmfmhd.so (HIP kernels for gfx950 + the C ABI) in-tree.

Each translation unit is compiled with its own floating-point rules:
  mfm_kernels_fp64.hip, mfm_init.hip, mfm_fill.hip, mfm_sync.hip :
                                       -ffp-contract=off
                                       (reference rounding)
  mfm_stage_fp64.hip                 : -ffp-contract=off
  mfm_kernels_fp32.hip, mfm_stage_fp32.hip
                                     : FMA contraction, fast FP32 div/sqrt,
                                       FP32 denormals flushed (a division
                                       is then one reciprocal and a multiply)
All HIP code is built with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmfmhd.so")
ARCH = os.environ.get("MFM_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-function",
]
UNITS = [
    ("mfm_kernels_fp64.hip", ["-ffp-contract=off"]),
    ("mfm_kernels_fp32.hip", ["-ffp-contract=fast",
                              "-fno-hip-fp32-correctly-rounded-divide-sqrt",
                              "-fgpu-flush-denormals-to-zero"]),
    ("mfm_stage_fp64.hip", ["-ffp-contract=off"]),
    ("mfm_stage_fp32.hip", ["-ffp-contract=fast",
                            "-fno-hip-fp32-correctly-rounded-divide-sqrt",
                            "-fgpu-flush-denormals-to-zero"]),
    ("mfm_init.hip", ["-ffp-contract=off"]),
    ("mfm_capi.cpp", ["-ffp-contract=off"]),
    ("mfm_serve.cpp", ["-ffp-contract=off"]),
    ("mfm_fill.hip", ["-ffp-contract=off"]),
    ("mfm_sync.hip", ["-ffp-contract=off"]),
    ("mfm_fill_plan.cpp", ["-ffp-contract=off"]),
    ("mfm_chombo.cpp", []),
    ("mfm_noalloc.hip", ["-ffp-contract=off"]),
    ("mfm_ceiling.hip", []),
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    lines = [l for l in p.stdout.splitlines()
             if "packed-fp32-ops" not in l or "error" in l]
    return p.returncode, "\n".join(lines)


def build(verbose=False):
    """Compile every unit and link libmfmhd.so; raises on failure."""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src, extra in UNITS:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        jobs.append((obj, [hipcc] + COMMON + extra +
                     ["-c", os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda j: _run(j[1]), jobs))
    for (obj, cmd), (rc, log) in zip(jobs, results):
        if verbose and log.strip():
            print(log)
        if rc != 0:
            raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + log)
    tmp = OUT + ".tmp"
    rc, log = _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-o", tmp] +
                   [o for o, _ in jobs])
    if rc != 0:
        raise RuntimeError("link failed:\n" + log)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
