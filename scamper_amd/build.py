"""Build libscamper_hip.so in-tree with hipcc for gfx950.

    python -m scamper_amd.build        (or scamper_amd.build.build())

Every HIP source is compiled with packed-FP32 instructions disabled; the
threshold tables, the cluster kernels (f64 O(3) bond probabilities and
reflections) and the exchange energies are compiled with -ffp-contract=off
so their f64 arithmetic rounds like the reference's.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libscamper_hip.so")
ARCH = os.environ.get("SCAMPER_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
PER_FILE = {
    "tables.cpp": ["-ffp-contract=off"],
    "capi.cpp": ["-ffp-contract=off"],
    "carlo.cpp": ["-ffp-contract=off"],
    "pool.cpp": ["-ffp-contract=off"],
    "k_cluster.hip": ["-ffp-contract=off"],
    "k_csr_real.hip": ["-ffp-contract=off"],
    "k_exchange.hip": ["-ffp-contract=off"],
}


def sources() -> list[str]:
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src + ".o")
    cmd = ["hipcc", *COMMON, *PER_FILE.get(src, []), "-x", "hip", "-c",
           os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = [ln for ln in (r.stderr or "").splitlines()
             if "packed-fp32-ops" not in ln and ln.strip()]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n" + "\n".join(noise))
    return obj


def build(verbose: bool = True) -> str:
    if shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources for {ARCH}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
