"""Build libcachip.so (HIP for gfx950) in-tree with hipcc.

Every .hip source under cac_amd/csrc is compiled to an object in parallel and
linked into cac_amd/libcachip.so, which exports the C ABI of include/cac_hip.h.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libcachip.so")
ARCH = os.environ.get("CAC_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CXXFLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 ops stay off: they are not needed and some forms misbehave on this part
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # contract a * b + c only inside one expression, so results do not depend
    # on unrolling or on the launch shape
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CXXFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    if r.stderr.strip():
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        if err.strip():
            sys.stderr.write(err + "\n")
    return obj


HOST = os.path.join(HERE, "host")
HOSTLIB = os.path.join(HERE, "libcachost.so")
RUNNER = os.path.join(HERE, "cac_run")
HOSTFLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", f"-I{os.path.join(ROOT, 'include')}"]


def build_host(verbose: bool = False) -> None:
    """The C++ host layer (Pair / Neighbor / FixNVE / Verlet mirrors, data-file
    reader) as libcachost.so over libcachip.so, and the cac_run driver."""
    steps = [
        (["g++", *HOSTFLAGS, "-shared", "-o", HOSTLIB + ".tmp", os.path.join(HOST, "cac_host.cpp"),
          os.path.join(HOST, "cac_dump.cpp"), f"-L{HERE}",
          "-lcachip", "-Wl,-rpath,$ORIGIN"], HOSTLIB),
        (["g++", *HOSTFLAGS, "-o", RUNNER + ".tmp", os.path.join(HOST, "cac_run.cpp"), f"-L{HERE}", "-lcachost",
          "-lcachip", "-Wl,-rpath,$ORIGIN"], RUNNER),
    ]
    for cmd, out in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
        os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {HOSTLIB} and {RUNNER}")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
