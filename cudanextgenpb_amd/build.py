"""In-tree build of the native libraries.

* ``lib/libngpb.so``          — HIP kernels for gfx950 + the C ABI (include/ngpb.h)
* ``lib/libngpb_workload.so`` — host-only synthetic workload generator
* ``oracle/_build/libngpb_oracle.so`` — the CPU oracle (test infrastructure), via
  ``oracle/Makefile``

Run ``python -m cudanextgenpb_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

from . import _paths

CSRC = os.path.join(_paths.PKG_DIR, "csrc")
INCLUDE = os.path.join(_paths.REPO_DIR, "include")
BUILD_DIR = os.path.join(_paths.PKG_DIR, "_build")

HIP_SOURCES = [
    "ngpb_api.hip",
    "ngpb_scan.hip",
    "ngpb_markers.hip",
    "ngpb_assemble.hip",
    "ngpb_krylov.hip",
    "ngpb_lis.hip",
    "ngpb_mixed.hip",
    "ngpb_energy.hip",
    "ngpb_batch.hip",
    "ngpb_rays.hip",
    "ngpb_ses.hip",
    "ngpb_locate.hip",
    "ngpb_octree.hip",
]

# -ffp-contract=off: every multiply and add is rounded on its own unless the source
# writes fma() explicitly, so the kernels round exactly like the host code they are
# compared with. packed-fp32-ops is disabled for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_workload():
    os.makedirs(_paths.LIB_DIR, exist_ok=True)
    out = _paths.workload_lib()
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-Wall",
          "-o", out + ".tmp", os.path.join(CSRC, "workload.cpp")])
    os.replace(out + ".tmp", out)


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(_paths.REPO_DIR, "oracle")])


def build_native(jobs: int = 8):
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(_paths.LIB_DIR, exist_ok=True)
    cc = hipcc()
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        _run([cc, *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = _paths.native_lib()
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)


def build_all():
    build_workload()
    build_oracle()
    build_native()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "workload": build_workload, "oracle": build_oracle, "native": build_native}[w]()
    print("built:", ", ".join(what))
