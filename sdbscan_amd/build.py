"""Builds the HIP library and the pybind11 drop-in module in-tree.

    python -m sdbscan_amd.build      (or: python sdbscan_amd/build.py)

Outputs (next to this file):
  libsdbscan_amd.so          -- HIP kernels + C ABI (include/sdbscan_amd.h), gfx950
  sDbscan.cpython-*.so       -- pybind11 module `sDbscan` wrapping the C++ class
  sdbscan                    -- command-line front end (the reference's main.cpp)
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["sdb_index.hip", "sdb_dist.hip", "sdb_cluster.hip", "sdb_optics.hip", "sdb_noise.hip", "sdb_sng.hip", "sdb_capi.hip", "sdb_batch.hip", "sdb_sweep.hip", "sdb_sweep_kernels.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SDB_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    # packed FP32 instructions disabled on the device
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(jobs=4):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c",
                                     os.path.join(CSRC, src), "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise SystemExit("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise SystemExit("hipcc failed")
    out = os.path.join(HERE, "libsdbscan_amd.so")
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out] + objs)
    return out


def build_module():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "sDbscan" + ext)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(CSRC, "sDbscan.cpp"), os.path.join(CSRC, "pybind_module.cpp"),
           "-L", HERE, "-lsdbscan_amd", "-Wl,-rpath,$ORIGIN", "-o", out]
    run(cmd)
    return out


def build_cli():
    out = os.path.join(HERE, "sdbscan")
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(CSRC, "sDbscan.cpp"), os.path.join(CSRC, "sdb_main.cpp"),
           "-L", HERE, "-lsdbscan_amd", "-Wl,-rpath,$ORIGIN", "-o", out]
    run(cmd)
    return out


def main():
    build_lib()
    build_module()
    build_cli()


if __name__ == "__main__":
    sys.exit(main())
