"""Build recipe of the native parts, run in-tree.

  libcampfire_hip.so   HIP kernels + C ABI (hipcc, --offload-arch=gfx950)
  js/campfire_addon.node  N-API addon over the C ABI (g++, Node headers)

usage: python -m campfire_amd.build [--lib] [--addon]
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcampfire_hip.so")
ADDON = os.path.join(HERE, "js", "campfire_addon.node")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CF_OFFLOAD_ARCH", "gfx950")

# No FMA contraction (keeps the reference's rounding order), no fast-math,
# packed-FP32 instructions disabled.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-command-line-argument"]
SOURCES = ["cf_kernels.hip", "cf_advect.hip", "cf_mg.hip", "cf_flame1d.hip", "cf_sim.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return err


def build_lib():
    # the translation units compile in parallel (one hipcc each)
    objs = [os.path.join(CSRC, src.replace(".hip", ".o")) for src in SOURCES]
    procs = [subprocess.Popen([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
             for src, obj in zip(SOURCES, objs)]
    failed = []
    for src, p in zip(SOURCES, procs):
        _, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write("\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l) + "\n")
            failed.append(src)
    if failed:
        raise RuntimeError("build failed: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_addon():
    """N-API addon (N-API v6+, Node 12 headers).  Skipped when Node headers are absent."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    src = os.path.join(HERE, "js", "addon.cc")
    if not os.path.exists(src):
        return None
    tmp = ADDON + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
          "-DNAPI_VERSION=6", src, "-o", tmp, "-L" + HERE, "-lcampfire_hip",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,--allow-shlib-undefined"])
    os.replace(tmp, ADDON)
    return ADDON


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def build_all():
    build_lib()
    build_oracle()
    build_addon()


if __name__ == "__main__":
    args = sys.argv[1:]
    if not args or "--lib" in args:
        print(build_lib())
    if not args or "--oracle" in args:
        build_oracle()
    if not args or "--addon" in args:
        print(build_addon())
