"""
Build recipe for the native parts of mars_amd (run from any directory):

  libmars_hip.so   hand-written HIP kernels for gfx950 + the C ABI of include/mars_hip.h
  _marslib<ext>    pybind11 host module (MarsAlgo over the C ABI), linked to libmars_hip.so

Both are written next to this file so they travel with the package.
"""

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["ortho.hip", "hinge.hip", "mixed.hip", "fit.hip", "forward.hip", "batch.hip", "post.hip"]
CXX_SOURCES = ["prune.cc"]  # host-only parts of libmars_hip
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MARS_HIP_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]

LIB = os.path.join(HERE, "libmars_hip.so")


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return out


def build_hip(jobs=8):
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for src in CXX_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-c", os.path.join(CSRC, src),
               "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write("\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l) + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", LIB])
    for o in objs:
        os.remove(o)
    return LIB


def build_pybind():
    import pybind11

    out = os.path.join(HERE, "_marslib" + ext_suffix())
    inc = [sysconfig.get_paths()["include"], pybind11.get_include(), os.path.join(HERE, "..", "include")]
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall"]
    cmd += ["-I" + i for i in inc]
    cmd += [os.path.join(CSRC, "marslib.cc"), "-L" + HERE, "-lmars_hip", "-Wl,-rpath,$ORIGIN", "-o", out]
    run(cmd)
    return out


def build():
    lib = build_hip()
    mod = build_pybind()
    return lib, mod


if __name__ == "__main__":
    for p in build():
        print(p)
