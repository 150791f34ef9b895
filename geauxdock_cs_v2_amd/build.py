"""Build libgeauxdock_amd.so in-tree.

Host C++ (loader, staging, C ABI host half) is compiled with g++ -- the same
compiler and libm the reference's CPU code uses, so staged tables round the
same way.  HIP sources are compiled with hipcc for gfx950 only.  Packed FP32
instructions are disabled for all device code (they are not used by hand and
hipcc is told not to form them).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libgeauxdock_amd.so")
PHASE_LIB = os.path.join(PKG, "libgeauxdock_amd_phase.so")
APP = os.path.join(PKG, "gd_dock")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HOST_FLAGS = ["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", f"-I{INC}"]
DEV_FLAGS = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", f"-I{INC}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # loop heads on 64-byte instruction-cache lines (the pair loops
             # measured 0.2-0.3 % faster)
             "-falign-loops=64",
             "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False, jobs=8):
    os.makedirs(os.path.join(BUILD, "host"), exist_ok=True)
    os.makedirs(os.path.join(BUILD, "dev"), exist_ok=True)
    host_src = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    dev_src = sorted(glob.glob(os.path.join(CSRC, "device", "*.hip")))
    cmds, objs = [], []
    for s in host_src:
        o = os.path.join(BUILD, "host", os.path.basename(s) + ".o")
        cmds.append(HOST_FLAGS + ["-c", s, "-o", o])
        objs.append(o)
    for s in dev_src:
        o = os.path.join(BUILD, "dev", os.path.basename(s) + ".o")
        cmds.append(DEV_FLAGS + ["-I", os.path.join(CSRC, "device"), "-c", s, "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    # the docking driver (the reference's 3_geauxdock over the C ABI), next
    # to the library it loads through $ORIGIN
    tmp = APP + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INC}", os.path.join(CSRC, "app", "gd_dock.cpp"),
          f"-L{os.path.dirname(LIB)}", "-lgeauxdock_amd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib",
          "-o", tmp])
    os.replace(tmp, APP)
    return LIB


def build_phase_profile(jobs=8):
    """The phase-profile build of the library (device code compiled with
    -DGD_PHASE_PROFILE: clock marks per phase of the MC step, exported through
    gd_phase_profile), beside the product library; phase_profile.py loads it.
    The product library does not contain the marks."""
    os.makedirs(os.path.join(BUILD, "phase"), exist_ok=True)
    host_objs = [os.path.join(BUILD, "host", os.path.basename(s) + ".o")
                 for s in sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))]
    missing = [o for o in host_objs if not os.path.exists(o)]
    if missing:
        build(jobs=jobs)
    cmds, objs = [], []
    for s in sorted(glob.glob(os.path.join(CSRC, "device", "*.hip"))):
        o = os.path.join(BUILD, "phase", os.path.basename(s) + ".o")
        cmds.append(DEV_FLAGS + ["-DGD_PHASE_PROFILE", "-I", os.path.join(CSRC, "device"), "-c", s, "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = PHASE_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + host_objs + objs + ["-lpthread"])
    os.replace(tmp, PHASE_LIB)
    return PHASE_LIB


def build_oracle():
    """CPU oracle (test infrastructure) and, where the reference tree exists,
    the reference's own CPU kernel under oracle/_ref."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_phase_profile())
    build_oracle()
    sys.exit(0)
