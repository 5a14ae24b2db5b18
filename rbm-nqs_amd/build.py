"""Builds the native parts in-tree:

  rbm-nqs_amd/lib/librbmgpu.so   HIP kernels + C ABI (include/rbmgpu.h), gfx950
  rbm-nqs_amd/lib/librbmhost.so  C++ host layer (include/rbmhost.h)
  rbm-nqs_amd/lib/rbm            the `rbm --train/--evaluate` driver

Run as `python rbm-nqs_amd/build.py`; every call recompiles.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-sign-compare"]

GPU_SOURCES = ["kern_f32.hip", "kern_f64.hip", "kern_struct.hip", "kern_learn.hip", "kern_eloc.hip",
               "kern_dense.hip", "kern_wave.hip", "rbmgpu_host.hip"]
HOST_SOURCES = ["lattice.cpp", "model.cpp", "ini.cpp", "setup.cpp", "optim.cpp", "capi.cpp"]
DRIVER_SOURCES = ["main.cpp", "adapters.cpp"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in GPU_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in HOST_SOURCES + [d for d in DRIVER_SOURCES if os.path.exists(os.path.join(HOST, d))]:
        o = os.path.join(OBJ, s.replace(".cpp", ".o"))
        jobs.append(["g++", *CXXFLAGS, "-c", os.path.join(HOST, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    gpu_objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in GPU_SOURCES]
    tmp = os.path.join(LIB, "librbmgpu.so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *gpu_objs, "-o", tmp])
    shutil.move(tmp, os.path.join(LIB, "librbmgpu.so"))
    host_objs = [os.path.join(OBJ, s.replace(".cpp", ".o")) for s in HOST_SOURCES]
    tmp = os.path.join(LIB, "librbmhost.so.tmp")
    _run(["g++", "-shared", "-fPIC", *host_objs, "-L" + LIB, "-lrbmgpu",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    shutil.move(tmp, os.path.join(LIB, "librbmhost.so"))
    drv = [os.path.join(OBJ, s.replace(".cpp", ".o")) for s in DRIVER_SOURCES
           if os.path.exists(os.path.join(HOST, s))]
    if len(drv) == len(DRIVER_SOURCES):
        tmp = os.path.join(LIB, "rbm.tmp")
        _run(["g++", *drv, "-L" + LIB, "-lrbmhost", "-lrbmgpu", "-Wl,-rpath,$ORIGIN", "-o", tmp])
        shutil.move(tmp, os.path.join(LIB, "rbm"))
    return LIB


def build_oracle():
    _run(["bash", os.path.join(ROOT, "oracle", "build.sh")])


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_oracle()
    print("built", LIB)
