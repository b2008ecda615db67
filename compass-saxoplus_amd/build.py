"""Build the native parts in-tree with hipcc for gfx950.

Products (all next to this file, git-ignored, shipped with the tree):
  lib/libshwfs.so        C ABI + HIP kernels        (include/shwfs.h)
  lib/libsutra_amd.so    C++ carma / sutra classes over the C ABI
  carma.<ext>, sutra.<ext>  pybind11 modules with the reference's names

Usage: python compass-saxoplus_amd/build.py [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("SHWFS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["shwfs_capi.hip", "shwfs_image_fft.hip", "shwfs_image_dft.hip",
               "shwfs_dft_bf16x3_q1.hip", "shwfs_dft_bf16x3_q2.hip", "shwfs_dft_fp16_q1.hip",
               "shwfs_dft_fp16_q2.hip", "shwfs_batch.hip", "shwfs_batch_quiet.hip",
               "shwfs_batch_noisy.hip", "shwfs_mlp.hip", "shwfs_raytrace.hip",
               "shwfs_image_ext.hip", "shwfs_batch_groups.hip", "shwfs_corr.hip",
               "shwfs_field_stop.hip", "shwfs_lgs.hip", "shwfs_dft_mlp.hip"]
CPP_SOURCES = ["carma/carma.cpp", "sutra/sutra.cpp", "sutra/sutra_centroider.cpp",
               "sutra/sutra_batch.cpp", "sutra/sutra_centroider_mlp.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# Packed FP32 instructions are disabled for every device compilation.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def compile_one(src, device):
    obj = os.path.join(OBJ, src.replace("/", "_") + ".o")
    cmd = [hipcc()] + COMMON + (DEVICE if device else ["-x", "c++", "-D__HIP_PLATFORM_AMD__"])
    if not device:
        cmd += ["-I" + os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include")]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    run(cmd)
    return obj


def pybind_flags():
    import pybind11
    return ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
            "-fvisibility=hidden"]


def build(jobs=None):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        fut_dev = [ex.submit(compile_one, s, True) for s in HIP_SOURCES]
        fut_host = [ex.submit(compile_one, s, False) for s in CPP_SOURCES]
        dev_objs = [f.result() for f in fut_dev]
        host_objs = [f.result() for f in fut_host]
    libshwfs = os.path.join(LIB, "libshwfs.so")
    run([hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", libshwfs] + dev_objs)
    libsutra = os.path.join(LIB, "libsutra_amd.so")
    run([hipcc(), "-shared", "-fPIC", "-o", libsutra] + host_objs +
        ["-L" + LIB, "-lshwfs", "-Wl,-rpath,$ORIGIN"])
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    mods = []
    with cf.ThreadPoolExecutor(2) as ex:
        for name in ("carma", "sutra"):
            src = os.path.join(CSRC, "python", name + "_module.cpp")
            out = os.path.join(HERE, name + ext)
            cmd = [hipcc(), "-x", "c++", "-D__HIP_PLATFORM_AMD__",
                   "-I" + os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include")] + \
                COMMON + pybind_flags() + [
                "-shared", src, "-o", out, "-L" + LIB, "-lsutra_amd", "-lshwfs",
                "-Wl,-rpath,$ORIGIN/lib"]
            mods.append(ex.submit(run, cmd))
        for m in mods:
            m.result()
    return [libshwfs, libsutra]


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    for p in build(a.j):
        print("built", p)
    sys.exit(0)
