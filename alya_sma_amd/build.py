"""In-tree build of the HIP library (gfx950) and the Fortran host layer
(the test-side checkers under oracle/ build from their own Makefile).

    python -m alya_sma_amd.build

Outputs (git-ignored, shipped with the tree):
  alya_sma_amd/libalya_sma_amd.so     HIP kernels + C ABI
  alya_sma_amd/fortran/*.so, drivers  iso_c_binding host layer (if amdflang)
"""
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_hip():
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    out = os.path.join(HERE, "libalya_sma_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-I" + os.path.join(ROOT, "include"),
                                "-I" + os.path.join(HERE, "csrc"), "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_fortran():
    """iso_c_binding host layer and the Fortran bar driver."""
    fdir = os.path.join(HERE, "fortran")
    srcs = [os.path.join(fdir, f) for f in ("mod_sld_dev.f90", "sld_stress_model_160_dev.f90")]
    if not os.path.exists(FLANG) or not all(os.path.exists(s) for s in srcs):
        return None
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    lib = os.path.join(HERE, "libalya_sma_amd.so")
    flib = os.path.join(fdir, "libalya_sma_f.so")
    _run([FLANG, "-O2", "-fPIC", "-shared", "-J" + bdir, "-o", flib] + srcs +
         ["-L" + HERE, "-lalya_sma_amd", "-Wl,-rpath,$ORIGIN/.."])
    for name in ("sld_bar_driver", "sld_law_driver"):
        drv = os.path.join(fdir, name + ".f90")
        if os.path.exists(drv):
            _run([FLANG, "-O2", "-J" + bdir, "-o", os.path.join(fdir, name), drv,
                  "-L" + fdir, "-lalya_sma_f", "-L" + HERE, "-lalya_sma_amd",
                  "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."])
    shutil.rmtree(bdir, ignore_errors=True)
    return flib


def build():
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build()
    print("built", os.path.join(HERE, "libalya_sma_amd.so"))
