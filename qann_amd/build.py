"""Builds the native parts of qann_amd in place (no cmake, no JIT cache):

* ``qann_amd/libqann_hip.so`` -- the HIP kernels and the C ABI of
  include/qann_hip.h, compiled by hipcc for gfx950;
* ``qann_amd/_qann*.so`` -- the host index objects and the pybind11 module,
  compiled by g++ and linked against libqann_hip.so.

Run ``python qann_amd/build.py`` (or ``__graft_entry__.build()``); it does not import
the package, which needs the libraries it builds.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys
import sysconfig

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
HIP_SRC = [PKG / "csrc" / "hip" / n for n in ("ivf_kernels.hip", "gemm_bf16.hip", "build_kernels.hip", "capi.hip")]
HOST_SRC = [PKG / "csrc" / "host" / n for n in ("qfile.cpp", "metadata.cpp", "train.cpp", "indexes.cpp", "serve.cpp", "bindings.cpp")]
HIP_LIB = PKG / "libqann_hip.so"
EXT = PKG / ("_qann" + sysconfig.get_config_var("EXT_SUFFIX"))
ARCH = os.environ.get("QANN_AMD_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add in the kernels is an explicit fmaf,
# so results keep the bits of the reference's CPU arithmetic.
# -packed-fp32-ops: the compiler emits no packed FP32 instructions (v_pk_add_f32,
# v_pk_mul_f32, v_pk_fma_f32); every FP32 operation is a scalar VALU one.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd[:3]))} ...")
    return r


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(h).exists():
        raise RuntimeError("hipcc not found")
    return h


def build_hip(jobs: int = 2):
    objs = []
    tmp = PKG / "csrc" / "_obj"
    tmp.mkdir(exist_ok=True)
    procs = []
    for src in HIP_SRC:
        obj = tmp / (src.stem + ".o")
        objs.append(obj)
        procs.append(subprocess.Popen([_hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)],
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    for p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("hipcc failed")
    out = HIP_LIB.with_suffix(".so.tmp")
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(out), *map(str, objs)])
    os.replace(out, HIP_LIB)


def build_host():
    import pybind11

    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    out = EXT.with_name(EXT.name + ".tmp")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-pthread", "-fvisibility=hidden", *inc,
          *map(str, HOST_SRC), "-o", str(out), f"-L{PKG}", "-lqann_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(out, EXT)


def build_oracle():
    """The CPU checkers (test infrastructure): liboracle.so, and the reference
    library where its sources are present."""
    oracle = ROOT / "oracle"
    targets = ["oracle"]
    if pathlib.Path("/root/reference/src").is_dir():
        targets.append("ref")
    _run(["make", "-C", str(oracle), "-j8", *targets])


def build_all():
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", HIP_LIB.name, EXT.name)
