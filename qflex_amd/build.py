"""Builds the native parts of qflex_amd in-tree.

* ``qflex_amd/lib/libqfx.so``  -- HIP kernels + the C-ABI of include/qfx.h,
  compiled by hipcc for gfx950 only.
* ``qflex_amd/qflex*.so``      -- the C++ host (Tensor, ContractionData,
  EvaluateCircuit, parsers) and the pybind11 module ``qflex`` whose
  ``simulate(dict)`` is the reference's Python entry point.  It is plain C++
  (g++) and reaches the GPU only through libqfx.so's C-ABI.
"""
import concurrent.futures
import os
import subprocess
import sys
import sysconfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qflex_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["qfx_runtime.hip", "qfx_permute.hip", "qfx_cgemm.hip", "qfx_mlp.hip", "qfx_fp8.hip"]
HOST_SOURCES = ["errors.cpp", "tensor.cpp", "contraction_utils.cpp", "read_circuit.cpp",
                "evaluate_circuit.cpp", "circuit.cpp", "grid.cpp", "ordering.cpp"]
PYBIND_SOURCE = "pybind_main.cpp"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
             # keep the compiler from emitting packed FP32 VALU ops
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I" + INCLUDE]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-fmacro-prefix-map=" + ROOT + "/=",
             "-I" + INCLUDE, "-I" + HOST]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def _hip_obj(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _cxx_obj(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    extra = []
    if src == PYBIND_SOURCE:
        import pybind11
        extra = ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"]]
    _run([CXX] + CXX_FLAGS + extra + ["-c", os.path.join(HOST, src), "-o", obj])
    return obj


def build_capi():
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_hip_obj, HIP_SOURCES))
    out = os.path.join(LIBDIR, "libqfx.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_cxx_obj, srcs + [PYBIND_SOURCE]))
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "qflex" + suffix)
    _run([CXX, "-shared", "-fPIC", "-o", out + ".tmp"] + objs +
         ["-L" + LIBDIR, "-lqfx", "-Wl,-rpath,$ORIGIN/lib"])
    os.replace(out + ".tmp", out)
    return out


def install_into_package(package_dir):
    """Copies the built module qflex (and lib/libqfx.so, which it loads from
    lib/ beside itself) into `package_dir`, e.g. an installed qflexcirq
    package, so that the reference's ``from qflexcirq import qflex``
    (qflexcirq/interface/qflex_simulator.py:5) imports this engine unchanged.
    Returns the module's path there."""
    import glob
    import shutil
    mods = glob.glob(os.path.join(PKG, "qflex" + sysconfig.get_config_var("EXT_SUFFIX")))
    lib = os.path.join(LIBDIR, "libqfx.so")
    if not mods or not os.path.exists(lib):
        raise RuntimeError("build first: qflex module or libqfx.so missing")
    os.makedirs(os.path.join(package_dir, "lib"), exist_ok=True)
    shutil.copy2(lib, os.path.join(package_dir, "lib", "libqfx.so"))
    dst = os.path.join(package_dir, os.path.basename(mods[0]))
    shutil.copy2(mods[0], dst)
    return dst


def build_all():
    lib = build_capi()
    mod = build_host() if os.path.exists(os.path.join(HOST, PYBIND_SOURCE)) else None
    return lib, mod


if __name__ == "__main__":
    print(build_all())
    sys.exit(0)
