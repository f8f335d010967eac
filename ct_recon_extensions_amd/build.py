"""Build the gfx950 shared libraries in-tree with hipcc.

    python -m ct_recon_extensions_amd.build          # libctrecon_hip.so + libnumpy_ctext.so

Outputs (git-ignored, shipped with the source tree):
    ct_recon_extensions_amd/lib/libctrecon_hip.so   C ABI (include/ctr_api.h) and the
                                                    ct_recon:: class API (include/*.h)
    ct_recon_extensions_amd/lib/libnumpy_ctext.so   the reference's numpy C ABI
                                                    (include/numpy_ctext.h)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
ARCH = os.environ.get("CTR_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Packed FP32 ops are disabled for all device code: hipcc -O3 would otherwise
# form v_pk_{add,mul,fma}_f32 by itself.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", f"-I{ROOT}", f"-I{CSRC}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

CORE_SOURCES = ["tables.cpp", "capi.cpp", "classes.cpp", "fan_bp.hip", "par_bp.hip",
                "fp_par.hip", "filter.hip", "ragged.hip", "toeplitz.hip", "par_svbp.hip",
                "fp_rc.hip", "fp_dd.hip", "helical.hip"]
NUMPY_SOURCES = ["npext.cpp"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *DEVICE_FLAGS, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "hip", *DEVICE_FLAGS, "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def _link(objs, out, extra=()):
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs, *extra]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in CORE_SOURCES + NUMPY_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = dict(zip(sources, ex.map(_compile, sources)))
    core = os.path.join(LIB, "libctrecon_hip.so")
    _link([objs[s] for s in CORE_SOURCES if s in objs], core)
    if "npext.cpp" in objs:
        _link([objs["npext.cpp"]], os.path.join(LIB, "libnumpy_ctext.so"),
              [f"-L{LIB}", "-lctrecon_hip", "-Wl,-rpath,$ORIGIN"])
    # C++ driver of the class API for tests/test_class_api.py
    drv_src = os.path.join(ROOT, "tests", "cpp", "class_api_main.cpp")
    if os.path.exists(drv_src):
        drv = os.path.join(ROOT, "tests", "cpp", "class_api_main")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", f"-I{ROOT}", drv_src,
               "-o", drv, f"-L{LIB}", "-lctrecon_hip", f"-Wl,-rpath,{LIB}",
               "-Wl,-rpath,$ORIGIN/../../ct_recon_extensions_amd/lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"class API driver build failed:\n{res.stderr}")
    if verbose:
        print(f"built {core} for {ARCH}")


if __name__ == "__main__":
    build()
    sys.exit(0)
