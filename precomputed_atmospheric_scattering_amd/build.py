"""Builds libatmo.so (HIP kernels for gfx950 + the C ABI + the C++ Model
facade) in-tree with hipcc.

    python -m precomputed_atmospheric_scattering_amd.build
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libatmo.so")
# Measurement tool of bench.py (on-device rate ceilings); not part of the
# library's interface.
PEAKS_LIB = os.path.join(PKG, "libatmo_peaks.so")

# -packed-fp32-ops: the compiler emits no packed FP32 instructions
# (v_pk_add/mul/fma_f32), some forms of which have given wrong results on
# MI355X without an error.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]

# The precompute kernels contract no multiply-add by themselves: each FMA in
# them is written as fma()/fmaf(). Implicit contraction depends on how many
# uses a product has, which differs between the tile-shape instantiations of
# one kernel (several nu outputs per thread share products), and the tile
# shapes must give bit-identical LUTs.
SOURCES = [
    ("precompute.hip", True, ["-ffp-contract=off"]),
    ("query.hip", True, []),
    ("runtime.cpp", True, []),   # host code using the HIP runtime
    ("server.cpp", True, []),    # serving loop (worker thread, HIP runtime)
    ("lut_io.cpp", True, []),    # LUT files (export, import, cached Init)
    ("autotune.cpp", True, []),  # tile-shape autotuner
    ("model.cc", False, []),     # C++ Model facade, host only
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: ROCm is required to build libatmo")
    return exe


def _compile(args):
    src, obj, is_hip, extra = args
    cmd = [hipcc()] + (["-x", "hip"] if is_hip else []) + HIP_FLAGS + extra + [
        "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    with tempfile.TemporaryDirectory(prefix="atmo_build_") as tmp:
        jobs = []
        for name, is_hip, extra in SOURCES:
            src = os.path.join(CSRC, name)
            if not os.path.exists(src):
                continue
            jobs.append((src, os.path.join(tmp, name + ".o"), is_hip, extra))
        with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
            outs = list(ex.map(_compile, jobs))
        if verbose:
            for o in outs:
                if o.strip():
                    print(o)
        tmp_lib = LIB + ".tmp%d" % os.getpid()
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib] + [
            j[1] for j in jobs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
        os.replace(tmp_lib, LIB)
        tmp_peaks = PEAKS_LIB + ".tmp%d" % os.getpid()
        cmd = [hipcc(), "-x", "hip"] + HIP_FLAGS + ["-shared", "-o", tmp_peaks,
                                                    os.path.join(CSRC, "peaks.hip")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
        os.replace(tmp_peaks, PEAKS_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
