"""Builds the C++ facade client tests/cpp/facade_client.cc with g++ against
include/atmo_model.h and libatmo.so (the way a C++ caller of the drop-in
would), into tests/cpp/facade_client."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "precomputed_atmospheric_scattering_amd")
SRC = os.path.join(HERE, "cpp", "facade_client.cc")
EXE = os.path.join(HERE, "cpp", "facade_client")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("atmo.h", "atmo_model.h")]


def build(out=EXE):
    if not os.path.exists(os.path.join(PKG, "libatmo.so")):
        raise RuntimeError("libatmo.so is not built (python -m "
                           "precomputed_atmospheric_scattering_amd.build)")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), SRC, "-o", out + ".tmp",
           "-L", PKG, "-latmo",
           # The client finds libatmo.so next to it in the tree wherever the
           # tree is; the HIP runtime libatmo.so needs comes from ROCm.
           "-Wl,-rpath,$ORIGIN/../../precomputed_atmospheric_scattering_amd",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade client build failed: %s\n%s" % (" ".join(cmd),
                                                                   r.stdout + r.stderr))
    os.replace(out + ".tmp", out)
    return out


def client():
    """The built client, rebuilt if missing or older than its sources."""
    srcs = [SRC, os.path.join(PKG, "libatmo.so")] + HEADERS
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < max(map(os.path.getmtime, srcs)):
        build()
    return EXE
