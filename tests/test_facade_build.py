"""The C++ facade client compiles with -Wall -Wextra -Werror against
include/atmo_model.h and links against libatmo.so (no GPU needed; the GPU
test tests/test_gpu_facade.py runs it)."""
import os
import subprocess

import facade_build


def test_client_compiles_and_links(tmp_path):
    exe = facade_build.build(str(tmp_path / "facade_client"))
    assert os.path.exists(exe)
    # Usage errors come back before any device call. (Built outside the
    # tree, the client finds libatmo.so through LD_LIBRARY_PATH.)
    env = dict(os.environ, LD_LIBRARY_PATH=facade_build.PKG)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 2 and "usage" in r.stderr
