"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU):
the FP64 oracle (tests/sanitize/oracle_driver.cc: Init at 3 and 15
wavelengths, the passes on texel subsets, the sky lookups) and libatmo's
host code (tests/sanitize/capi_driver.cc: argument validation, parameter
lowering, illuminance iterations, spectrum conversion, FP16 scales, tile
table save/load with malformed files). Built by tests/sanitize/Makefile;
any sanitizer report aborts the driver."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
ENV = dict(os.environ,
           # The HIP runtime keeps allocations until exit, by design.
           ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-j8", "-C", HERE], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("driver", ["oracle_driver", "capi_driver"])
def test_driver_clean_under_sanitizers(built, driver, tmp_path):
    r = subprocess.run([os.path.join(HERE, driver), str(tmp_path)], capture_output=True,
                       text=True, timeout=300, env=ENV)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0 and "0 failures" in r.stdout, out[-4000:]
