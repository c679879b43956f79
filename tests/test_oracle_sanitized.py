"""The oracle's C++ (test infrastructure) under AddressSanitizer and
UndefinedBehaviorSanitizer: the same sources built by `make -C oracle
sanitized`, loaded through A3D_ORACLE_LIB in a child Python with the ASan
runtime preloaded, run the oracle's own tests.  A negative control passes a
buffer one element short and must be reported, so the check cannot pass with
the sanitizer silently off.  Host code only; no GPU."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "oracle", "_san", "liba3d_oracle.so")
ORACLE_TESTS = ["tests/test_oracle_bvh.py", "tests/test_oracle_clip.py", "tests/test_oracle_flood.py",
                "tests/test_oracle_ray.py"]


def _runtime():
    gcc = shutil.which("gcc")
    if gcc is None:
        return None
    path = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_env():
    asan = _runtime()
    if asan is None:
        pytest.skip("no ASan runtime for the host compiler")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitized", "-j8"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", A3D_ORACLE_LIB=SAN_LIB)
    return env


def _py(code, env, timeout=300):
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=timeout)


def test_sanitized_library_is_the_one_loaded(san_env):
    r = _py("import oracle; oracle.lib(); print(oracle.LIB_PATH);"
            "print(any('_san/liba3d_oracle.so' in l for l in open('/proc/self/maps')))", san_env)
    assert r.returncode == 0, r.stderr[-2000:]
    path, mapped = r.stdout.split()[-2:]
    assert path == SAN_LIB and mapped == "True"


def test_negative_control_is_reported(san_env):
    """orc_tri_closest_fp64 reads 9 doubles of the triangle: given 8, the read
    past the end is a heap overflow the sanitizer reports."""
    code = ("import ctypes, numpy as np, oracle\n"
            "t = np.ones(8, np.float64); p = np.zeros(3, np.float64); o = np.zeros(3, np.float64)\n"
            "oracle.lib().orc_tri_closest_fp64.argtypes = [ctypes.c_void_p] * 3\n"
            "oracle.lib().orc_tri_closest_fp64(t.ctypes.data, p.ctypes.data, o.ctypes.data)\n")
    r = _py(code, san_env)
    assert r.returncode != 0
    assert "AddressSanitizer" in r.stderr and "heap-buffer-overflow" in r.stderr


def test_oracle_tests_clean_under_sanitizers(san_env):
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *ORACLE_TESTS], cwd=ROOT,
                       env=san_env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout[-3000:] + r.stderr[-3000:])
    assert r.returncode == 0, tail
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
