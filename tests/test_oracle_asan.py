"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle asan):
a C driver of every walk option, tally and trace, then the oracle's own CPU tests with the
sanitizer build loaded in place of the plain one."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
BUILD = os.path.join(ORACLE, "_build")


@pytest.fixture(scope="module")
def asan_build():
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    return os.path.join(BUILD, "oracle_asan_check"), os.path.join(BUILD, "liboracle_asan.so")


def test_c_driver_is_clean(asan_build):
    exe, _ = asan_build
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1"))
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all cases ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_oracle_suite_under_sanitizers(asan_build):
    if os.environ.get("LD_PRELOAD"):
        pytest.skip("a preload is already set in this environment; the sanitizer runtime must come first")
    _, so = asan_build
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                             check=True).stdout.strip()
    assert os.path.isabs(libasan), "gcc has no libasan"
    env = dict(os.environ, ORACLE_SO=so, LD_PRELOAD=libasan,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    files = [os.path.join(HERE, f) for f in ("test_oracle_physics.py", "test_oracle_paths.py",
                                             "test_oracle_sources_tallies.py", "test_oracle_known_answers.py",
                                             "test_oracle_rng.py")]
    probe = ("from oracle import oracle as og; og.lib(); m = open('/proc/self/maps').read(); "
             "print('liboracle_asan.so' in m and 'libasan' in m)")
    r = subprocess.run([sys.executable, "-c", probe], capture_output=True, text=True, timeout=120, env=env, cwd=ROOT)
    assert r.stdout.strip() == "True", (r.stdout, r.stderr[-2000:])     # the sanitizer build is the one loaded
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *files],
                       capture_output=True, text=True, timeout=1200, env=env, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
