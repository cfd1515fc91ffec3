"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code
only): builds oracle/liboracle_asan.so (`make -C oracle asan`) and runs the
oracle's pin tests against it in a child process, with the ASan runtime
preloaded ahead of the (uninstrumented) Python interpreter.  Any sanitizer
report fails the child: UBSan halts on the first error, ASan aborts on its
first report."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
ASAN_LIB = os.path.join(ORACLE, "liboracle_asan.so")
PINS = ["test_oracle_pins.py", "test_oracle_energy_pins.py", "test_oracle_pole_pins.py",
        "test_oracle_variance_pins.py", "test_wall_cpu.py", "test_golden_cpu.py"]


def _env():
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    if not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("gcc's libasan.so not found")
    return dict(os.environ, QEM_ORACLE_LIB=ASAN_LIB, LD_PRELOAD=libasan,
                ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.mark.slow
def test_oracle_pins_under_sanitizers():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc/make absent")
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    env = _env()
    # the child really maps the sanitized build
    probe = ("import sys; sys.path.insert(0, %r); import oracle as O; O.legendre(5); "
             "print(any('liboracle_asan.so' in l for l in open('/proc/self/maps')))" % ORACLE)
    r = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "True", r.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider", *[os.path.join(ROOT, "tests", t) for t in PINS]],
                       env=env, capture_output=True, text=True, timeout=1200, cwd=ROOT)
    out = r.stdout[-3000:] + r.stderr[-3000:]
    assert r.returncode == 0, out
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, out
