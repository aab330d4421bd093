"""Host code under the compiler's sanitizers (CPU only, g++/gcc):

* pattern.cpp (CRS pattern, colouring, incidence, Gauss tables; its
  multi-threaded row gathering) with tests/sanitize/pattern_check.cpp under
  AddressSanitizer + UndefinedBehaviorSanitizer, and under ThreadSanitizer
  on a mesh large enough for the worker threads;
* oracle.c built with AddressSanitizer + UndefinedBehaviorSanitizer and the
  oracle test modules re-run against that build in a child process.

The reference's counterpart is its gDEBUG bounds checks
(FEM/matrix_class.cpp:1180-1187).  A planted out-of-bounds write shows that
the sanitizer build really reports."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ogs5_amd", "csrc")
ROCM_INC = "/opt/rocm/include"

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None
                                or not os.path.isdir(ROCM_INC),
                                reason="needs gcc/g++ and the HIP headers")

ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _run(cmd, env=None, timeout=600):
    p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def _build_pattern_check(tmp_path, flags):
    exe = str(tmp_path / "pattern_check")
    rc, out = _run(["g++", "-std=c++17", "-O1", "-g", *flags, "-D__HIP_PLATFORM_AMD__",
                    "-I" + ROCM_INC, "-I" + CSRC,
                    os.path.join(ROOT, "tests", "sanitize", "pattern_check.cpp"),
                    os.path.join(CSRC, "pattern.cpp"), "-o", exe, "-lpthread"])
    assert rc == 0, out
    return exe


def test_pattern_under_address_and_ub_sanitizers(tmp_path):
    exe = _build_pattern_check(tmp_path, ASAN)
    for args in ([], ["big"]):
        rc, out = _run([exe, *args], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
        print(out)
        assert rc == 0, out


def test_pattern_under_thread_sanitizer(tmp_path):
    exe = _build_pattern_check(tmp_path, ["-fsanitize=thread"])
    rc, out = _run([exe, "big"], env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    print(out)
    assert rc == 0, out
    assert "ThreadSanitizer" not in out


def _asan_runtime():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("gcc's AddressSanitizer runtime not found")
    return path


def test_planted_overflow_is_reported(tmp_path):
    src = tmp_path / "planted.c"
    src.write_text("#include <stdlib.h>\nint main(void){int*a=malloc(4*sizeof(int));"
                   "a[4]=1;int r=a[0];free(a);return r;}\n")
    exe = str(tmp_path / "planted")
    rc, out = _run(["gcc", "-O0", "-g", *ASAN, str(src), "-o", exe])
    assert rc == 0, out
    rc, out = _run([exe])
    assert rc != 0 and "heap-buffer-overflow" in out


def test_oracle_tests_under_address_and_ub_sanitizers(tmp_path):
    lib = str(tmp_path / "liboracle_san.so")
    rc, out = _run(["gcc", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", *ASAN,
                    os.path.join(ROOT, "oracle", "oracle.c"), "-o", lib, "-lm"])
    assert rc == 0, out
    env = dict(os.environ, OGSX_ORACLE_LIB=lib, LD_PRELOAD=_asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    rc, out = _run([sys.executable, "-c",
                    "from oracle import ffi; ffi.lib(); print(ffi.LIB_PATH)"], env=env)
    assert rc == 0 and out.strip().endswith("liboracle_san.so"), out
    mods = [os.path.join(ROOT, "tests", m) for m in
            ("test_oracle_math.py", "test_oracle_ref.py", "test_oracle_velocity.py",
             "test_deck.py")]
    rc, out = _run([sys.executable, "-m", "pytest", "-p", "no:cacheprovider", "-q", "-x",
                    "-m", "not gpu", *mods], env=dict(env, PYTHONPATH=ROOT))
    print(out[-3000:])
    assert rc == 0, out[-3000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out
