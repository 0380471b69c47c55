"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/fc3_oracle.c is compiled with -fsanitize=address,undefined (no recovery) and every
C entry point runs once on small LJ and SW systems in a child process
(tests/sanitized_oracle_run.py) with the ASan runtime preloaded ahead of Python. A planted
out-of-bounds write checks that the set-up does catch errors.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SCRIPT = os.path.join(HERE, "sanitized_oracle_run.py")


def _runtime(name: str):
    gcc = shutil.which("gcc")
    if gcc is None:
        return None
    p = subprocess.run([gcc, f"-print-file-name={name}"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def sanitized_lib(tmp_path_factory):
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    out = str(tmp_path_factory.mktemp("asan") / "liboracle_san.so")
    cmd = ["gcc", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-shared", "-o", out,
           os.path.join(ROOT, "oracle", "fc3_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    return out, asan


def _run(lib_path: str, asan: str, *args: str):
    env = dict(os.environ)
    env.update(FC3_ORACLE_LIB=lib_path, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    return subprocess.run([sys.executable, SCRIPT, *args], capture_output=True, text=True,
                          env=env, timeout=600)


def test_oracle_entry_points_run_clean(sanitized_lib):
    p = _run(*sanitized_lib)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "ran clean" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr


def test_planted_overflow_is_reported(sanitized_lib):
    p = _run(*sanitized_lib, "--plant")
    assert p.returncode != 0
    assert "heap-buffer-overflow" in p.stderr
    assert "not detected" not in p.stdout
