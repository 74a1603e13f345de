"""AddressSanitizer + UndefinedBehaviorSanitizer build of the pure host code (complex helpers,
small eigen-solver, mode solver and mode files: 3d_fdtd_amd/csrc/host/{complex,matrix,mode,
mode_solve}.c) and of the FP64 oracle (oracle/fdtd_oracle.c, serial), driven by
tests/sanitize_driver.c on the CPU.  Any sanitizer report makes the run fail."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "3d_fdtd_amd", "csrc", "host")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_host_code_and_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_driver")
    srcs = [os.path.join(ROOT, "tests", "sanitize_driver.c"), os.path.join(ROOT, "oracle", "fdtd_oracle.c")]
    srcs += [os.path.join(HOST, f) for f in ("complex.c", "matrix.c", "mode.c", "mode_solve.c")]
    cmd = ["gcc", "-O1", "-g", "-std=gnu99", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"),
           "-I" + HOST, *srcs, "-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "eigen 1 mode 1 oracle 1" in r.stdout
