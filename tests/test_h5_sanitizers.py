"""The HDF5/XDMF writer (scarl_amd/csrc/scarl_h5.cpp, host C++) under AddressSanitizer and
UndefinedBehaviorSanitizer with leak checking: tests/h5_abi_exercise.cpp drives every entry point
of include/scarl_h5.h, including short buffers, small read caps and the error paths.

The exerciser is a standalone program: a sanitizer runtime cannot start late inside a Python
process that has already mapped memory where its shadow must go."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DIR = os.environ.get("H5DIR", "/opt/conda")


def _have_toolchain():
    return shutil.which("g++") is not None and os.path.exists(os.path.join(H5DIR, "include", "hdf5.h"))


@pytest.mark.skipif(not _have_toolchain(), reason="g++ or the HDF5 headers are not available")
def test_h5_abi_under_address_and_ub_sanitizers(tmp_path):
    exe = tmp_path / "h5_abi_exercise"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-static-libstdc++", "-isystem", os.path.join(H5DIR, "include"),
           "-o", str(exe), os.path.join(ROOT, "tests", "h5_abi_exercise.cpp"),
           os.path.join(ROOT, "scarl_amd", "csrc", "scarl_h5.cpp"), os.path.join(H5DIR, "lib", "libhdf5.so"),
           "-Wl,-rpath," + os.path.join(H5DIR, "lib")]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    report = out.stdout + out.stderr
    assert "AddressSanitizer" not in report and "LeakSanitizer" not in report, report[-4000:]
    assert "runtime error" not in report, report[-4000:]
    assert out.returncode == 0 and "OK (0 failures)" in out.stdout, report[-4000:]
