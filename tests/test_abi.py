"""CPU checks of the C ABI: libscarl_hip.so loads, exports exactly what include/scarl_hip.h
declares, and fails cleanly (status + message, no crash) where no GPU is present."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "scarl_hip.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^int (scarl_hip_\w+)\(", txt, flags=re.M)))


def test_header_and_binding_agree():
    from scarl_amd import _capi as A
    assert declared() == sorted(A.SIGNATURES)


def test_library_exports_every_declared_symbol():
    from scarl_amd import _capi as A
    L = C.CDLL(A.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (scarl_hip_\w+)", out))
    assert exported == set(declared())


def _run_cpu(code, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    env["HIP_VISIBLE_DEVICES"] = env.get("HIP_VISIBLE_DEVICES", "")
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, env=env,
                          timeout=120)


def test_errors_are_status_codes_not_crashes():
    code = r'''
import numpy as np, ctypes as C
from scarl_amd import _capi as A
rc = A.lib().scarl_hip_init(0)
print("init", rc, A.last_error())
out = np.zeros((4, 4, 4), order="F")
Np = np.array([4, 4, 4], np.int32); xr = np.ones(3); cl = np.ones(3)
rc2 = A.lib().scarl_hip_gen_std_gauss_fft(A.ptr(out), A.ptr(Np), A.ptr(xr), A.ptr(cl), 1, 1, 0, None, 0)
print("gen", rc2, A.last_error())
'''
    r = _run_cpu(code, {"SCARL_FAULT_INJECT": "init"})
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines())
    assert lines["init"].startswith("2 ") and "fault injection" in lines["init"]
    assert lines["gen"].startswith("2 ")


def test_python_mirror_raises_with_message():
    code = r'''
from scarl_amd import normal_ScaRL, ScarlError
try:
    normal_ScaRL.r8vec_uniform_01(4, 1)
except ScarlError as e:
    print("raised", e.code, str(e))
'''
    r = _run_cpu(code, {"SCARL_FAULT_INJECT": "init"})
    assert "raised 2" in r.stdout, r.stdout + r.stderr


def test_missing_library_fails_loudly(tmp_path):
    code = r'''
import scarl_amd._capi as A
A.LIB_PATH = "/nonexistent/libscarl_hip.so"
try:
    A.lib()
except A.ScarlLibraryMissing as e:
    print("missing", e)
'''
    r = _run_cpu(code)
    assert "missing" in r.stdout
