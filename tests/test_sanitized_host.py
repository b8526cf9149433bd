"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The descriptor parser and plan compiler (plan_compile.cpp), the kernel source
generator (spec_codegen.cpp) and the C ABI take caller-supplied sizes and
arrays; this builds the library once more with its host code sanitized
(tensorqec.jl_amd/build.py --sanitize, device code unchanged) and runs the
C-ABI CPU tests (tests/test_lib_cpu.py: valid and malformed descriptors,
null and out-of-range arguments, statistics, generated kernel sources)
against it in a child Python with the ASan runtime preloaded.  Any
sanitizer report aborts the child, failing this test.  No GPU is used.
"""
import importlib.util
import os
import shutil
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "tensorqec.jl_amd")


def _build_module():
    spec = importlib.util.spec_from_file_location("tqec_build_san", os.path.join(PKG, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_c_abi_cpu_tests_pass_under_asan_and_ubsan():
    if not (os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    b = _build_module()
    try:
        rt = b.sanitizer_runtime()
    except RuntimeError as e:
        pytest.skip(str(e))
    lib = b.build(sanitize=True)
    env = dict(os.environ)
    env.update({"LD_PRELOAD": rt, "TQEC_LIB": lib,
                "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    probe = ("import sys; sys.path.insert(0, %r); from tensorqec import lib as L; L.load(); "
             "print(L.LIB_PATH)" % PKG)
    r = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == lib, r.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(REPO, "tests", "test_lib_cpu.py"), "-q",
                        "-p", "no:cacheprovider", "-m", "not gpu"],
                       env=env, cwd=REPO, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out
