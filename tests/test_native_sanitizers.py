"""Sanitizer builds of the host code, run on the CPU.

- The server's host core (mesoengine_amd/csrc/serve_core.h: admission,
  ring entries, completion, take) under ThreadSanitizer and under
  AddressSanitizer + UndefinedBehaviorSanitizer, driven by client threads and
  a fake device thread (tests/native/serve_core_stress.cpp).
- The CPU oracle (oracle/meso_oracle.c) under ASan + UBSan, and its threaded
  raymarch and mesher under TSan (tests/native/oracle_sanitize.c).
- The engine adapter's host code (mesoengine_amd/engine/meso_instance.cpp)
  under ASan + UBSan (tests/native/engine_sanitize.cpp).

The reference's own worker pools and queues race (Runtimes/Thread/ThreadPool.h,
Runtimes/Thread/DoubleBufferQueue.h); these builds are how this code is kept
free of that.  ROCm's clang is used: its TSan runtime intercepts the
pthread_cond_clockwait that libstdc++'s timed waits call, which the GCC 11
runtime in this image does not (that gap makes every timed wait look like a
double lock).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
CLANG = "/opt/rocm/llvm/bin/clang"
CLANGXX = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not (os.path.exists(CLANGXX) and os.path.exists(CLANG)),
                                reason="ROCm clang not installed")

ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
TSAN = ["-fsanitize=thread"]


def _build_and_run(tmp_path, compiler, flags, sources, extra, name, env=None, timeout=300):
    exe = str(tmp_path / name)
    cmd = [compiler, "-O1", "-g"] + flags + sources + extra + ["-o", exe]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, " ".join(cmd) + "\n" + p.stderr[-4000:]
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=e)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "ThreadSanitizer" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    assert "runtime error" not in r.stderr, r.stderr[-6000:]
    return r.stdout


@pytest.mark.parametrize("san", ["tsan", "asan"])
def test_serve_core_under_sanitizers(tmp_path, san):
    out = _build_and_run(tmp_path, CLANGXX, ["-std=c++17", "-pthread"] + (TSAN if san == "tsan" else ASAN),
                         [os.path.join(NATIVE, "serve_core_stress.cpp")],
                         ["-I" + os.path.join(ROOT, "include"),
                          "-I" + os.path.join(ROOT, "mesoengine_amd", "csrc")],
                         "serve_core_" + san, env={"TSAN_OPTIONS": "halt_on_error=1"})
    assert "serve_core_stress ok" in out


@pytest.mark.parametrize("san", ["asan", "tsan"])
def test_oracle_under_sanitizers(tmp_path, san):
    out = _build_and_run(tmp_path, CLANG, ["-std=c11", "-ffp-contract=off", "-pthread"] +
                         (TSAN if san == "tsan" else ASAN),
                         [os.path.join(NATIVE, "oracle_sanitize.c"),
                          os.path.join(ROOT, "oracle", "meso_oracle.c")],
                         ["-I" + os.path.join(ROOT, "oracle"), "-lm"], "oracle_" + san,
                         env={"TSAN_OPTIONS": "halt_on_error=1"})
    assert "oracle_sanitize ok" in out


def test_engine_host_code_under_asan(tmp_path):
    build = os.path.join(ROOT, "mesoengine_amd", "_build")
    if not os.path.exists(os.path.join(build, "libmeso_hip.so")):
        pytest.skip("libmeso_hip.so not built")
    out = _build_and_run(tmp_path, CLANGXX, ["-std=c++17"] + ASAN,
                         [os.path.join(NATIVE, "engine_sanitize.cpp"),
                          os.path.join(ROOT, "mesoengine_amd", "engine", "meso_instance.cpp")],
                         ["-I" + os.path.join(ROOT, "include"),
                          "-I" + os.path.join(ROOT, "mesoengine_amd", "engine"),
                          "-L" + build, "-lmeso_hip", "-Wl,-rpath," + build],
                         "engine_asan", env={"HIP_VISIBLE_DEVICES": ""})
    assert "engine_sanitize ok" in out
