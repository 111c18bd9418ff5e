"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(g++ instrumentation of api.cpp, scene_build.cpp, scene_blob.cpp, synth.cpp,
hosek.cpp; the device code is not instrumented): a host driver that builds
octrees, damages scene blobs and feeds corrupted node arrays to the
validation, and the host-side CPU tests run against the instrumented library.
No GPU is needed."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def sanitized():
    sys.path.insert(0, os.path.join(ROOT, "droplet-render_amd"))
    try:
        import build
    finally:
        sys.path.pop(0)
    if not os.path.exists(os.path.join(build.OBJDIR, "render.hip.o")):
        build.build()
    exe = build.build_sanitized()
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True).stdout.strip()
    return exe, build.SAN_LIB, asan


def _env(asan_lib=None, lib=None):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=%d:abort_on_error=1" % (0 if asan_lib else 1)
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    if asan_lib:
        env["LD_PRELOAD"] = asan_lib
    if lib:
        env["DRK_LIBRARY"] = lib
    return env


def test_host_driver_is_clean(sanitized, tmp_path):
    exe, _, _ = sanitized
    out = subprocess.run(["nm", exe], stdout=subprocess.PIPE, text=True).stdout
    assert "__asan_init" in out and "__ubsan_handle" in out     # really instrumented
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=600, env=_env())
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "host_sanitize ok" in r.stdout


def test_host_cpu_tests_under_sanitizers(sanitized):
    """The CPU tests of the octree build, scene blobs, SDF scenes and the C
    ABI's argument checks, with the instrumented library."""
    _, lib, asan = sanitized
    tests = [os.path.join(ROOT, "tests", t) for t in
             ("test_octree_build.py", "test_blob.py", "test_sdf.py", "test_abi.py", "test_chunks.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + tests,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=900, cwd=ROOT,
                       env=_env(asan, lib))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-4000:]
    assert "passed" in r.stdout
