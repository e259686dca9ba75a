"""AddressSanitizer and UndefinedBehaviorSanitizer over the host-side C and
C++ code: the oracle (oracle/m2d_oracle.c) and the C++ host layer
(mag2d_amd/host/mag2d_host.cpp, whose CPU paths include the reference's
text and checkpoint formats and their error edges).  Both are rebuilt with
-fsanitize=address,undefined into a temporary directory, and the CPU tests
that drive them run again in a child process that loads those builds; any
sanitizer report ends the child with a non-zero status.  Device code is not
instrumented: GPU sanitizers are not available for it here."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _runtime(name):
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(_runtime("libasan.so") is None, reason="gcc's AddressSanitizer runtime is not installed")
def test_oracle_and_host_layer_under_asan_and_ubsan(tmp_path):
    oracle = tmp_path / "liboracle_san.so"
    host = tmp_path / "libmag2d_host_san.so"
    subprocess.run(["gcc", *SAN, "-fPIC", "-shared", "-std=c11", "-fopenmp", str(ROOT / "oracle" / "m2d_oracle.c"),
                    "-o", str(oracle), "-lm"], check=True)
    pkg = ROOT / "mag2d_amd"
    subprocess.run(["g++", *SAN, "-fPIC", "-shared", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{pkg / 'host'}",
                    str(pkg / "host" / "mag2d_host.cpp"), "-o", str(host), f"-L{pkg}", "-lmag2d_amd",
                    f"-Wl,-rpath,{pkg}"], check=True)
    env = dict(os.environ)
    # libstdc++ right after the ASan runtime, so that ASan finds the C++
    # runtime's __cxa_throw (the host layer throws across its C entry points)
    preload = [_runtime("libasan.so")] + ([_runtime("libstdc++.so.6")] if _runtime("libstdc++.so.6") else [])
    env.update(LD_PRELOAD=" ".join(preload), M2D_ORACLE_LIB=str(oracle), M2D_HOST_LIB=str(host),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    tests = sorted(str(p) for p in (ROOT / "tests").glob("test_oracle_*.py")) + [
        str(ROOT / "tests" / "test_host_files.py"), str(ROOT / "tests" / "test_host_shim.py"),
        str(ROOT / "tests" / "test_adjoint.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", *tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert " passed" in out
