"""Build recipe of the mag2d_amd native libraries.

* ``mag2d_amd/libmag2d_amd.so`` -- the HIP kernels and the C ABI
  (include/mag2d_amd.h), compiled for gfx950 with hipcc.
* ``mag2d_amd/host/libmag2d_host.so`` -- the C++ host layer
  (mag2d_amd/host/mag2d.hpp) exercised through a small C entry point.
* the oracle (``oracle/liboracle.so``) and, when the reference sources are
  present, ``oracle/_ref/`` -- test infrastructure, built by oracle/Makefile.

Objects are compiled in parallel into ``_build/`` and linked in place, so the
libraries travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "mag2d_amd"
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libmag2d_amd.so"
HOST_LIB = PKG / "host" / "libmag2d_host.so"
REFERENCE = Path("/root/reference")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", f"-I{ROOT / 'include'}",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a * b + c is fused where the source writes it in one expression, never
    # across statements: the fusing then does not depend on the code around an
    # inlined function, so kernels that share a device function (the single-run
    # and batched passes, forward and adjoint replay) compute the same bits
    "-ffp-contract=on",
    # no SLP vectorisation: it packs the particle pass's scalar compares into
    # vector masks that lower to shifts and selects, costing VALU issue slots
    # and registers (the pass ran 5% faster without it, results unchanged)
    "-fno-slp-vectorize",
]


def _sources_hip():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.hpp")) + sorted((ROOT / "include").glob("*.h"))


def _sources_host():
    return sorted((PKG / "host").glob("*.cpp")) + sorted((PKG / "host").glob("*.hpp")) + \
        sorted((ROOT / "include").glob("*.h"))


def _sources_oracle():
    return sorted((ROOT / "oracle").glob("*.c")) + sorted((ROOT / "oracle").glob("*.h")) + \
        [ROOT / "oracle" / "Makefile"]


def source_hash(files) -> str:
    """sha256 over the names and contents of a library's sources"""
    h = hashlib.sha256()
    for f in files:
        h.update(str(f.relative_to(ROOT)).encode())
        h.update(b"\0")
        h.update(f.read_bytes())
    return h.hexdigest()


def _stamp(lib: Path) -> Path:
    return lib.with_name(lib.name + ".srchash")


# library -> its sources; build() stamps each library with their hash
LIBRARIES = {
    "hip": (LIB, _sources_hip),
    "host": (HOST_LIB, _sources_host),
    "oracle": (ROOT / "oracle" / "liboracle.so", _sources_oracle),
}


def stale_libraries() -> list:
    """names of the libraries that are missing or older than their sources
    (by content hash, not file times)"""
    out = []
    for name, (lib, srcs) in LIBRARIES.items():
        st = _stamp(lib)
        if not lib.exists() or not st.exists() or st.read_text().strip() != source_hash(srcs()):
            out.append(name)
    return out


def _write_stamp(name: str) -> None:
    lib, srcs = LIBRARIES[name]
    _stamp(lib).write_text(source_hash(srcs()) + "\n")


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{out}")
    return out


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_hip(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    _write_stamp("hip")
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_host(verbose: bool = False) -> Path:
    src = PKG / "host" / "mag2d_host.cpp"
    if not src.exists():
        return HOST_LIB
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{ROOT / 'include'}",
          f"-I{PKG / 'host'}", str(src), "-o", str(tmp), f"-L{PKG}", "-lmag2d_amd",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, HOST_LIB)
    _write_stamp("host")
    if verbose:
        print(f"built {HOST_LIB}")
    return HOST_LIB


def build_oracle(verbose: bool = False) -> None:
    """Test infrastructure: the CPU oracle, and oracle/_ref when the
    reference sources exist in this container."""
    _run(["make", "-B", "-C", str(ROOT / "oracle"), "-s", "all"])
    _write_stamp("oracle")
    if REFERENCE.exists():
        _run(["make", "-C", str(ROOT / "oracle"), "-s", "ref", f"REF={REFERENCE}"])
    if verbose:
        print("built oracle")


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
