"""Build recipe for libstreamingbc_hip.so (gfx950) and the CPU oracle.

Run as ``python streamingbc_amd/build.py`` or through ``__graft_entry__.build()`` (both
work before the library exists; importing the package needs it).
Always recompiles.  Outputs:

* ``streamingbc_amd/lib/libstreamingbc_hip.so`` -- kernels, host runtime and
  the host graph generators, the product library;
* ``oracle/liboracle.so`` -- the CPU oracle (test infrastructure);
* ``oracle/_ref/libsbcref.so`` -- the reference built from its own sources,
  only when ``/root/reference`` is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "streamingbc_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(ROOT, "_build")
LIBNAME = "libstreamingbc_hip.so"

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # FP64 expressions round exactly as written (no FMA contraction), so the
    # delta terms match a plain C evaluation bit for bit.
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_library() -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    o = os.path.join(BUILD, "sbc_graphgen.o")
    _run(["gcc", "-std=gnu99", "-O2", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"),
          "-c", os.path.join(CSRC, "sbc_graphgen.c"), "-o", o])
    objs.append(o)
    for src in ("sbc_kernels.hip", "sbc_engine.cpp", "sbc_mlp.hip"):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        o = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        _run([hipcc()] + HIP_FLAGS + ["-c", path, "-o", o])
        objs.append(o)
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, out)
    # the C driver (a reference-style caller of the C ABI)
    exe = os.path.join(LIBDIR, "sbc_stream")
    _run(["gcc", "-std=gnu99", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "sbc_stream_main.c"), "-o", exe + ".tmp",
          "-L" + LIBDIR, "-lstreamingbc_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    if os.path.isdir("/root/reference/src"):
        _run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")])


def build_all() -> None:
    build_library()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
