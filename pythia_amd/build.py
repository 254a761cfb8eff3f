"""Build recipes for the native parts of the package.

* ``libpgpu.so``: every HIP source under ``csrc/`` compiled for gfx950 into
  one shared library exposing the C ABI of ``include/pgpu.h``.
* ``libpythia_host.so``: the C++ operator host (``host/``), which reaches the
  GPU only through ``libpgpu.so``.

Both are written in-tree so they travel with the repository snapshot.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBPGPU = os.path.join(PKG, "libpgpu.so")
LIBHOST = os.path.join(PKG, "libpythia_host.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def hip_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def host_sources():
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))


def build_pgpu(jobs: int = 8) -> str:
    """Compile each .hip to an object in parallel, then link libpgpu.so."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in hip_sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        procs.append((src, subprocess.Popen([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj],
                                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(f"{src}:\n{out}")
    if errors:
        raise RuntimeError("HIP compile failed:\n" + "\n".join(errors))
    tmp = LIBPGPU + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIBPGPU)
    return LIBPGPU


def build_host() -> str | None:
    srcs = host_sources()
    if not srcs:
        return None
    tmp = LIBHOST + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", f"-I{INCLUDE}", f"-I{HOST}",
          *srcs, "-o", tmp, f"-L{PKG}", "-lpgpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, LIBHOST)
    return LIBHOST


def build_all() -> None:
    build_pgpu()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", LIBPGPU)
