"""Links the library's objects against the stand-in HIP runtime (fake_hip.cpp)
into tests/fakehip/_build/libmtcheck_fakehip.so, for CPU tests of the host code.

The objects are the ones mtcheck_amd/build.py compiles (mtcheck_amd/_build); they
are built first when missing.  Test infrastructure only.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libmtcheck_fakehip.so")


def _objects():
    from mtcheck_amd import build as B
    objs = [os.path.join(B.OBJ, os.path.splitext(s)[0] + ".o") for s in B.SOURCES]
    srcs = [os.path.join(B.CSRC, f) for f in os.listdir(B.CSRC)]
    newest_src = max(os.path.getmtime(p) for p in srcs)
    if not all(os.path.exists(o) for o in objs) or min(os.path.getmtime(o) for o in objs) < newest_src:
        B.build()
    return objs, B


def build() -> str:
    objs, B = _objects()
    os.makedirs(OUT, exist_ok=True)
    fake_o = os.path.join(OUT, "fake_hip.o")
    src = os.path.join(HERE, "fake_hip.cpp")
    deps = [src] + [os.path.join(B.CSRC, f) for f in ("args.hpp", "internal.hpp", "device.hpp")]
    if not os.path.exists(fake_o) or os.path.getmtime(fake_o) < max(os.path.getmtime(p) for p in deps):
        cmd = [B.HIPCC, "-x", "hip", f"--offload-arch={B.ARCH}", "-O1", "-std=c++17", "-fPIC", "-w",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I", os.path.join(ROOT, "include"), "-c", src, "-o", fake_o + f".{os.getpid()}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"fake_hip.cpp failed to compile:\n{r.stdout}\n{r.stderr}")
        os.replace(fake_o + f".{os.getpid()}", fake_o)  # test processes may build at once
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(p) for p in objs + [fake_o]):
        tmp = LIB + f".{os.getpid()}"
        cmd = ["g++", "-shared", "-Wl,-Bsymbolic", "-o", tmp, fake_o, *objs, "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link of the fake-HIP library failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
