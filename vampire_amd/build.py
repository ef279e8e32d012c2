"""Builds libvmx.so (hipcc, gfx950) and the C++ gpu:: shim object check.

    python -m vampire_amd.build            # compile everything, in-tree

Flags common to every HIP unit: -O3, C++17, gfx950 only, packed FP32 instructions disabled.
The FP64 unit and the context unit use -ffp-contract=off so that device arithmetic rounds
where the reference's C++ rounds; the FP32 unit allows contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvmx.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("VMX_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-I{INCLUDE}"]
UNITS = [
    ("vmx_f64.hip", ["-ffp-contract=off"]),
    ("vmx_f32.hip", ["-ffp-contract=fast"]),
    ("vmx_api.hip", ["-ffp-contract=off"]),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)


def build(verbose: bool = False) -> str:
    """Compiles every HIP unit and links libvmx.so; returns its path."""
    os.makedirs(OUT, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        objs.append(obj)
        procs.append((src, _run([cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj], verbose)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature flag; that warning is expected
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append((src, "\n".join(lines)))
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(f"[{s}]\n{t}" for s, t in failed))
    tmp = LIB + ".tmp"
    link = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    p = _run(link, verbose)
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + out.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
