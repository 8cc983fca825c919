"""Build libplato_hip.so (HIP, gfx950) in-tree with hipcc.

    python -m plato_amd.build            # compile every csrc/*.hip into one shared library
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libplato_hip.so")
ARCH = os.environ.get("PLATO_HIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build(verbose: bool = True, jobs: int = 8, out: str | None = None,
          defines: list[str] | None = None) -> str:
    """Compile each translation unit to an object in parallel, then link."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = out or LIB
    defines = [f"-D{d}" for d in (defines or [])]
    objdir = os.path.join(os.path.dirname(out), "_build" + ("_" + "_".join(d[2:] for d in defines)
                                                          if defines else ""))
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    t0 = time.time()
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc, *[f for f in HIPCC_FLAGS if f != "-shared"], *defines, "-c", *inc, src, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            time.sleep(0.05)
    failed = False
    for src, p in procs:
        stdout, _ = p.communicate()
        text = stdout.decode(errors="replace")
        # the host pass does not know the device-only target feature; harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {os.path.basename(src)} ---\n" + "\n".join(lines) + "\n")
        elif verbose and any("warning" in l for l in lines):
            sys.stderr.write(f"--- {os.path.basename(src)} (warnings) ---\n" + "\n".join(lines[:40]) + "\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = out + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(link, check=True)
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} in {time.time() - t0:.1f}s")
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    build(out=a.out, defines=a.defines)
