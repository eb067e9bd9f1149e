"""Builds the HIP libraries in-tree for gfx950.

  liblsmhip.so      LSMHIP_REAL=double: the FP64 drop-in (reference operation
                    order, -ffp-contract=off so that no FMA changes a rounding)
  liblsmhip_f32.so  LSMHIP_REAL=float: FP32 storage and arithmetic (contraction
                    allowed)

Both export the same symbols, as the reference's double and single precision
builds do (USE_SINGLE_PRECISION, CMakeLists.txt:36-41).  Objects go to
lsmlib_amd/_build/, the shared libraries next to this file.

    python -m lsmlib_amd.build [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

INCLUDE = os.path.join(os.path.dirname(HERE), "include")
COMMON = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-I" + INCLUDE,
          # packed FP32 VALU forms are kept out of every kernel
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
VARIANTS = {
    "liblsmhip.so": ["-DLSMHIP_REAL=double", "-ffp-contract=off"],
    "liblsmhip_f32.so": ["-DLSMHIP_REAL=float", "-ffp-contract=fast"],
}


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _compile(src, obj, flags):
    cmd = [HIPCC] + COMMON + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    return r.returncode, " ".join(cmd), msgs


def build(jobs: int = 8, verbose: bool = False) -> list[str]:
    os.makedirs(BUILD, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    newest_header = max([os.path.getmtime(h) for h in headers] + [0])
    tasks = []
    for lib, flags in VARIANTS.items():
        tag = os.path.splitext(lib)[0]
        for src in sources():
            obj = os.path.join(BUILD, tag + "__" + os.path.basename(src) + ".o")
            stale = (not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(src)
                     or os.path.getmtime(obj) < newest_header)
            tasks.append((lib, src, obj, flags, stale))
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, src, obj, flags): (lib, src) for lib, src, obj, flags, stale in tasks if stale}
        for f in cf.as_completed(futs):
            rc, cmd, msgs = f.result()
            if rc != 0:
                failed.append(futs[f][1])
                sys.stderr.write("FAILED: %s\n%s\n" % (cmd, msgs))
            elif verbose and msgs.strip():
                sys.stderr.write(msgs + "\n")
    if failed:
        raise RuntimeError("HIP compilation failed for: %s" % ", ".join(sorted(set(failed))))
    built = []
    for lib, flags in VARIANTS.items():
        objs = [t[2] for t in tasks if t[0] == lib]
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
        os.replace(tmp, out)
        built.append(out)
    return built


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    for p in build(a.j, a.v):
        print(p)


if __name__ == "__main__":
    main()
