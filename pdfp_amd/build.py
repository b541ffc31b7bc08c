"""Builds libpdfp_hip.so in-tree with hipcc for gfx950.

    python -m pdfp_amd.build            # build the library
    python -m pdfp_amd.build --check    # print the exported C-ABI symbols

The library is written next to this file (pdfp_amd/libpdfp_hip.so), and the
benchmark's gather probe beside it as a library of its own
(pdfp_amd/libpdfp_probe.so, include/pdfp_probe.h).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpdfp_hip.so")
SOURCES = ["pdfp_hip.hip", "pdlp_hip.hip", "batch_hip.hip", "pack_hip.hip", "scale_hip.hip",
           "check_hip.hip", "rounder.cpp"]
PROBE = os.path.join(HERE, "libpdfp_probe.so")
PROBE_SOURCES = ["bench_hip.hip"]

# -ffp-contract=off: no a*b+c contraction, so the FP64 form performs the
# reference's operations one by one.  The packed-FP32 target feature is
# switched off for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _link(objs, out, verbose):
    tmp = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed with exit code {res.returncode}")
    os.replace(tmp, out)


def build(verbose: bool = False, out: str = LIB, defines: tuple = (),
          probe: str | None = PROBE) -> str:
    """Compiles every source to an object in parallel (one hipcc per file,
    under pdfp_amd/_build/), then links the library and (probe) the probe
    library."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    psrcs = [os.path.join(CSRC, s) for s in PROBE_SOURCES] if probe else []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    common = [*flags, *[f"-D{d}" for d in defines], "-I", os.path.join(ROOT, "include")]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + (".%d.o" % os.getpid()))
        cmd = [hipcc(), *common, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        return obj, subprocess.run(cmd, capture_output=True, text=True)

    allsrc = srcs + psrcs
    with ThreadPoolExecutor(max_workers=min(len(allsrc), os.cpu_count() or 1, 8)) as ex:
        results = list(ex.map(compile_one, allsrc))
    failed = [(o, r) for o, r in results if r.returncode != 0]
    objs = [o for o, _ in results]
    try:
        if failed:
            for _, r in failed:
                sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {len(failed)} source(s)")
        _link(objs[:len(srcs)], out, verbose)
        if probe:
            _link(objs[len(srcs):], probe, verbose)
    finally:
        for o in objs:
            if os.path.exists(o):
                os.remove(o)
    return out


def exported_symbols(path: str = LIB) -> list[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return sorted(line.split()[-1] for line in out.splitlines() if " T " in line)


if __name__ == "__main__":
    if "--check" in sys.argv:
        print("\n".join(s for s in exported_symbols() if s.startswith(("pdfp_", "pdlp_"))))
    else:
        args = sys.argv[1:]
        out = args[args.index("--out") + 1] if "--out" in args else LIB
        defs = tuple(args[i + 1] for i, a in enumerate(args) if a == "-D")
        print(build(verbose=True, out=out, defines=defs))
