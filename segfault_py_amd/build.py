"""Build libsfeng.so (the HIP library) in-tree for gfx950 with hipcc.

    python -m segfault_py_amd.build        # or: from segfault_py_amd.build import build

Each .hip translation unit is compiled in parallel to an object under
segfault_py_amd/_build/, then linked into segfault_py_amd/libsfeng.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsfeng.so")
# (source, extra defines, object name): k_eri.hip is compiled once per
# quartet class so the 21 classes build in parallel.
# The d-shell classes (6..20) are further split into their J/K, gradient
# and McMurchie-Davidson parts (SF_PART), the heaviest compiles.
# f-shell classes with compiled kernels (the SF_FOR_EACH_FCLASS list of
# k_eri.hip): class id and (la, lb, lc, ld); J/K only, one unit each.
FCLASSES = [(21, 3, 0, 0, 0), (22, 3, 0, 1, 0), (23, 3, 0, 1, 1), (24, 3, 0, 2, 0),
            (25, 3, 0, 2, 1), (26, 3, 0, 2, 2), (27, 3, 0, 3, 0), (28, 3, 1, 0, 0),
            (29, 3, 1, 1, 0), (30, 3, 1, 1, 1), (31, 3, 1, 2, 0), (32, 3, 1, 2, 1),
            (34, 3, 1, 3, 0), (36, 3, 2, 0, 0), (37, 3, 2, 1, 0), (38, 3, 2, 1, 1),
            (39, 3, 2, 2, 0), (45, 3, 3, 0, 0), (46, 3, 3, 1, 0)]
UNITS = ([("k_eri.hip", [], "k_eri_dispatch.o")]
         + [("k_eri.hip", [f"-DSF_CLASS={c}"], f"k_eri_c{c:02d}.o") for c in range(6)]
         + [("k_eri.hip", [f"-DSF_CLASS={c}", f"-DSF_PART={p}"], f"k_eri_c{c:02d}p{p}.o")
            for c in range(6, 21) for p in range(3)]
         + [("k_eri.hip", ["-DSF_CLASS=99", f"-DSF_FCLASS={c}", f"-DSF_FLA={a}", f"-DSF_FLB={b}",
                           f"-DSF_FLC={cc}", f"-DSF_FLD={d}"], f"k_eri_f{c:02d}.o")
            for c, a, b, cc, d in FCLASSES]
         + [("k_eri_gen.hip", [], "k_eri_gen.o"),
            ("k_util.hip", [], "k_util.o"), ("k_onee.hip", [], "k_onee.o"),
            ("sf_api.hip", [], "sf_api.o")])
ARCH = os.environ.get("SF_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled on purpose: keep them out of all code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _newest_header() -> float:
    """Latest modification time of the headers every unit includes."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    paths = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".inc", ".h"))]
    paths += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return max(os.path.getmtime(p) for p in paths)


def build(verbose: bool = False, extra_flags: list[str] | None = None,
          incremental: bool = False) -> str:
    """Compile every unit and link libsfeng.so.  incremental=True skips a
    unit whose object is newer than its source and every header."""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    cc = hipcc()
    flags = HIPCC_FLAGS + list(extra_flags or [])
    objs = []
    cmds = []
    hdr = _newest_header() if incremental else 0.0
    for src, defs, oname in UNITS:
        obj = os.path.join(bdir, oname)
        objs.append(obj)
        if incremental and os.path.exists(obj):
            t = os.path.getmtime(obj)
            if t > hdr and t > os.path.getmtime(os.path.join(CSRC, src)):
                continue
        cmds.append((oname, [cc, *flags, *defs, "-c", os.path.join(CSRC, src), "-o", obj]))
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    failed = []
    running = []
    pending = list(cmds)
    # longest units first: the d-shell classes (their gradient and J/K parts)
    # dominate the build time
    def cost(c):
        name = c[0]
        if name.startswith("k_eri_f"):
            # integrals per quartet of the f class (compile time grows with it)
            cid = int(name[7:9])
            ncart = lambda L: (L + 1) * (L + 2) // 2  # noqa: E731
            ls = next(f[1:] for f in FCLASSES if f[0] == cid)
            n = ncart(ls[0]) * ncart(ls[1]) * ncart(ls[2]) * ncart(ls[3])
            return n / 20.0
        if not name.startswith("k_eri_c"):
            return 0
        part = name[10:11] if name[9:10] == "p" else "0"
        return int(name[7:9]) * (0.1 if part == "2" else 1.0)
    pending.sort(key=lambda c: -cost(c))
    while pending or running:
        while pending and len(running) < jobs:
            name, cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd), flush=True)
            log = tempfile.TemporaryFile()
            running.append((name, subprocess.Popen(cmd, stdout=log, stderr=subprocess.STDOUT), log))
        # refill as soon as any compile ends (not the oldest one)
        k = next((k for k, r in enumerate(running) if r[1].poll() is not None), None)
        if k is None:
            time.sleep(0.1)
            continue
        name, p, log = running.pop(k)
        log.seek(0)
        out = log.read()
        log.close()
        if p.returncode != 0:
            failed.append((name, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"--- {s} ---\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = OUT + ".tmp"
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
