"""Build recipe of liboal.so (gfx950) and, for the test harness, the CPU oracle.

`python -m openaerolab_alpha_01_amd.build` compiles in-tree; the built library sits next to this
file so that it travels with the source tree.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "liboal.so"
# the step kernels of each storage form sit in their own translation unit and compile in parallel
SOURCES = [CSRC / "oal_step_f32.hip", CSRC / "oal_step_f16s.hip", CSRC / "oal_step_bf16s.hip",
           CSRC / "oal_step_f64.hip", CSRC / "oal_kernels.hip", CSRC / "oal_surface.hip", CSRC / "oal_plan.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every multiply and add rounds separately, as in the reference arithmetic.
# -packed-fp32-ops off: no v_pk_*_f32 instructions are generated.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
]


def build_library(verbose: bool = False, extra: list[str] | None = None, jobs: int | None = None,
                  out: Path | None = None) -> Path:
    """Compile liboal.so for gfx950 (always recompiles): one object per source, in parallel, then link."""
    lib = Path(out).resolve() if out else LIB
    jobs = jobs or min(len(SOURCES), max(1, len(os.sched_getaffinity(0))))
    with tempfile.TemporaryDirectory(prefix="oal_build_") as tmpdir:
        objs = [Path(tmpdir) / (src.stem + ".o") for src in SOURCES]
        cmds = [[HIPCC, *HIP_FLAGS, *(extra or []), "-c", "-o", str(o), str(src)] for src, o in zip(SOURCES, objs)]

        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, cwd=str(CSRC), capture_output=True, text=True)

        with ThreadPoolExecutor(jobs) as pool:
            results = list(pool.map(run, cmds))
        for r in results:
            if r.stdout or r.stderr:
                sys.stderr.write(r.stdout + r.stderr)
        bad = [c for c, r in zip(cmds, results) if r.returncode != 0]
        if bad:
            raise subprocess.CalledProcessError(1, bad[0])
        tmp = lib.with_suffix(".so.tmp")
        link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
        if verbose:
            print(" ".join(link), flush=True)
        subprocess.run(link, check=True, cwd=str(CSRC))
    os.replace(tmp, lib)
    return lib


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    oracle = ROOT / "oracle"
    cmd = ["make", "-C", str(oracle), "-B", "liboal_oracle.so"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=None if verbose else subprocess.DEVNULL)
    return oracle / "liboal_oracle.so"


def scratch_kernels(report: str) -> list[tuple[str, int]]:
    """(kernel, bytes per lane) of every kernel that the compiler's resource-usage remarks show with
    scratch (register spills)"""
    out = []
    for block in report.split("remark: Function Name: ")[1:]:
        name = block.split()[0]
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", block)
        if m and int(m.group(1)) > 0:
            out.append((name, int(m.group(1))))
    return out


def main(argv: list[str]) -> int:
    verbose = "-v" in argv
    if "--resources" in argv:
        # every kernel's registers and scratch; fails when any kernel spills to scratch
        import io
        import contextlib
        err = io.StringIO()
        with contextlib.redirect_stderr(err):
            build_library(verbose=verbose, extra=["-Rpass-analysis=kernel-resource-usage"])
        report = err.getvalue()
        sys.stderr.write(report)
        spills = scratch_kernels(report)
        for name, b in spills:
            print(f"scratch: {b} B/lane in {name}")
        print(f"{report.count('Function Name: ')} kernels, {len(spills)} with scratch")
        if spills:
            return 1
    else:
        build_library(verbose=verbose)
    if "--no-oracle" not in argv:
        build_oracle(verbose=verbose)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
