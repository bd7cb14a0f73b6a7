"""Build libm1march.so for gfx950 with hipcc (run from any directory).

    python proton-dose-m1_amd/build_native.py

The kernel source is compiled twice (m1::exact with separately rounded
arithmetic, m1::fast with fused multiply-adds); host code is compiled with
-ffp-contract=off so that the per-step physics scalars round like the
reference's Python arithmetic.  Objects go to _build/, the library to
proton_solver/_lib/.
"""

from __future__ import annotations

import os
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "proton_solver" / "_lib"
LIB = LIBDIR / "libm1march.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("M1_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-I", str(ROOT / "include"), "-I", str(CSRC),
    # Packed FP32 instructions stay off (they are not used by this code and
    # some of their forms are unreliable on the target boxes).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def batch_shapes():
    """Tile shapes of the batched kernels, read from M1_BATCH_SHAPES in
    csrc/m1_device.h (the one list the library is built from)."""
    text = (CSRC / "m1_device.h").read_text()
    m = re.search(r"#define M1_BATCH_SHAPES\(X\)(.*)", text)
    shapes = [tuple(int(v) for v in t) for t in re.findall(r"X\((\d+), (\d+), (\d+)\)", m.group(1))]
    if not shapes:
        raise RuntimeError("no M1_BATCH_SHAPES in m1_device.h")
    return shapes


def _units():
    waves = [f"-DM1_STAGE_WAVES={int(os.environ['M1_STAGE_WAVES'])}"] \
        if os.environ.get("M1_STAGE_WAVES") else []
    # waves per SIMD of the FP64 face kernels (default in m1_kernels.hip)
    fwaves = [f"-DM1_F64_SPLIT_WAVES={int(os.environ['M1_F64_SPLIT_WAVES'])}"] \
        if os.environ.get("M1_F64_SPLIT_WAVES") else []
    return [
        (CSRC / "m1_kernels.hip", BUILD / "m1_kernels_exact.o",
         ["-DM1_EXACT=1", "-ffp-contract=off", *fwaves]),
        (CSRC / "m1_kernels.hip", BUILD / "m1_kernels_fast.o",
         ["-DM1_EXACT=0", "-ffp-contract=fast", *fwaves]),
        # float kernels: fused multiply-adds only where written (fmaf), so the
        # four moments of a cell round alike (m1_f32_common.h)
        (CSRC / "m1_kernels_f32.hip", BUILD / "m1_kernels_f32.o", ["-ffp-contract=off", *waves]),
        (CSRC / "m1_march.hip", BUILD / "m1_march.o", ["-ffp-contract=off", *waves]),
        (CSRC / "m1_box_kernels.hip", BUILD / "m1_box_kernels.o", ["-ffp-contract=off"]),
        (CSRC / "m1_gamma.hip", BUILD / "m1_gamma.o", ["-ffp-contract=off"]),
        *[(CSRC / "m1_batch_kernels.hip", BUILD / f"m1_batch_kernels_{x}x{y}x{z}.o",
           ["-ffp-contract=off", *waves, f"-DM1_TILE_TX={x}", f"-DM1_TILE_TY={y}",
            f"-DM1_BATCH_NS=batch_{x}x{y}x{z}", f"-DM1_BATCH_TABLE=batch_kernels_{x}x{y}x{z}"])
          for x, y, z in batch_shapes()],
        (CSRC / "m1_batch_march.hip", BUILD / "m1_batch_march.o", ["-ffp-contract=off", *waves]),
        (CSRC / "m1_mlp.hip", BUILD / "m1_mlp.o", ["-ffp-contract=off"]),
        (CSRC / "m1_batch_steps.hip", BUILD / "m1_batch_steps.o", ["-ffp-contract=off"]),
    ]


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    procs = []
    for src, obj, extra in _units():
        cmd = [HIPCC, *COMMON, *extra, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    phys_obj = BUILD / "m1_physics.o"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", str(ROOT / "include"),
           "-I", str(CSRC), "-c", str(CSRC / "m1_physics.cpp"), "-o", str(phys_obj)]
    procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        noise = "is not a recognized feature for this target"
        lines = [ln for ln in text.splitlines() if noise not in ln and "(ignoring feature)" not in ln]
        if p.returncode != 0 or (verbose and lines):
            print("\n".join(lines), file=sys.stderr)
        failed |= p.returncode != 0
    if failed:
        raise RuntimeError("libm1march.so: compilation failed")
    objs = [str(o) for _, o, _ in _units()] + [str(phys_obj)]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
