"""Build step of the native libraries (run on a machine with ROCm; no GPU needed).

Produces, in-tree:
  openphase_amd/libmpf_hip.so   -- the MI355X step (HIP kernels for gfx950 + the C ABI of include/mpf.h)
and, for tests only, oracle/libmpf_oracle.so (the CPU oracle) and, when the reference sources are
present, oracle/_ref/ref_driver (the golden-vector driver built from them).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "openphase_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libmpf_hip.so"

# -ffp-contract=off: no FMA contraction, so the FP64 path performs exactly the reference's operations.
# -packed-fp32-ops disabled: packed FP32 instructions are not used by any kernel here.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip() -> Path:
    """Compiles each source to an object in parallel, then links the shared library."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    # (source, object name, extra defines): the per-cell step kernels build as four variants
    units = [(CSRC / "mpf_kernels.hip", "mpf_kernels", []), (CSRC / "mpf_tile.hip", "mpf_tile", []),
             (CSRC / "mpf_adjoint.hip", "mpf_adjoint", []), (CSRC / "mpf_dgavg.hip", "mpf_dgavg", []),
             (CSRC / "mpf_energy.hip", "mpf_energy", []),
             (CSRC / "mpf_capi.cpp", "mpf_capi", [])]
    for fp32 in (0, 1):
        for full in (0, 1):
            units.append((CSRC / "mpf_kernels_variant.hip", f"mpf_kernels_v{32 if fp32 else 64}_{'full' if full else 'plain'}",
                          [f"-DMPF_VARIANT_FP32={fp32}", f"-DMPF_VARIANT_FULL={full}"]))
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src, name, defs in units:
        obj = objdir / (name + ".o")
        objs.append(str(obj))
        cmd = [hipcc, *compile_flags, *defs, "-x", "hip", "-c", str(src), "-o", str(obj)]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    if any(p.wait() != 0 for p in procs):
        raise RuntimeError("hipcc failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_examples() -> None:
    """The C++ drop-in examples (OpenPhase-style classes over the C ABI), linked to libmpf_hip.so."""
    cxx = shutil.which("g++") or "g++"
    for name in ("normal_gg", "driving_force", "regularization"):
        src = PKG / "examples" / f"{name}.cpp"
        out = PKG / "examples" / name
        _run([cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(src), f"-L{PKG}", "-lmpf_hip",
              "-Wl,-rpath,$ORIGIN/..", "-o", str(out)])


def build_oracle() -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    ref = Path(os.environ.get("OPENPHASE_REFERENCE", "/root/reference"))
    if (ref / "src" / "PhaseField.cpp").exists():
        jobs = str(min(8, os.cpu_count() or 1))
        _run(["make", "-s", "-f", str(ROOT / "oracle" / "Makefile.ref"), f"REF={ref}", f"-j{jobs}"])


def build_all() -> None:
    build_hip()
    build_examples()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"hip": build_hip, "examples": build_examples, "oracle": build_oracle, "all": build_all}[what]()
