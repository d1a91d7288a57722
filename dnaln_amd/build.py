"""Builds the native parts in-tree.

    python -m dnaln_amd.build            # everything
    python -m dnaln_amd.build hip host   # selected parts

Outputs (git-ignored, but they travel with the tree):
  dnaln_amd/lib/libdnaln_hip.so   HIP kernels + the C-ABI of include/dnaln_hip.h (gfx950)
  dnaln_amd/lib/libdnaln_host.so  host C/C++: index build, windows, SAM, generator, C++ mirror
  dnaln_amd/bin/dnaln-amd         `map` CLI with the reference's flags, on the HIP path
  dnaln_amd/lib/libdnaln_emu.so   CPU build of the per-lane device programs (test infrastructure)
  dnaln_amd/bin/graph_emu         the library over a CPU emulation of HIP, with ASan (test infrastructure)
  oracle/liboracle.so             C oracle (test infrastructure)
  oracle/_ref/                    the reference built from its sources (only where it is mounted)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "dnaln_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
INCLUDE = ROOT / "include"
REFERENCE = Path("/root/reference")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-fp32 VALU forms are kept out of every code object.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(INCLUDE), "-I", str(CSRC),
]
HOST_CXX = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-parameter", "-I", str(INCLUDE), "-I", str(CSRC)]
HOST_CC = ["gcc", "-std=gnu11", "-O2", "-fPIC", "-Wall", "-Wno-unused-parameter", "-I", str(INCLUDE), "-I", str(CSRC)]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def hip_sources() -> list[Path]:
    return sorted((CSRC / "device").glob("*.hip"))


def build_hip() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libdnaln_hip.so"
    objs = []
    objdir = ROOT / "_build" / "hip"
    objdir.mkdir(parents=True, exist_ok=True)
    for src in hip_sources():
        obj = objdir / (src.stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        objs.append(obj)
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    objdir = ROOT / "_build" / "host"
    objdir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in sorted((CSRC / "host").glob("*.c")):
        obj = objdir / (src.stem + ".o")
        _run([*HOST_CC, "-c", src, "-o", obj])
        objs.append(obj)
    for src in sorted((CSRC / "host").glob("*.cpp")):
        if src.name.endswith("_main.cpp"):
            continue
        obj = objdir / (src.stem + ".o")
        _run([*HOST_CXX, "-c", src, "-o", obj])
        objs.append(obj)
    out = LIB / "libdnaln_host.so"
    tmp = out.with_suffix(".so.tmp")
    hip = LIB / "libdnaln_hip.so"
    link = ["-L", str(LIB), "-ldnaln_hip", f"-Wl,-rpath,$ORIGIN"] if hip.exists() else []
    _run(["g++", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp, *link, "-lz", "-lm"])
    os.replace(tmp, out)
    mains = sorted((CSRC / "host").glob("*_main.cpp"))
    if mains and hip.exists():
        BIN.mkdir(parents=True, exist_ok=True)
        for m in mains:
            exe = BIN / ("dnaln-amd" if m.stem == "cli_main" else m.stem.replace("_main", ""))
            _run([*HOST_CXX, m, "-o", exe, "-L", str(LIB), "-ldnaln_host", "-ldnaln_hip",
                  "-Wl,-rpath,$ORIGIN/../lib", "-lz", "-lm"])
    # Test drivers of the C++ interface (tests/cpp/*.cpp -> tests/cpp/bin/).
    checks = sorted((ROOT / "tests" / "cpp").glob("*.cpp"))
    if checks and hip.exists():
        (ROOT / "tests" / "cpp" / "bin").mkdir(parents=True, exist_ok=True)
        for m in checks:
            _run([*HOST_CXX, m, "-o", ROOT / "tests" / "cpp" / "bin" / m.stem, "-L", str(LIB), "-ldnaln_host",
                  "-ldnaln_hip", "-Wl,-rpath,$ORIGIN/../../../dnaln_amd/lib", "-lz", "-lm"])
    return out


def build_emu() -> Path:
    """CPU builds of the device code (test infrastructure: kernel logic checked without a GPU).

    libdnaln_emu.so holds the per-lane programs; bin/graph_emu is the whole library (HIP runtime
    calls and kernels) over the CPU emulation of HIP in csrc/emu/hip, with AddressSanitizer and
    exact-size device buffers, so the batched form's kernel sequence (direct and captured) runs
    with every access bounds-checked."""
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libdnaln_emu.so"
    tmp = out.with_suffix(".so.tmp")
    _run([*HOST_CXX, "-Wno-unknown-pragmas", "-O2", "-shared", CSRC / "emu" / "emu.cpp", "-o", tmp])
    os.replace(tmp, out)
    BIN.mkdir(parents=True, exist_ok=True)
    exe = BIN / "graph_emu"
    tmp = exe.with_suffix(".tmp")
    _run(["g++", "-std=c++20", "-O1", "-g1", "-fsanitize=address", "-fno-omit-frame-pointer", "-DDN_EMU",
          "-DDN_GRAPH_CAPTURE", "-Wno-attributes", "-Wno-unknown-pragmas", "-I", CSRC / "emu", "-I", INCLUDE,
          "-I", CSRC, "-x", "c++", CSRC / "device" / "dnaln_hip.hip", "-x", "none", CSRC / "host" / "host_api.cpp",
          CSRC / "emu" / "graph_emu_main.cpp", "-o", tmp, "-lz", "-pthread"])
    os.replace(tmp, exe)
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", ROOT / "oracle"])
    if (REFERENCE / "src" / "mapper.cpp").exists():
        _run(["make", "-s", "-j8", "-f", ROOT / "oracle" / "ref.mk", f"REF={REFERENCE}"])


def build_all() -> None:
    build_hip()
    build_host()
    build_emu()
    build_oracle()


def main(argv: list[str]) -> int:
    parts = argv or ["hip", "host", "emu", "oracle"]
    for p in parts:
        {"hip": build_hip, "host": build_host, "emu": build_emu, "oracle": build_oracle}[p]()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
