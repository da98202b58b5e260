"""Builds libqoa_hip.so (gfx950) in-tree with hipcc.

    python -m qoa_amd.build            # compile every source, link the library

Each source under qoa_amd/csrc is compiled on its own (in parallel) into
qoa_amd/_build/, then linked into qoa_amd/libqoa_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libqoa_hip.so"
ARCH = os.environ.get("QOA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # Every FP64 kernel must evaluate the reference's unfused products.
    "-ffp-contract=off",
    # Packed FP32 instructions stay off for every kernel in this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # The register-resident tile keeps its 16 amplitudes in VGPRs only if
    # switch arms that differ in register indices are not merged (merging
    # them leaves a dynamically indexed array in scratch memory).
    "-mllvm", "-simplifycfg-sink-common=false",
    # Op dispatch in the register kernels branches on wave-uniform op
    # words. Structurizing those uniform if/else arms into exec-masked flow
    # blocks leaves phis the register coalescer cannot join, which costs a
    # copy of every live amplitude register per op; uniform regions keep
    # their plain scalar branches.
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I", str(PKG.parent / "include"),
]

# A kernel whose code fits in +-128 KiB needs no long jumps (s_getpc /
# s_add / s_addc / s_setpc for one branch), but the compiler sizes each inline
# asm statement at 16 bytes, about twice the real size of the register
# kernels' arithmetic, and turns branches it believes far into long jumps
# (128 of them in the fused register kernel, most inside its op loop). Telling
# it the branch field reaches twice as far removes them; the assembler still
# rejects any branch whose real offset does not fit ("branch size exceeds
# simm16"), and such a source is compiled again without the option.
NEAR_BRANCHES = ["-mllvm", "-amdgpu-s-branch-bits=17"]


# Headers that kernels compiled at run time (hipRTC, csrc/specialize.cpp)
# include, by the name they are included under.
RTC_HEADERS = [
    ("jit_device.hpp", "jit_device.hpp"),
    ("tile_device.hpp", "tile_device.hpp"),
    ("planner.hpp", "planner.hpp"),
    ("traj_noise.hpp", "traj_noise.hpp"),
    ("../../include/qoa_hip.h", "../../include/qoa_hip.h"),
]


def rtc_headers_source() -> Path:
    """Writes _build/rtc_headers.cpp: the texts of RTC_HEADERS as string
    literals (kRtcHeaderNames / kRtcHeaderTexts), so the library carries the
    device code it compiles at run time."""
    out = BUILD / "rtc_headers.cpp"
    parts = ["// generated by qoa_amd/build.py from the headers named below", "namespace qoa {",
             f"extern const int kRtcHeaderCount = {len(RTC_HEADERS)};",
             "extern const char* const kRtcHeaderNames[] = {"]
    parts += [f'    "{name}",' for name, _ in RTC_HEADERS]
    parts += ["};", "extern const char* const kRtcHeaderTexts[] = {"]
    for _, rel in RTC_HEADERS:
        text = (CSRC / rel).read_text()
        delim = "QOAHDR"
        assert f"){delim}\"" not in text
        # long texts as concatenated literals (compilers limit one literal)
        chunks = [text[i:i + 12000] for i in range(0, len(text), 12000)] or [""]
        parts.append("    " + " ".join(f'R"{delim}({c}){delim}"' for c in chunks) + ",")
    parts += ["};", "}  // namespace qoa", ""]
    text = "\n".join(parts)
    if not out.exists() or out.read_text() != text:
        out.write_text(text)
    return out


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, *NEAR_BRANCHES, "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0 and "branch size exceeds simm16" in res.stderr:
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
        res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources() + [rtc_headers_source()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
           "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
