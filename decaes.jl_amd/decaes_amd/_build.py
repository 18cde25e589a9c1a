"""Builds libdecaes_amd.so (gfx950 kernels + C ABI) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent            # decaes.jl_amd/
REPO = ROOT.parent
CSRC = ROOT / "csrc"
LIB = PKG / "libdecaes_amd.so"
# a C host program of the C ABI (csrc/dmi_run.c), linked against LIB
RUN = PKG / "dmi_run"
# Same sources built with -DDMI_PHASE_PROFILE: per-phase shader-clock counters (decaes_amd.phase_profile)
PHASE_LIB = PKG / "libdecaes_amd_phase.so"

# dmi_kernels_split.hip: the voxel kernel of the split schedule (dmi_kernels.hip compiled once more); the instrumented
# build keeps both schedules in dmi_kernels.hip
SOURCES = ["dmi_kernels.hip", "dmi_kernels_split.hip", "dmi_host.cpp", "dmi_api.cpp", "dmi_server.cpp", "dmi_synth.cpp"]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # fma only where the reference writes muladd; never contract elsewhere
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(Path(rocm) / "bin" / "hipcc")


def build(verbose: bool = False, phase_profile: bool = False) -> Path:
    """Compile every source into decaes_amd/libdecaes_amd.so (always recompiles). With phase_profile, build the
    instrumented variant decaes_amd/libdecaes_amd_phase.so instead."""
    objs, cmds = [], []
    obj_dir = ROOT / ("_build_phase" if phase_profile else "_build")
    obj_dir.mkdir(exist_ok=True)
    out = PHASE_LIB if phase_profile else LIB
    extra = ["-DDMI_PHASE_PROFILE"] if phase_profile else []
    for src in SOURCES:
        if phase_profile and src == "dmi_kernels_split.hip":
            continue
        obj = obj_dir / (src + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *extra, "-c", "-I", str(REPO / "include"), "-I", str(CSRC), str(CSRC / src),
               "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = [c for c in cmd if c not in ("--offload-arch=gfx950",)]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(str(obj))
    # the sources compile independently: one hipcc process each, at most MAX_JOBS (or the CPU count) at once
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", 0)) or os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            f.result()
    tmp = out.with_suffix(".so.tmp")
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp), "-lpthread"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, out)
    if not phase_profile:
        run = [os.environ.get("CC", "gcc"), "-O2", "-std=c11", "-Wall", "-I", str(REPO / "include"),
               str(CSRC / "dmi_run.c"), "-o", str(RUN.with_suffix(".tmp")), "-L", str(PKG), "-ldecaes_amd",
               "-Wl,-rpath,$ORIGIN", "-Wl,--allow-shlib-undefined", "-lm"]
        if verbose:
            print(" ".join(run), flush=True)
        subprocess.run(run, check=True)
        os.replace(RUN.with_suffix(".tmp"), RUN)
    return out


if __name__ == "__main__":
    import sys

    print(build(verbose=True, phase_profile="--phase" in sys.argv[1:]))
