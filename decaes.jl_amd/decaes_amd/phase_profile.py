"""Per-phase cycle breakdown of the voxel kernel, from the instrumented build libdecaes_amd_phase.so.

Each wave accumulates shader-clock cycles (clock64) and call counts per phase in LDS and adds them to a
device array at exit; the table divides by the number of voxels processed. Phases nest: the NNLS inner
steps (construct / apply / trisolve / delete / dual) are inside nnls, nnls and init are inside each solve,
and the solves run inside the flip-angle search and the regularization search.

    python decaes.jl_amd/decaes_amd/_build.py --phase      # build the instrumented library
    python -m decaes_amd.phase_profile gcv lcurve          # (with decaes.jl_amd on sys.path)
"""
from __future__ import annotations

import ctypes as C
import sys
import time

from . import _build, _lib

# (shape, nTE, TE, nT2, T2Range, Reg, nRefAngles, compartments): slabs of the 240x240 volumes
WORKLOADS = {
    "gcv": ((240, 240, 4), 48, 10e-3, 40, (10e-3, 2.0), "gcv", 32, 3),
    "lcurve": ((240, 240, 2), 56, 7.9e-3, 60, (8e-3, 2.0), "lcurve", 64, 4),
    "chi2": ((64, 64, 8), 32, 10e-3, 40, (10e-3, 2.0), "chi2", 32, 3),
}
PHASES = [(1, "construct"), (2, "apply"), (3, "trisolve"), (4, "delete"), (5, "dual"), (6, "init"), (7, "nnls"),
          (16, " select+constr"), (18, " post-apply"), (17, " feasibility"),
          (10, "flip search"), (11, "voxel basis"), (13, "gram eigs"), (8, " gram matrix"), (9, " tridiagonal"),
          (21, " eig brackets"), (12, " sturm"), (14, "regularization"),
          (15, "reg + save")]
NPH = 24


def run(names):
    _lib.LIB_PATH = _build.PHASE_LIB
    L = _lib.load()
    L.dmi_phase_read.argtypes = [C.POINTER(C.c_ulonglong)]
    L.dmi_tikh_rows_read.argtypes = [C.POINTER(C.c_ulonglong)]
    L.dmi_eig_steps_read.argtypes = [C.POINTER(C.c_ulonglong)]
    import decaes_amd as d

    for wl in names:
        shape, m, TE, n, T2R, reg, K, nc = WORKLOADS[wl]
        img, nin = d.synth_mse(shape, m, TE, ncomp=nc, mask_fraction=0.6, seed=1)
        opts = d.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K, Chi2Factor=1.02,
                                         Silent=True)
        d.T2mapSEcorr(img, opts)  # warm-up: grid bases and buffers
        L.dmi_phase_reset()
        t = time.time()
        d.T2mapSEcorr(img, opts)
        dt = time.time() - t
        st = _lib.context().stats()
        buf = (C.c_ulonglong * (2 * NPH))()
        L.dmi_phase_read(buf)
        voxel = buf[10] + buf[11] + buf[15]
        print(f"== {wl}: {shape} nTE={m} nT2={n} Reg={reg} K={K}: {nin} voxels, kernel {st['kernel_ms']:.1f} ms "
              f"({nin / st['kernel_ms'] * 1e3:.0f} voxels/s), wall {dt:.2f} s, {st['nnls_solves'] / nin:.1f} "
              "NNLS solves/voxel")
        print(f"   {'phase':>15s} {'cycles/voxel':>13s} {'% voxel':>8s} {'calls/voxel':>12s} {'cycles/call':>12s}")
        for k, name in PHASES:
            if buf[NPH + k] == 0:
                continue
            print(f"   {name:>15s} {buf[k] / nin:13.0f} {100 * buf[k] / max(voxel, 1):7.1f}% "
                  f"{buf[NPH + k] / nin:12.2f} {buf[k] / buf[NPH + k]:12.0f}", flush=True)
        steps = (C.c_ulonglong * 32)()
        L.dmi_eig_steps_read(steps)
        if sum(steps):
            hist = ", ".join(f"{k}: {steps[k]}" for k in range(32) if steps[k])
            print(f"   GCV eigenvalue Laguerre steps per voxel: {hist}", flush=True)
        rows = (C.c_ulonglong * 65)()
        L.dmi_tikh_rows_read(rows)
        tot = sum(rows)
        if tot:
            cum, p50, p99, mx = 0, None, None, max(i for i in range(65) if rows[i])
            for i in range(65):
                cum += rows[i]
                if p50 is None and cum >= 0.5 * tot:
                    p50 = i
                if p99 is None and cum >= 0.99 * tot:
                    p99 = i
            over = ", ".join(f">{c}: {sum(rows[c + 1:]) / tot:.2e}" for c in (16, 24, 32, 40, 48) if c < n)
            print(f"   Tikhonov rows appended per solve ({tot} solves): median {p50}, p99 {p99}, max {mx}; "
                  f"fraction of solves {over}", flush=True)


if __name__ == "__main__":
    run(sys.argv[1:] or ["gcv", "lcurve"])
