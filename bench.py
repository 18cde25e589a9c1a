"""Throughput of the gfx950 T2mapSEcorr + T2partSEcorr path on one GPU.

Default workload: the 240x240x113 56-echo volume (TE 7.9 ms, nT2 60 over 8 ms-2 s, L-curve regularization, 64 flip
angles, 4 compartments, 60% ellipsoid mask) in mixed precision (FP32 EPG bases, FP64 NNLS and regularization). The
volume is resident in HBM as 8 z-slabs; one step = T2mapSEcorr over one slab (flip-angle search, EPG bases,
regularized NNLS, maps) followed by T2partSEcorr on its distribution, steps cycling through the slabs.
value = processed (above-threshold) voxels per second over the timed steps. The JSON line also carries the plain
FP64 path timed over the whole volume (plain_fp64), the metric's accuracy clause (every voxel's MWF of the mixed run
against the plain path, and the plain path's own change under a 1-ulp input change), the p50/p99 latency of a served
stream below capacity, the roofline object and the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c3chi2|c2|c1] [--precision mixed|fp64]
                    [--slabs N] [--no-cpu-baseline] [--no-served] [--no-plain] [--no-ulp]
    python bench.py --copy-probe BYTES    # plain device copy of known size, the HBM-counter calibration run
    python bench.py --serve [--qps 2,8,32] [--requests 16] [--max-slices 8]   # served stream: p50/p99 latency
    python bench.py --batch [--volumes 64]   # batched uneven workload of seeded volumes through the request loop
    python bench.py --cli [--workload c2]    # the command line end to end: NIfTI in, MAT v7.3 out
    python bench.py --host [--workload c2] --precision fp64 --steps 3 --warmup 1   # host arrays, PCIe included
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "decaes.jl_amd"))

WORKLOADS = {
    # plain FP64 drop-in volume: 240x240x48, nTE 48, TE 10 ms, nT2 40 over 10 ms-2 s, gcv, 32 angles, 60% mask
    "c2": dict(shape=(240, 240, 48), nTE=48, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="gcv", nRefAngles=32,
               ncomp=3, mask=0.6, Chi2Factor=None, slabs=1),
    # reference-CPU-sized case: 32x32x4, nTE 32, chi2 1.02, 32 angles, every voxel
    "c1": dict(shape=(32, 32, 4), nTE=32, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="chi2", nRefAngles=32,
               ncomp=3, mask=1.0, Chi2Factor=1.02, slabs=1),
    # 240x240x113, nTE 56, TE 7.9 ms, nT2 60 over 8 ms-2 s, lcurve, 64 angles, 4 compartments
    "c3": dict(shape=(240, 240, 113), nTE=56, TE=7.9e-3, nT2=60, T2Range=(8e-3, 2.0), Reg="lcurve", nRefAngles=64,
               ncomp=4, mask=0.6, Chi2Factor=None, slabs=8),
    # the same volume with chi2 1.02, the regularization of the MATLAB toolbox that the reference's published
    # whole-volume timings compare against
    "c3chi2": dict(shape=(240, 240, 113), nTE=56, TE=7.9e-3, nT2=60, T2Range=(8e-3, 2.0), Reg="chi2", nRefAngles=64,
                   ncomp=4, mask=0.6, Chi2Factor=1.02, slabs=8),
    # the legacy path (legacy = true: the Dierckx spline flip-angle surrogate over the reference's legacy default of
    # 8 angles, and the chi2 doubling search) on the 240x240x48 volume with chi2 1.02
    "c2legacy": dict(shape=(240, 240, 48), nTE=48, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="chi2", nRefAngles=8,
                     ncomp=3, mask=0.6, Chi2Factor=1.02, slabs=1, legacy=True),
}
SPWIN = (None, 40e-3)  # short-peak window upper edge (lower edge = T2Range[0])
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
METRIC = "voxels/s (T2map+T2part) at MWF err\u22641e-3 vs plain; served p50/p99 latency"
TRAFFIC_FILE = {("c2", "fp64"): ROOT / "profiles" / "t2map_kernel_gcv_240x240x48_hbm_traffic.md",
                ("c3", "mixed"): ROOT / "profiles" / "t2map_kernel_lcurve_240x240x113_mixed_hbm_traffic.md",
                ("c3", "fp64"): ROOT / "profiles" / "t2map_kernel_lcurve_240x240x113_mixed_hbm_traffic.md"}


def measured_traffic(workload: str, precision: str):
    """HBM bytes per t2map_kernel launch (one step) from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    (calibrated against a plain 1 GiB copy), or None when no such measurement exists for the workload."""
    import re

    f = TRAFFIC_FILE.get((workload, precision))
    if f is None or not f.exists():
        return None
    m = re.search(rf"^traffic_bytes_per_launch\[{workload}-{precision}\]: (\d+)", f.read_text(), re.M)
    if m is None and workload == "c2":
        m = re.search(r"^traffic_bytes_per_launch\[c2\]: (\d+)", f.read_text(), re.M)
    return int(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(img, opts_kw, seconds: float):
    """The FP64 oracle on host cores over a bounded, seeded sample of the same in-mask voxels."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib

    oracle_lib.build()
    w = opts_kw
    shape = img.shape[:3]
    legacy = w.get("legacy", False)
    oo = oracle_lib.make_t2map_opts(shape, w["nTE"], w["TE"], w["nT2"], w["T2Range"], w["Reg"],
                                    nRefAngles=w["nRefAngles"], nRefAnglesMin=w["nRefAngles"] if legacy else None,
                                    Chi2Factor=w["Chi2Factor"], legacy=legacy)
    # the oracle is FP64 throughout: the baseline is the reference's plain CPU path
    first = img[..., 0].reshape(-1, order="F")
    inmask = np.flatnonzero(first > 0.0)
    rng = np.random.default_rng(1234)
    order = rng.permutation(inmask)
    # every host core this process may use, up to the CPU share of one GPU on a shared GPU host (16)
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(avail, 16))
    done, t0, chunk = 0, time.perf_counter(), 256 * threads
    while time.perf_counter() - t0 < seconds and done < len(order):
        vox = np.sort(order[done:done + chunk])
        oracle_lib.t2map_subset(img, oo, vox, nthreads=threads)
        done += len(vox)
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="voxels/s", cores=threads, kind="port",
                sample=f"{done} in-mask voxels drawn (seed 1234) from the same volume, {dt:.1f} s of CPU work",
                cpu=f"{cpu_model()}, {threads} of {avail} logical CPUs available to the process, "
                    f"{os.cpu_count()} on the host; the FP64 oracle (a C restatement of the reference), "
                    "64-voxel blocks per thread as the reference's worker pool")


def copy_probe(nbytes: int):
    """Three plain device-to-device copies of `nbytes`: the known-size run that scales FETCH_SIZE/WRITE_SIZE."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / 3
    print(json.dumps({"copy_bytes": a.numel() * 4, "GBps_read_plus_write": 2 * a.numel() * 4 / dt / 1e9}), flush=True)


def alone_rate(items, tag="alone") -> dict:
    """The same requests one at a time through the device-array entry points (image and outputs resident in HBM):
    total voxels over the summed device time of T2mapSEcorr + T2partSEcorr, and over the summed kernel time.
    items: iterable of (image, T2mapOptions, T2partOptions)."""
    import numpy as np
    import torch

    import decaes_amd as dmi
    from decaes_amd.device import alloc_t2map_outputs, t2map_device, t2part_device

    ctx = dmi.context(0)
    t_k, t_all, nv = 0.0, 0.0, 0
    for img, o, po in items:
        d_img = torch.from_numpy(np.ascontiguousarray(img.transpose(3, 2, 1, 0))).to("cuda:0")
        out = alloc_t2map_outputs(o, "cuda:0")
        X, Y, Z = img.shape[:3]
        pout = {k: torch.full((Z, Y, X), float("nan"), dtype=torch.float64, device="cuda:0")
                for k in ("sfr", "sgm", "mfr", "mgm")}
        torch.cuda.synchronize()
        ta = time.perf_counter()
        stt = t2map_device(d_img, o, out, ctx=ctx)
        t2part_device(out["dist"], po, pout, ctx=ctx)
        torch.cuda.synchronize()
        t_all += time.perf_counter() - ta
        t_k += stt["kernel_ms"] * 1e-3
        nv += stt["voxels"]
        log(f"[{tag}] {tuple(img.shape[:3])} nTE={o.nTE} {o.Reg}: {stt['voxels']} voxels, "
            f"kernel {stt['kernel_ms']:.0f} ms")
        del d_img, out, pout
    return {"voxels": nv, "kernel_voxels_per_s": nv / t_k, "device_voxels_per_s": nv / t_all}


def serve_levels(args) -> list:
    """Served stream through the request loop: Poisson arrivals at each QPS of requests of 1..max_slices slices of
    240x240 (nTE in {32, 48, 56}, Reg in {chi2, gcv, lcurve}, T2part on each); p50/p99 latency and voxels/s per
    QPS."""
    import threading

    import numpy as np

    import decaes_amd as dmi
    from decaes_amd.server import Server

    rng = np.random.default_rng(args.seed)
    pool = {}
    for nte, te in ((32, 10e-3), (48, 10e-3), (56, 7.9e-3)):
        img, _ = dmi.synth_mse((240, 240, args.max_slices), nte, te, ncomp=4 if nte == 56 else 3, mask_fraction=0.6,
                               seed=args.seed + nte)
        pool[nte] = (img, te)
    levels = []
    for qps in [float(q) for q in args.qps.split(",")]:
        reqs = []
        for i in range(args.requests):
            nte = int(rng.choice([32, 48, 56]))
            reg = str(rng.choice(["chi2", "gcv", "lcurve"]))
            ns = int(rng.integers(1, args.max_slices + 1))
            z0 = int(rng.integers(0, args.max_slices - ns + 1))
            img, te = pool[nte]
            sub = np.asfortranarray(img[:, :, z0:z0 + ns, :])
            o = dmi.T2mapOptions.from_image(sub, TE=te, nT2=40 if nte < 56 else 60,
                                            T2Range=(10e-3, 2.0) if nte < 56 else (8e-3, 2.0), Reg=reg,
                                            nRefAngles=32 if nte < 56 else 64, Chi2Factor=1.02, Silent=True)
            po = dmi.T2partOptions.from_t2map(o, SPWin=(o.T2Range[0], 40e-3), MPWin=(40e-3, o.T2Range[1]))
            reqs.append((sub, o, po))
        gaps = rng.exponential(1.0 / qps, size=len(reqs))
        lat = []
        with Server(order=args.order) as srv:
            handles = []
            t0 = time.perf_counter()
            waiter_done = threading.Event()

            def waiter():
                k = 0
                while k < len(reqs):
                    if k < len(handles):
                        _, _, _, info = srv.wait(handles[k])
                        lat.append(info["latency_ms"])
                        log(f"[serve] qps {qps}: request {k} done, {info['voxels']} voxels, "
                            f"latency {info['latency_ms']:.0f} ms")
                        k += 1
                    else:
                        time.sleep(0.001)
                waiter_done.set()

            th = threading.Thread(target=waiter)
            th.start()
            t_next = t0
            for (sub, o, po), g in zip(reqs, gaps):
                t_next += g
                time.sleep(max(0.0, t_next - time.perf_counter()))
                handles.append(srv.submit(sub, o, po))
            th.join()
            wall = time.perf_counter() - t0
            st = srv.stats()
        levels.append({"qps": qps, "requests": len(reqs), "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                       "voxels": st["voxels"], "voxels_per_s_busy": st["voxels"] / (st["busy_ms"] * 1e-3),
                       "voxels_per_s_wall": st["voxels"] / wall, "failed": st["failed"],
                       "order": getattr(args, "order", "shortest")})
        if getattr(args, "compare_alone", False):
            a = alone_rate(reqs)
            a["served_busy_over_alone"] = levels[-1]["voxels_per_s_busy"] / a["device_voxels_per_s"]
            levels[-1]["alone"] = a
        log(f"[serve] qps {qps}: p50 {st['p50_ms']:.0f} ms p99 {st['p99_ms']:.0f} ms, {st['voxels']} voxels, "
            f"{levels[-1]['voxels_per_s_busy']:.0f} voxels/s busy")
    return levels


def serve(args):
    """One JSON line with the served stream's p50/p99 latency and voxels/s per QPS (serve_levels)."""
    levels = serve_levels(args)
    print(json.dumps({"metric": "served latency (T2map+T2part requests)", "unit": "ms", "levels": levels,
                      "config": {"requests_per_level": args.requests, "slices": f"1..{args.max_slices} of 240x240",
                                 "nTE": [32, 48, 56], "Reg": ["chi2", "gcv", "lcurve"], "seed": args.seed},
                      "data": "synthetic"}), flush=True)


def batch(args):
    """Batched uneven workload: seeded volumes with matrix sizes from {96x96x16, 128x128x32, 240x240x48}, nTE from
    {24, 32, 48, 56}, TE from {7, 10} ms, mask fraction U(0.2, 0.8), Reg from {chi2, gcv, lcurve}, nT2 40 and the
    default 64 flip angles, all submitted to the request loop (at most `window` in flight so host memory stays
    bounded); prints one JSON line with the aggregate voxels/s."""
    import numpy as np

    import decaes_amd as dmi
    from decaes_amd.server import Server

    rng = np.random.default_rng(args.seed)
    sizes = [(96, 96, 16), (128, 128, 32), (240, 240, 48)]
    specs = []
    for v in range(args.volumes):
        specs.append(dict(shape=sizes[int(rng.integers(0, 3))], nTE=int(rng.choice([24, 32, 48, 56])),
                          TE=float(rng.choice([7e-3, 10e-3])), mask=float(rng.uniform(0.2, 0.8)),
                          Reg=str(rng.choice(["chi2", "gcv", "lcurve"])), seed=args.seed * 1000 + v))
    voxels, per_reg = 0, {}
    t_synth = 0.0
    with Server(max_streams=4) as srv:
        inflight = []
        t0 = time.perf_counter()
        for i, sp in enumerate(specs):
            ts = time.perf_counter()
            img, nin = dmi.synth_mse(sp["shape"], sp["nTE"], sp["TE"], ncomp=3, mask_fraction=sp["mask"], seed=sp["seed"])
            t_synth += time.perf_counter() - ts
            o = dmi.T2mapOptions.from_image(img, TE=sp["TE"], nT2=40, T2Range=(10e-3, 2.0), Reg=sp["Reg"],
                                            Chi2Factor=1.02, Silent=True)
            po = dmi.T2partOptions.from_t2map(o, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0))
            inflight.append((srv.submit(img, o, po), sp))
            while len(inflight) >= args.window or (i == len(specs) - 1 and inflight):
                h, spd = inflight.pop(0)
                _, _, _, info = srv.wait(h)
                voxels += info["voxels"]
                per_reg[spd["Reg"]] = per_reg.get(spd["Reg"], 0) + info["voxels"]
                log(f"[batch] {spd['shape']} nTE={spd['nTE']} {spd['Reg']}: {info['voxels']} voxels, "
                    f"latency {info['latency_ms']:.0f} ms")
        wall = time.perf_counter() - t0
        st = srv.stats()
    alone = None
    if args.compare_alone:
        def items():
            for sp in specs:
                img, _ = dmi.synth_mse(sp["shape"], sp["nTE"], sp["TE"], ncomp=3, mask_fraction=sp["mask"],
                                       seed=sp["seed"])
                o = dmi.T2mapOptions.from_image(img, TE=sp["TE"], nT2=40, T2Range=(10e-3, 2.0), Reg=sp["Reg"],
                                                Chi2Factor=1.02, Silent=True)
                yield img, o, dmi.T2partOptions.from_t2map(o, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0))
        alone = alone_rate(items())
        alone["batched_busy_over_alone"] = (voxels / (st["busy_ms"] * 1e-3)) / alone["device_voxels_per_s"]
    print(json.dumps({"metric": "voxels/s (T2map+T2part) batched uneven workload", "value": voxels / wall,
                      "alone": alone,
                      "unit": "voxels/s", "volumes": len(specs), "voxels": voxels, "wall_s": wall,
                      "busy_s": st["busy_ms"] * 1e-3, "voxels_per_s_busy": voxels / (st["busy_ms"] * 1e-3),
                      "host_synthesis_s": t_synth, "voxels_by_reg": per_reg,
                      "config": {"sizes": [list(x) for x in sizes], "nTE": [24, 32, 48, 56], "TE": [7e-3, 10e-3],
                                 "mask": "U(0.2,0.8)", "Reg": ["chi2", "gcv", "lcurve"], "nT2": 40, "nRefAngles": 64,
                                 "window": args.window, "seed": args.seed},
                      "data": "synthetic"}), flush=True)


def cli_run(args):
    """The command line layer end to end on one GPU: the workload's volume is written as a NIfTI file, and
    decaes_amd.cli runs T2mapSEcorr and T2partSEcorr on it and writes the MAT v7.3 results, as the reference's CLI
    does (src/main.jl:292-420); each stage's wall time comes from the run's own log."""
    import tempfile
    from datetime import datetime

    import decaes_amd as dmi
    from decaes_amd import cli, fileio

    w = WORKLOADS[args.workload]
    img, nin = dmi.synth_mse(w["shape"], w["nTE"], w["TE"], ncomp=w["ncomp"], mask_fraction=w["mask"], seed=2024)
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "volume.nii"
        fileio.write_nifti(src, img)
        nbytes_in = src.stat().st_size
        del img
        argv = [str(src), "--T2map", "--T2part", "--TE", repr(w["TE"]), "--nT2", str(w["nT2"]), "--T2Range",
                *map(repr, w["T2Range"]), "--SPWin", repr(w["T2Range"][0]), "40e-3", "--MPWin", "40e-3",
                repr(w["T2Range"][1]), "--Reg", w["Reg"], "--nRefAngles", str(w["nRefAngles"]), "-o",
                str(Path(d) / "out"), "--quiet"]
        if w["Chi2Factor"] is not None:
            argv += ["--RegParams", repr(w["Chi2Factor"])]
        if w.get("legacy"):
            argv += ["--legacy"]
        t = time.perf_counter()
        rc = cli.main(argv)
        wall = time.perf_counter() - t
        if rc != 0:
            raise SystemExit(f"the command line failed ({rc})")
        # stage boundaries from the log's timestamps ("[YYYY-mm-dd HH:MM:SS,mmm] LEVEL message")
        marks = []
        for line in (Path(d) / "out" / "volume.log").read_text().splitlines():
            if line.startswith("["):
                stamp, msg = line[1:].split("] ", 1)
                marks.append((datetime.strptime(stamp, "%Y-%m-%d %H:%M:%S,%f"), msg.split(" ", 1)[1]))
        def at(prefix):
            return next(ts for ts, m in marks if m.startswith(prefix))
        stages = {
            "load_s": (at("Running T2mapSEcorr") - at("Loading input file")).total_seconds(),
            "t2map_s": (at("Saving T2 distribution") - at("Running T2mapSEcorr")).total_seconds(),
            "save_t2map_s": (at("Running T2partSEcorr") - at("Saving T2 distribution")).total_seconds(),
            "t2part_s": (at("Saving T2 parts") - at("Running T2partSEcorr")).total_seconds(),
            "save_t2parts_s": (at("Finished") - at("Saving T2 parts")).total_seconds(),
        }
        out_bytes = sum(f.stat().st_size for f in (Path(d) / "out").iterdir())
    res = {"cli": {"workload": args.workload, "voxels_in_mask": nin, "input_nifti_bytes": nbytes_in,
                   "output_mat_bytes": out_bytes, "wall_s": wall, "voxels_per_s_end_to_end": nin / wall,
                   **stages, "argv": " ".join(argv[1:])}}
    print(json.dumps(res), flush=True)


def host_run(args):
    """The host-array entry point (dmi_t2map_f64 through T2mapSEcorr) on the workload's whole volume, PCIe copies
    included: the image goes in and the maps come back on every call. The context is created before the warmup calls
    and each call's wall time is taken around T2mapSEcorr, beside the library's own kernel and call totals."""
    import numpy as np

    import decaes_amd as dmi

    w = WORKLOADS[args.workload]
    img, nin = dmi.synth_mse(w["shape"], w["nTE"], w["TE"], ncomp=w["ncomp"], mask_fraction=w["mask"], seed=2024)
    extra = {} if w["Chi2Factor"] is None else {"Chi2Factor": w["Chi2Factor"]}
    opts = dmi.T2mapOptions.from_image(img, TE=w["TE"], nT2=w["nT2"], T2Range=w["T2Range"], Reg=w["Reg"],
                                       nRefAngles=w["nRefAngles"], legacy=bool(w.get("legacy")),
                                       precision=args.precision, Silent=True, **extra)
    t = time.perf_counter()
    ctx = dmi.context(0)
    ctx_s = time.perf_counter() - t
    walls, kms, tms, out_bytes = [], [], [], 0
    for i in range(args.warmup + args.steps):
        t = time.perf_counter()
        maps, dist = dmi.T2mapSEcorr(img, opts)
        wall = time.perf_counter() - t
        st = ctx.stats()
        phase = "warmup" if i < args.warmup else "step"
        log(f"[host] {phase} {i}: wall {wall * 1e3:.1f} ms, library {st['total_ms']:.1f} ms, kernel "
            f"{st['kernel_ms']:.1f} ms, {st['voxels']} voxels")
        if i >= args.warmup:
            walls.append(wall)
            kms.append(st["kernel_ms"])
            tms.append(st["total_ms"])
        out_bytes = dist.nbytes + sum(v.nbytes for k, v in maps.items() if v.shape[:3] == tuple(w["shape"]))
        del maps, dist
    wall = float(np.mean(walls))
    res = {"host": {"workload": args.workload, "precision": args.precision, "voxels_in_mask": nin,
                    "image_bytes": img.nbytes, "output_bytes": out_bytes, "context_s": ctx_s, "calls": len(walls),
                    "wall_ms_per_call": wall * 1e3, "library_ms_per_call": float(np.mean(tms)),
                    "kernel_ms_per_call": float(np.mean(kms)), "voxels_per_s_pcie_inclusive": nin / wall,
                    "voxels_per_s_kernel": nin / (float(np.mean(kms)) * 1e-3),
                    "schedule": {"lds_rows": st["lds_rows"], "workgroups_per_cu": st["workgroups_per_cu"]}}}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--slabs", type=int, default=0, help="z-slabs per volume, one per step (0: the workload's own)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison of a mixed run")
    ap.add_argument("--no-ulp", action="store_true", help="skip the plain path's 1-ulp sensitivity")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the short served stream reported beside the value")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES")
    ap.add_argument("--serve", action="store_true")
    ap.add_argument("--qps", default="2,8,32")
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--max-slices", type=int, default=8)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--order", default="shortest", choices=["shortest", "turns"],
                    help="served stream: which request of a stream gets the next chunk")
    ap.add_argument("--batch", action="store_true")
    ap.add_argument("--volumes", type=int, default=64)
    ap.add_argument("--window", type=int, default=6)
    ap.add_argument("--compare-alone", action="store_true",
                    help="--batch / --serve: also run each volume or request alone on the device")
    ap.add_argument("--cli", action="store_true",
                    help="the command line end to end on the workload's volume written as NIfTI (load, T2map, "
                         "T2part, MAT v7.3 output)")
    ap.add_argument("--host", action="store_true",
                    help="the host-array entry point on the workload's whole volume, PCIe copies included")
    args = ap.parse_args()
    # PyTorch's HIP runtime is loaded before the library, which then binds to it (one runtime per process)
    import torch  # noqa: F401

    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.serve:
        if args.gpus != 1:
            print("the served stream runs on one GPU", file=sys.stderr)
            sys.exit(2)
        return serve(args)
    if args.batch:
        if args.gpus != 1:
            print("the batched workload runs on one GPU", file=sys.stderr)
            sys.exit(2)
        return batch(args)
    if args.cli:
        return cli_run(args)
    if args.host:
        return host_run(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); voxels shard across GPUs with no exchange "
              "step, so a multi-GPU figure is N independent single-GPU runs.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import decaes_amd as dmi
    from decaes_amd.device import alloc_t2map_outputs, t2map_device, t2part_device

    w = WORKLOADS[args.workload]
    X, Y, Z = w["shape"]
    nslab = args.slabs if args.slabs > 0 else w["slabs"]
    log(f"[bench] synthesizing {args.workload} {w['shape']} nTE={w['nTE']}")
    t = time.perf_counter()
    img, nin = dmi.synth_mse(w["shape"], w["nTE"], w["TE"], ncomp=w["ncomp"], mask_fraction=w["mask"], seed=2024)
    log(f"[bench] synth done in {time.perf_counter() - t:.1f}s, {nin} voxels in mask")
    dev = torch.device("cuda:0")
    ctx = dmi.context(0)
    # the volume as nslab z-slabs, each a (X,Y,Zs,nTE) column-major array == contiguous (nTE,Zs,Y,X) on the device
    bounds = np.linspace(0, Z, nslab + 1).round().astype(int)
    slabs = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        sub = img[:, :, a:b, :]
        kw = dict(MatrixSize=(X, Y, int(b - a)), nTE=w["nTE"], TE=w["TE"], nT2=w["nT2"], T2Range=w["T2Range"],
                  Reg=w["Reg"], nRefAngles=w["nRefAngles"], Chi2Factor=w["Chi2Factor"], legacy=w.get("legacy", False),
                  Silent=True)
        o = dmi.T2mapOptions(precision=args.precision, **kw)
        o64 = dmi.T2mapOptions(precision="fp64", **kw)
        po = dmi.T2partOptions.from_t2map(o, SPWin=(w["T2Range"][0], 40e-3), MPWin=(40e-3, w["T2Range"][1]))
        d_img = torch.from_numpy(np.ascontiguousarray(sub.transpose(3, 2, 1, 0))).to(dev)
        Zs = int(b - a)
        pout = {k: torch.full((Zs, Y, X), float("nan"), dtype=torch.float64, device=dev) for k in ("sfr", "sgm", "mfr", "mgm")}
        slabs.append(dict(z=(int(a), int(b)), opts=o, opts64=o64, popts=po, img=d_img, out=alloc_t2map_outputs(o, dev),
                          pout=pout))
    torch.cuda.synchronize()

    def run(sl, opts, out, pout):
        st = t2map_device(sl["img"], opts, out, ctx=ctx)
        t2part_device(out["dist"], sl["popts"], pout, ctx=ctx)
        return st

    def step(i):
        sl = slabs[i % nslab]
        return run(sl, sl["opts"], sl["out"], sl["pout"])

    for i in range(args.warmup):
        st = step(i)
        log(f"[bench] warmup {i}: slab {i % nslab}, kernel {st['kernel_ms']:.1f} ms, {st['voxels']} voxels")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step(args.warmup + i)
        stats.append(st)
        log(f"[bench] step {i}: slab {(args.warmup + i) % nslab}, kernel {st['kernel_ms']:.1f} ms, "
            f"total {st['total_ms']:.1f} ms, {st['voxels']} voxels")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nvox = int(sum(s["voxels"] for s in stats))
    kernel_s = float(sum(s["kernel_ms"] for s in stats)) * 1e-3
    flops64 = float(sum(s["fp64_flops"] for s in stats))
    flops32 = float(sum(s["fp32_flops"] for s in stats))
    achieved = flops64 / kernel_s / 1e12
    launches = len(stats)
    # every slab has mixed-precision (or plain) outputs once all slabs were stepped
    ran_all = args.warmup + args.steps >= nslab
    res = {
        "metric": METRIC,
        "value": nvox / dt,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 EPG + f64 NNLS",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {X}x{Y}x{Z} nTE={w['nTE']} TE={w['TE']} nT2={w['nT2']} "
                               f"T2Range={w['T2Range']} Reg={w['Reg']} nRefAngles={w['nRefAngles']} "
                               f"{'legacy=true ' if w.get('legacy') else ''}"
                               f"mask={w['mask']} precision={args.precision}; one step = one of {nslab} z-slabs "
                               f"of the volume (slices {[sl['z'] for sl in slabs]})",
                   "voxels_processed": nvox, "voxels_in_mask": int(nin), "voxels_volume": X * Y * Z,
                   "nnls_solves_per_voxel": float(sum(s["nnls_solves"] for s in stats)) / max(nvox, 1),
                   "fp64_flops_per_voxel": flops64 / max(nvox, 1), "fp32_flops_per_voxel": flops32 / max(nvox, 1),
                   "schedule": {"lds_rows": int(stats[-1]["lds_rows"]), "rows": w["nTE"] + w["nT2"],
                                "workgroups_per_cu": int(stats[-1]["workgroups_per_cu"]),
                                "chosen_by": "autotuner (measured on the first warmup slab)"}},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(args.workload, args.precision),
                     "kernel": f"t2map_kernel<{dmi.api._lib.REG[w['Reg']] + (8 if w.get('legacy') else 0)}>", "kernel_ms": 1e3 * kernel_s / launches,
                     "flops_per_launch": flops64 / launches, "fp32_tflops": flops32 / kernel_s / 1e12},
    }
    if args.precision == "mixed" and ran_all and not args.no_plain:
        # the plain FP64 path over the whole volume (its own voxels/s), then the metric's accuracy clause: every
        # voxel's MWF (sfr) of the timed mixed run against the plain path; beside it the plain path's own MWF change
        # when the image moves by one ulp (the L-curve corner search is discontinuous at the rounding level)
        log("[bench] plain FP64 path over the whole volume")

        def plain(scale=None):
            outs, t_k, nv = [], 0.0, 0
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for sl in slabs:
                o64 = alloc_t2map_outputs(sl["opts64"], dev)
                p64 = {k: torch.full_like(v, float("nan")) for k, v in sl["pout"].items()}
                imgs = sl["img"] if scale is None else sl["img"] * scale
                st64 = run(dict(sl, img=imgs), sl["opts64"], o64, p64)
                t_k += st64["kernel_ms"] * 1e-3
                nv += st64["voxels"]
                outs.append(p64["sfr"].cpu().numpy())
            torch.cuda.synchronize()
            return outs, nv / (time.perf_counter() - t1), nv / t_k

        b64, v64, v64k = plain()
        mixed = [sl["pout"]["sfr"].cpu().numpy() for sl in slabs]

        def err_stats(a, b):
            a, b = np.concatenate([x.ravel() for x in a]), np.concatenate([x.ravel() for x in b])
            ok = np.isfinite(b)
            e = np.abs(a[ok] - b[ok])
            return {"voxels": int(ok.sum()), "max": float(e.max()), "p99": float(np.quantile(e, 0.99)),
                    "p99.9": float(np.quantile(e, 0.999)), "count_gt_1e-3": int(np.sum(e > 1e-3)),
                    "frac_gt_1e-3": float(np.mean(e > 1e-3))}

        res["plain_fp64"] = {"value": v64, "unit": "voxels/s", "kernel_voxels_per_s": v64k,
                             "speed_ratio_mixed_over_plain": res["value"] / v64}
        res["config"]["mwf_err_vs_fp64"] = err_stats(mixed, b64)
        if not args.no_ulp:
            log("[bench] plain FP64 path on the volume scaled by 1 + 2^-52")
            b64p, _, _ = plain(1 + 2.0 ** -52)
            res["config"]["mwf_err_fp64_1ulp"] = err_stats(b64p, b64)
    elif args.precision == "fp64":
        res["config"]["mwf_err_vs_fp64"] = 0.0
    if not args.no_served:
        # served p50/p99 (the metric's latency clause): a Poisson stream through the request loop at a rate below its
        # capacity, so the percentiles are service latencies rather than backlog drain times
        log("[bench] served stream (12 requests of 1-4 slices of 240x240 at QPS 2)")
        try:
            lv = serve_levels(argparse.Namespace(qps="2", requests=12, max_slices=4, seed=7, order=args.order))[0]
            res["served"] = {"qps": lv["qps"], "requests": lv["requests"], "p50_ms": lv["p50_ms"],
                             "p99_ms": lv["p99_ms"], "voxels_per_s_busy": lv["voxels_per_s_busy"],
                             "failed": lv["failed"], "order": lv["order"],
                             "stream": "1-4 slices of 240x240, nTE 32/48/56, chi2/gcv/lcurve"}
        except Exception as e:  # the throughput line is still printed
            res["served"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle on host cores)")
        res["cpu_baseline"] = cpu_baseline(img, w, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
