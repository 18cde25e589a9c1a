"""dmi_t2map_f64_device_noalloc (no device allocation in the call; buffers in the device's static scratch) against
dmi_t2map_f64_device on the same device arrays: identical outputs on two calls in a row, the second one smaller and
with other options (the grids in the static scratch are rebuilt). The first volume is larger than one selection
chunk (262144 voxels), so the chunked voxel selection runs too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CALLS = [
    # shape, nTE, TE, nT2, Reg, nRefAngles, extra, mask fraction
    ((96, 96, 32), 48, 10e-3, 40, "gcv", 32, {}, 0.3),
    ((9, 7, 3), 32, 10e-3, 40, "chi2", 32, dict(Chi2Factor=1.02), 0.8),
]


def _outputs(dmi, torch, opts):
    from decaes_amd import device as D

    return D.alloc_t2map_outputs(opts, device="cuda")


def test_noalloc_matches_device_entry(dmi):
    import torch
    from decaes_amd import device as D

    for shape, m, TE, n, reg, K, extra, frac in CALLS:
        print(f"[noalloc] shape={shape} nTE={m} {reg}", flush=True)
        img, nin = dmi.synth_mse(shape, m, TE, mask_fraction=frac, seed=31)
        opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=(10e-3, 2.0), Reg=reg, nRefAngles=K,
                                           SaveRegParam=True, SaveResidualNorm=True, SaveDecayCurve=True, Silent=True,
                                           **extra)
        # (X,Y,Z,m) column-major == contiguous (m,Z,Y,X)
        dimg = torch.from_numpy(np.ascontiguousarray(img.transpose(3, 2, 1, 0))).to("cuda")
        ref = D.alloc_t2map_outputs(opts, device="cuda")
        got = D.alloc_t2map_outputs(opts, device="cuda")
        s_ref = D.t2map_device(dimg, opts, ref)
        s_got = D.t2map_device(dimg, opts, got, noalloc=True)
        torch.cuda.synchronize()
        assert s_got["voxels"] == s_ref["voxels"] == nin
        for k in ref:
            a, b = got[k].cpu().numpy(), ref[k].cpu().numpy()
            assert np.array_equal(a, b, equal_nan=True), k
        assert s_got["nnls_solves"] == s_ref["nnls_solves"]


def test_noalloc_rejects_sizes_beyond_limits(dmi):
    import torch
    from decaes_amd import _lib, api
    from decaes_amd import device as D

    img, _ = dmi.synth_mse((2, 2, 1), 32, 10e-3, mask_fraction=1.0, seed=1)
    opts = dmi.T2mapOptions.from_image(img, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="none", Silent=True)
    out = D.alloc_t2map_outputs(opts, device="cuda")
    dimg = torch.from_numpy(np.ascontiguousarray(img.transpose(3, 2, 1, 0))).to("cuda")
    oc = api.t2map_opts_c(opts)
    oc.nT2 = 65
    ctx = _lib.context(0)
    oo = _lib.T2mapOutC(*(D._dptr(out.get(k)) for k in ("dist", "gdn", "ggm", "gva", "fnr", "snr", "alpha",
                                                        "resnorm", "decaycurve", "mu", "chi2factor", "decaybasis")))
    import ctypes as C
    rc = _lib.load().dmi_t2map_f64_device_noalloc(ctx.handle, C.c_void_p(dimg.data_ptr()), C.byref(oc), None,
                                                  C.byref(oo))
    assert rc != 0
    assert b"exceeds" in _lib.load().dmi_last_error(ctx.handle)


def _t2part_outputs(torch, opts):
    X, Y, Z = opts.MatrixSize
    return {k: torch.full((Z, Y, X), float("nan"), dtype=torch.float64, device="cuda")
            for k in ("sfr", "sgm", "mfr", "mgm")}


GRAPH_CASES = {
    # shape, nTE, TE, nT2, T2Range, Reg, nRefAngles, extra options, mask fraction, B1 map
    # two selection chunks: two selections and two voxel-kernel launches that read the selected count on the device
    "gcv_two_chunks": ((96, 96, 32), 48, 10e-3, 40, (10e-3, 2.0), "gcv", 32, {}, 0.3, False),
    # the split schedule: workspace rows beyond the LDS part in the workgroups' global slots
    "lcurve_split": ((60, 60, 4), 56, 7.9e-3, 60, (8e-3, 2.0), "lcurve", 64, dict(schedule="split"), 0.8, False),
    # a B1 map: maps["alpha"] holds the map on every voxel, at or below the threshold too, at every replay
    "chi2_b1": ((48, 48, 8), 32, 10e-3, 40, (10e-3, 2.0), "chi2", 24, dict(Chi2Factor=1.02), 0.5, True),
}


@pytest.mark.parametrize("name", list(GRAPH_CASES))
def test_noalloc_graph_replays_match_eager_calls(dmi, name):
    """dmi_t2map_f64_device_noalloc and dmi_t2part_f64_device recorded into a HIP graph (torch.cuda.graph, the
    library on the capturing stream through dmi_set_stream): every replay equals the eager calls bit for bit, with
    the program's own eager work and waits between replays, and a replay after the image changed in place equals
    eager calls on the new image."""
    import torch
    from decaes_amd import device as D

    shape, m, TE, n, T2R, reg, K, extra, frac, b1 = GRAPH_CASES[name]
    print(f"[graph] case={name} shape={shape} nTE={m} {reg}", flush=True)
    img, nin = dmi.synth_mse(shape, m, TE, mask_fraction=frac, seed=37)
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K, SaveRegParam=True,
                                       SaveResidualNorm=True, SaveDecayCurve=True, Silent=True, **extra)
    popts = dmi.T2partOptions.from_t2map(opts, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0))
    dimg = torch.from_numpy(np.ascontiguousarray(img.transpose(3, 2, 1, 0))).to("cuda")
    dalpha = None
    if b1:
        dalpha = torch.from_numpy(np.random.default_rng(5).uniform(130, 180, shape[::-1])).to("cuda")

    def eager():
        out, parts = D.alloc_t2map_outputs(opts, device="cuda"), _t2part_outputs(torch, opts)
        if b1:
            out["alpha"].copy_(dalpha)
        s = D.t2map_device(dimg, opts, out, alpha=dalpha, noalloc=True)
        D.t2part_device(out["dist"], popts, parts)
        torch.cuda.synchronize()
        return s, {k: v.cpu().numpy() for k, v in {**out, **parts}.items()}

    s_ref, ref = eager()
    assert s_ref["voxels"] == nin
    if extra.get("schedule") == "split":
        assert s_ref["lds_rows"] < m + n
    if b1:
        assert nin < np.prod(shape) and np.array_equal(ref["alpha"], dalpha.cpu().numpy())
    out, parts = D.alloc_t2map_outputs(opts, device="cuda"), _t2part_outputs(torch, opts)
    g = D.capture(dimg, opts, out, alpha=dalpha, t2part=(popts, parts))
    other = torch.linspace(0.0, 1.0, 1 << 20, dtype=torch.float64, device="cuda")
    for r in range(3):
        print(f"[graph] replay {r}", flush=True)
        for t in list(out.values()) + list(parts.values()):
            t.fill_(7.0)
        other = torch.sin(other) + 0.5
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, v in {**out, **parts}.items():
            assert np.array_equal(v.cpu().numpy(), ref[k], equal_nan=True), (r, k)
    dimg.mul_(1.5)
    _, ref2 = eager()
    assert not np.array_equal(ref2["gdn"], ref["gdn"], equal_nan=True)
    print("[graph] replay on the changed image", flush=True)
    g.replay()
    torch.cuda.synchronize()
    for k, v in {**out, **parts}.items():
        assert np.array_equal(v.cpu().numpy(), ref2[k], equal_nan=True), k


# the refusal comes before anything is recorded, so torch reports the captured graph as empty
@pytest.mark.filterwarnings("ignore:The CUDA Graph is empty")
def test_capture_needs_the_grids_built_outside_capture(dmi):
    """Under capture, options whose grids are not in the static scratch are refused before anything is recorded."""
    import torch
    from decaes_amd import device as D

    print("[graph] refusal under capture", flush=True)
    img, nin = dmi.synth_mse((4, 4, 2), 32, 10e-3, mask_fraction=1.0, seed=2)
    opts = dmi.T2mapOptions.from_image(img, TE=10e-3, nT2=36, T2Range=(10e-3, 2.0), Reg="none", nRefAngles=12,
                                       Silent=True)
    out = D.alloc_t2map_outputs(opts, device="cuda")
    dimg = torch.from_numpy(np.ascontiguousarray(img.transpose(3, 2, 1, 0))).to("cuda")
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(AssertionError, match="outside capture first"):
        with torch.cuda.graph(g):
            D.t2map_device(dimg, opts, out, noalloc=True, stream=torch.cuda.current_stream())
    # the context runs on its own stream again and an eager call works
    s = D.t2map_device(dimg, opts, out, noalloc=True)
    assert s["voxels"] == nin
