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
