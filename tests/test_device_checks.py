"""The device-array entry points (decaes_amd.device) refuse tensors whose dtype, layout, size or device differ from
what the kernels read and write, and the request loop refuses host arrays of other sizes than its options'. Both do
so on the host, before any library call, so these run without a GPU: every size here is checked before any device."""
import pytest

torch = pytest.importorskip("torch")


def _opts(dmi, **kw):
    base = dict(MatrixSize=(4, 3, 2), nTE=32, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="chi2", Chi2Factor=1.02,
                Silent=True)
    return dmi.T2mapOptions(**{**base, **kw})


def _image(opts, dtype=torch.float64):
    X, Y, Z = opts.MatrixSize
    return torch.zeros((opts.nTE, Z, Y, X), dtype=dtype)


def test_t2map_device_refuses_mismatched_tensors(dmi):
    from decaes_amd import device as D

    o = _opts(dmi)
    V = 4 * 3 * 2
    out = D.alloc_t2map_outputs(o, device="cpu")
    with pytest.raises(ValueError, match=rf"image must be a contiguous float64 tensor of {V * 32} elements"):
        D.t2map_device(torch.zeros(V * 31, dtype=torch.float64), o, out)
    with pytest.raises(ValueError, match="image .*torch.float32"):
        D.t2map_device(_image(o, torch.float32), o, out)
    with pytest.raises(ValueError, match="not contiguous"):
        D.t2map_device(_image(o).transpose(1, 3), o, out)
    with pytest.raises(ValueError, match="alpha"):
        D.t2map_device(_image(o), o, out, alpha=torch.zeros(V - 1, dtype=torch.float64))
    short = {**out, "dist": torch.zeros(V * 39, dtype=torch.float64)}
    with pytest.raises(ValueError, match=r'out\["dist"\] must be .* of 960 elements'):
        D.t2map_device(_image(o), o, short)
    with pytest.raises(ValueError, match=r'out\["dist"\] is required'):
        D.t2map_device(_image(o), o, {k: v for k, v in out.items() if k != "dist"})
    with pytest.raises(ValueError, match=r'out\["gva"\]'):
        D.t2map_device(_image(o), o, {**out, "gva": torch.zeros(V + 1, dtype=torch.float64)})
    # every size right: the tensors are refused for being on the host, not on the context's device
    with pytest.raises(ValueError, match="must be on cuda:0"):
        D.t2map_device(_image(o), o, out)


def test_t2map_device_output_sizes_follow_the_options(dmi):
    from decaes_amd import device as D

    V = 4 * 3 * 2
    o = _opts(dmi, SaveDecayCurve=True, SaveNNLSBasis=True, SaveRegParam=True, SaveResidualNorm=True)
    out = D.alloc_t2map_outputs(o, device="cpu")
    assert out["decaycurve"].numel() == V * 32 and out["decaybasis"].numel() == V * 32 * 40
    with pytest.raises(ValueError, match="must be on cuda:0"):
        D.t2map_device(_image(o), o, out)
    with pytest.raises(ValueError, match=r'out\["decaycurve"\]'):
        D.t2map_device(_image(o), o, {**out, "decaycurve": torch.zeros(V, dtype=torch.float64)})
    # SetFlipAngle: one shared nTE x nT2 basis
    o = _opts(dmi, SetFlipAngle=165.0, SaveNNLSBasis=True)
    out = D.alloc_t2map_outputs(o, device="cpu")
    assert out["decaybasis"].numel() == 32 * 40
    with pytest.raises(ValueError, match="must be on cuda:0"):
        D.t2map_device(_image(o), o, out)
    with pytest.raises(ValueError, match=r'out\["decaybasis"\] .* of 1280 elements'):
        D.t2map_device(_image(o), o, {**out, "decaybasis": torch.zeros(V * 32 * 40, dtype=torch.float64)})


def test_t2part_device_refuses_mismatched_tensors(dmi):
    from decaes_amd import device as D

    o = _opts(dmi)
    po = dmi.T2partOptions.from_t2map(o, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0))
    V = 4 * 3 * 2
    parts = {k: torch.zeros(V, dtype=torch.float64) for k in ("sfr", "sgm", "mfr", "mgm")}
    with pytest.raises(ValueError, match=r"dist must be a contiguous float64 tensor of 960 elements"):
        D.t2part_device(torch.zeros(V * 41, dtype=torch.float64), po, parts)
    with pytest.raises(ValueError, match=r'out\["mfr"\]'):
        D.t2part_device(torch.zeros(V * 40, dtype=torch.float64), po, {**parts, "mfr": torch.zeros(V - 2)})
    with pytest.raises(ValueError, match="must be on cuda:0"):
        D.t2part_device(torch.zeros(V * 40, dtype=torch.float64), po, parts)


def test_request_arrays_refuse_mismatched_host_arrays(dmi):
    """The request loop's host arrays: the library reads V x nTE image doubles and V B1-map doubles, so a B1 map that
    numpy would broadcast (a scalar, one slice) is refused rather than passed with its own, shorter buffer."""
    import numpy as np

    from decaes_amd.server import request_arrays

    o = _opts(dmi, SaveRegParam=True)
    img = np.zeros((4, 3, 2, 32))
    with pytest.raises(AssertionError, match="image shape"):
        request_arrays(np.zeros((4, 3, 2, 31)), o)
    for bad in (np.float64(150.0), np.full((4, 3, 1), 150.0), np.full((2, 3, 4), 150.0)):
        with pytest.raises(AssertionError, match="B1 map shape"):
            request_arrays(img, o, bad)
    b1 = np.random.default_rng(0).uniform(130, 180, (4, 3, 2))
    image, alpha, out, dist = request_arrays(img, o, b1)
    assert image.flags.f_contiguous and alpha.flags.f_contiguous
    assert np.array_equal(out["alpha"], b1) and np.isnan(out["mu"]).all() and dist.shape == (4, 3, 2, 40)
