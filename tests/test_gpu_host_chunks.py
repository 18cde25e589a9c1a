"""dmi_t2map_f64 on volumes of several chunks (host arrays staged chunk by chunk, the copies of one chunk overlapped
with the voxel kernel of another) against dmi_t2map_f64_device on the same image: every output bit for bit. The
cases cover a B1 map, SetFlipAngle's shared basis, and an option set whose launch schedule is measured on the
first chunk of the host call."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 524288  # voxels per chunk of the host-array path

CASES = {
    # shape, nTE, TE, nT2, T2Range, Reg, nRefAngles, extra options, mask fraction, B1 map
    "chi2_b1": ((128, 128, 72), 32, 10e-3, 40, (10e-3, 2.0), "chi2", 24, dict(Chi2Factor=1.02), 0.5, True),
    "gcv_setflip_basis": ((128, 128, 70), 32, 10e-3, 40, (10e-3, 2.0), "gcv", 32,
                          dict(SetFlipAngle=160.0, SaveNNLSBasis=True), 0.5, False),
    # 56 echoes, nT2 60: split schedules are candidates; 40 angles: no schedule measured for these sizes yet
    "lcurve_tuned": ((120, 120, 80), 56, 7.9e-3, 60, (8e-3, 2.0), "lcurve", 40, {}, 0.5, False),
}


def _host(a):
    return a.cpu().numpy().transpose(tuple(reversed(range(a.dim()))))


@pytest.mark.parametrize("name", list(CASES))
def test_host_chunks_equal_device_path(dmi, name):
    import torch

    from decaes_amd import device as D

    shape, m, TE, n, T2R, reg, K, extra, frac, b1 = CASES[name]
    assert np.prod(shape) > 2 * CHUNK
    print(f"[host-chunks] case={name} shape={shape}", flush=True)
    img, nin = dmi.synth_mse(shape, m, TE, mask_fraction=frac, seed=41)
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K, SaveRegParam=True,
                                       SaveResidualNorm=True, SaveDecayCurve=True, Silent=True, **extra)
    alpha = np.asfortranarray(np.random.default_rng(3).uniform(130, 180, shape)) if b1 else None
    # the host-array call first, so that for lcurve_tuned its first chunk runs the schedule measurement
    maps, dist = dmi.T2mapSEcorr(img, opts, alpha=alpha)
    st_host = dmi.context(0).stats()
    assert st_host["voxels"] == nin

    dimg = torch.from_numpy(np.ascontiguousarray(img.transpose(3, 2, 1, 0))).to("cuda")
    out = D.alloc_t2map_outputs(opts, device="cuda")
    dalpha = None
    if b1:
        dalpha = torch.from_numpy(np.ascontiguousarray(alpha.transpose(2, 1, 0))).to("cuda")
        out["alpha"].copy_(dalpha)
    st_dev = D.t2map_device(dimg, opts, out, alpha=dalpha)
    torch.cuda.synchronize()
    print(f"[host-chunks] host schedule {st_host['lds_rows']} rows x {st_host['workgroups_per_cu']} per CU, "
          f"device {st_dev['lds_rows']} x {st_dev['workgroups_per_cu']}", flush=True)
    assert st_dev["voxels"] == nin
    assert st_host["nnls_solves"] == st_dev["nnls_solves"] and st_host["fp64_flops"] == st_dev["fp64_flops"]
    assert (st_host["lds_rows"], st_host["workgroups_per_cu"]) == (st_dev["lds_rows"], st_dev["workgroups_per_cu"])
    assert np.array_equal(_host(out["dist"]), dist, equal_nan=True)
    for k, v in out.items():
        if k == "dist":
            continue
        assert np.array_equal(_host(v), maps[k], equal_nan=True), k
    if b1:
        assert np.array_equal(maps["alpha"], alpha)
