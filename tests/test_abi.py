"""CPU checks of the C ABI library and the host-side mirror (no compute calls need a GPU here)."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "decaes_amd.h").read_text()
    return sorted(set(re.findall(r"\b(dmi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(dmi):
    lib = dmi.load_library()
    names = _declared_symbols()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(dmi.EXPORTS)


def test_no_cpu_fallback_in_product():
    # the product package never imports the oracle or carries a host compute path
    pkg = ROOT / "decaes.jl_amd"
    for p in list(pkg.rglob("*.py")) + list(pkg.rglob("*.cpp")) + list(pkg.rglob("*.hip")):
        t = p.read_text()
        assert "oracle" not in t.lower() or "never imports the oracle" in t, p


def test_synth_deterministic_across_threads(dmi):
    a, na = dmi.synth_mse((12, 10, 3), 32, 10e-3, mask_fraction=0.6, seed=42, nthreads=1)
    b, nb = dmi.synth_mse((12, 10, 3), 32, 10e-3, mask_fraction=0.6, seed=42, nthreads=7)
    assert na == nb and np.array_equal(a, b)
    c, _ = dmi.synth_mse((12, 10, 3), 32, 10e-3, mask_fraction=0.6, seed=43, nthreads=7)
    assert not np.array_equal(a, c)
    assert abs(na / (12 * 10 * 3) - 0.6) <= 0.03
    inside = a[..., 0] > 0
    assert np.all(a[~inside] == 0) and np.all(a[inside] > 0)


def test_synth_signal_model(dmi, oracle):
    # noiseless limit of the generator: huge SNR is not exposed, so check the envelope: every in-mask voxel is a
    # positive decaying train whose first echo is within noise of a convex combination of EPG curves (<= 1)
    img, n = dmi.synth_mse((8, 8, 2), 48, 10e-3, mask_fraction=1.0, seed=3)
    assert n == 128
    assert np.all(img[..., 0] < 1.1) and np.all(img[..., -1] < img[..., 0])


def test_options_mirror_reference_asserts(dmi):
    base = dict(MatrixSize=(2, 2, 2), nTE=32, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="chi2", Chi2Factor=1.02)
    o = dmi.T2mapOptions(**base)
    assert o.nRefAngles == 64 and o.nRefAnglesMin == 5 and o.T1 == 1.0 and o.Threshold == 0.0
    bad = [dict(nTE=3), dict(TE=0.0), dict(nT2=1), dict(T2Range=(2.0, 1.0)), dict(Threshold=-1.0),
           dict(MinRefAngle=190.0), dict(nRefAngles=1), dict(nRefAnglesMin=65), dict(Reg="foo"),
           dict(Chi2Factor=1.0), dict(Reg="mdp"), dict(RefConAngle=-1.0), dict(SetFlipAngle=181.0)]
    for kw in bad:
        with pytest.raises(AssertionError):
            dmi.T2mapOptions(**{**base, **kw})
    with pytest.raises(AssertionError, match="At least four echoes"):
        dmi.T2mapOptions(**{**base, "nTE": 2})
    assert dmi.T2mapOptions(**{**base, "Threshold": -math.inf}).Threshold == -math.inf
    o = dmi.T2mapOptions.from_image(np.zeros((3, 4, 5, 24)), TE=7e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="gcv")
    assert o.MatrixSize == (3, 4, 5) and o.nTE == 24
    p = dmi.T2partOptions.from_t2map(o, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0))
    assert p.nT2 == 40 and p.MatrixSize == (3, 4, 5)


def test_grids_match_oracle(dmi, oracle):
    o = dmi.T2mapOptions(MatrixSize=(1, 1, 1), nTE=48, TE=10e-3, nT2=60, T2Range=(8e-3, 2.0), Reg="gcv", nRefAngles=64)
    assert np.array_equal(o.T2_component_times(), oracle.logrange(8e-3, 2.0, 60))
    assert np.array_equal(o.flip_angles(), oracle.linrange(50.0, 180.0, 64))


def test_context_without_device_fails_loudly(dmi):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(dmi.DmiError) as e:
        dmi.Context(0)
    assert e.value.code == -5


def test_server_config_layout_and_order_check():
    """dmi_server_config as declared: int64 budget, then four int32 (streams, chunk, order, reserved)."""
    import ctypes as C

    from decaes_amd import _lib
    from decaes_amd.server import Server

    assert C.sizeof(_lib.ServerConfigC) == 24
    assert [f[0] for f in _lib.ServerConfigC._fields_] == ["hbm_budget_bytes", "max_streams", "chunk_voxels", "order",
                                                           "reserved"]
    with pytest.raises(ValueError, match="order"):
        Server(order="fifo")
    # the C entry point rejects an unknown order before it looks for a device; a valid one gets that far
    L = _lib.load()
    h = C.c_void_p()
    assert L.dmi_server_create(C.byref(h), 0, C.byref(_lib.ServerConfigC(0, 0, 0, 7, 0))) == _lib.DMI_EINVAL
    assert not h.value
    for order in (0, 1):
        rc = L.dmi_server_create(C.byref(h), 0, C.byref(_lib.ServerConfigC(0, 0, 0, order, 0)))
        assert rc != _lib.DMI_EINVAL
        if rc == _lib.DMI_OK:
            assert L.dmi_server_destroy(h) == _lib.DMI_OK
