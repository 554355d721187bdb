"""gendaymtx on the GPU against the reference gendaymtx: the full float
matrices for every option case, the ASCII output, and the device-resident
matrix feeding the DC x sky product."""

import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

META = json.loads((GOLDEN / "sky_cases.json").read_text())
SKY = np.load(GOLDEN / "sky.npz")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def L():
    from pyradiance_amd import _lib
    _lib.lib()
    return _lib


@pytest.mark.parametrize("name", sorted(META["cases"]))
def test_sky_matrix_matches_reference(L, name):
    from pyradiance_amd import gen
    _say(f"gendaymtx {name}")
    case = META["cases"][name]
    mtx, sky, f, o = gen.sky_matrix(META["tapes"][case["tape"]], case["args"])
    sky.close()
    want = SKY[name].reshape(mtx.shape)
    # FP32 matrix as the reference's; the per-patch Perez terms are evaluated
    # in FP64 on both sides with different math libraries
    assert np.allclose(mtx, want, rtol=2e-5, atol=1e-7 * float(np.abs(want).max()))


def test_ascii_output_matches_reference(L, tmp_path):
    from pyradiance_amd import gendaymtx
    _say("gendaymtx ascii")
    (tmp_path / "wea1.txt").write_text(META["tapes"]["wea1"])
    out = gendaymtx(tmp_path / "wea1.txt").decode()
    head, body = out.split("\n\n", 1)
    rhead, rbody = META["ascii_wea1_m1"].split("\n\n", 1)
    assert [ln for ln in head.splitlines() if "=" in ln] == [ln for ln in rhead.splitlines() if "=" in ln]
    a = np.array(body.split(), dtype=np.float64)
    b = np.array(rbody.split(), dtype=np.float64)
    assert a.shape == b.shape
    # three significant digits: equal, or one unit of the last digit apart
    assert np.allclose(a, b, rtol=6e-3, atol=0)
    assert (body.splitlines() == rbody.splitlines()) or np.mean(np.array(body.splitlines()) != np.array(rbody.splitlines())) < 0.01


def test_sky_matrix_device_feeds_dcsky(L):
    """The sky matrix written straight into device memory is the B operand of
    the DC x sky product: weather tape -> sky -> result without a host copy."""
    import torch
    from pyradiance_amd import gen
    _say("gendaymtx -> dcsky on device")
    mtx, sky, f, o = gen.sky_matrix(META["tapes"]["wea1"], ["-m", "1"])
    K, N = sky.npatch, sky.ncols
    B = torch.empty((K, N, 3), dtype=torch.float32, device="cuda")
    sky.matrix_device(B.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(B.cpu().numpy(), mtx)
    rng = np.random.default_rng(0)
    A = torch.from_numpy(rng.random((64, K, 3), dtype=np.float32)).cuda()
    Cm = torch.empty((64, N, 3), dtype=torch.float32, device="cuda")
    L.dcsky_device(A.data_ptr(), B.data_ptr(), Cm.data_ptr(), 64, K, N)
    want = np.einsum("mkc,knc->mnc", A.cpu().numpy().astype(np.float64), mtx.astype(np.float64))
    assert np.allclose(Cm.cpu().numpy(), want, rtol=1e-4, atol=1e-3)
    sky.close()
