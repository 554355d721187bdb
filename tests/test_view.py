"""The image path on the CPU: view options, view set-up and viewray() of
every view type against the reference vwrays, aspect handling, and the
Radiance picture codec on the reference rpict's own pictures."""

import json

import numpy as np
import pytest

from conftest import GOLDEN
from pyradiance_amd import view as V

META = json.loads((GOLDEN / "image.json").read_text())
VW = np.load(GOLDEN / "vwrays.npz")


@pytest.mark.parametrize("name", sorted(META["vwrays"]))
def test_vwrays_match_reference(name):
    out = V.vwrays(view=META["vwrays"][name], xres=16, yres=12, outform="d")
    got = np.frombuffer(out, dtype="<f8").reshape(-1, 6)
    assert got.shape == VW[name].shape
    assert np.allclose(got, VW[name], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", sorted(META["vwrays"]))
def test_dimensions_match_reference(name):
    assert V.vwrays(view=META["vwrays"][name], xres=16, yres=12, dimensions=True).decode() == META["dims"][name]


def test_pixel_aspect_zero_keeps_resolution():
    out = V.vwrays(view=META["vwrays"]["per_offset"], xres=16, yres=12, pixel_aspect=0, outform="d")
    got = np.frombuffer(out, dtype="<f8").reshape(-1, 6)
    assert np.allclose(got, VW["per_offset_pa0"], rtol=1e-12, atol=1e-12)


def test_view_from_picture_header():
    """The view of the reference rpict picture, read back from its header,
    sets up to the same vectors as the options it was rendered with."""
    data = (GOLDEN / "image" / "persp.hdr").read_bytes()
    path = GOLDEN / "image" / "persp.hdr"
    a = V.View()
    assert V.view_file(path, a) > 0
    b, rest = V.parse_view(META["rpict"]["persp"])
    assert not rest
    a.setup()
    b.setup()
    for k in ("vp", "vdir", "hvec", "vvec"):
        assert np.allclose(getattr(a, k), getattr(b, k), atol=1e-6)
    assert b"VIEW=" in data


@pytest.mark.parametrize("name", ["persp", "fisheye"])
def test_picture_codec_round_trip(name):
    """Reference pictures decode (run-length scanlines), re-encode to the
    same RGBE bytes (setcolr quantisation) and decode again unchanged."""
    img, head = V.read_picture((GOLDEN / "image" / f"{name}.hdr").read_bytes())
    v, _ = V.parse_view(META["rpict"][name])
    v.setup()
    _, x, y = V.normaspect(v.aspect(), 1.0, 48, 48)
    assert img.shape == (y, x, 3) and any(h.startswith("VIEW=") for h in head)
    again, _ = V.read_picture(V.write_picture(img, ["VIEW= -vtv"]))
    assert np.array_equal(again, img)


def test_view_errors():
    with pytest.raises(ValueError, match="parallel"):
        V.parse_view(["-vtv", "-vd", "0", "0", "1"])[0].setup()
    with pytest.raises(ValueError, match="horizontal"):
        V.parse_view(["-vtv", "-vh", "180"])[0].setup()
