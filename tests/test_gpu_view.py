"""rpict on the GPU against the reference rpict: pictures of the mesh scene
(smooth plastic, mirror metal, pure trans, glass, a lamp and the floor) for
a perspective and an angular fisheye view, every pixel sampled at its centre
(-ps 1 -pj 0) with deterministic shading."""

import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

META = json.loads((GOLDEN / "image.json").read_text())
SHADING = ["-ab", "0", "-dj", "0", "-ds", "0", "-dt", "0", "-dc", "1", "-dr", "0", "-st", "1", "-lr", "8",
           "-lw", "1e-4", "-av", "0", "0", "0"]


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", ["persp", "fisheye"])
def test_rpict_matches_reference(name):
    from pyradiance_amd import rpict
    from pyradiance_amd import view as V
    _say(f"rpict {name}")
    out = rpict(META["rpict"][name], GOLDEN / "mesh" / "scene.rad", xres=48, yres=48,
                params=["-ps", "1", "-pj", "0", *SHADING])
    img, head = V.read_picture(out)
    ref, rhead = V.read_picture((GOLDEN / "image" / f"{name}.hdr").read_bytes())
    assert img.shape == ref.shape
    vw = [h for h in head if h.startswith("VIEW=")]
    assert vw and vw[0] == [h for h in rhead if h.startswith("VIEW=")][0]
    # both quantised to RGBE (8-bit mantissa): equal, or one mantissa step apart
    step = np.maximum(ref, img).max(axis=2, keepdims=True) / 128.0
    assert (np.abs(img - ref) <= step + 1e-12).all()
    assert (img == ref).mean() > 0.98


def test_rpict_jitter_and_resolution():
    """Default jitter (-pj .67) moves each ray within its pixel only: the
    picture keeps its resolution and agrees with the centred one on average."""
    from pyradiance_amd import rpict
    from pyradiance_amd import view as V
    _say("rpict jitter")
    a, _ = V.read_picture(rpict(META["rpict"]["persp"], GOLDEN / "mesh" / "scene.rad", xres=96, yres=96,
                                params=SHADING))
    b, _ = V.read_picture(rpict(META["rpict"]["persp"], GOLDEN / "mesh" / "scene.rad", xres=96, yres=96,
                                params=["-pj", "0", *SHADING]))
    assert a.shape == b.shape and a.shape[1] == 96
    assert abs(a.mean() - b.mean()) < 0.05 * b.mean()
