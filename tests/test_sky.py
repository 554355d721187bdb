"""gendaymtx on the CPU side: weather reading, solar position and the
per-step Perez set-up (host code of libradgpu.so, no GPU needed), pinned by
the reference gendaymtx's sun descriptions (-D), which carry every placed
sun's position and direct illuminance, and by its matrix shapes."""

import json

import numpy as np
import pytest

from conftest import GOLDEN
from pyradiance_amd import _lib, gen

META = json.loads((GOLDEN / "sky_cases.json").read_text())
SKY = np.load(GOLDEN / "sky.npz")


def _sky(tape, args):
    o, f, _ = gen._parse_args(list(args))
    return _lib.Sky(META["tapes"][tape], o), f, o


def test_sun_descriptions_match_reference():
    sky, f, o = _sky("wea1", ["-D", "-"])
    assert not f["matrix"]
    text, _ = gen._sun_text(*sky.suns(), list(o.sun_color))
    assert text == META["suns_wea1"]


def test_sun_file_and_modifiers_match_reference_epw():
    sky, f, o = _sky("epw", ["-m", "1", "-D", "suns.rad", "-M", "mods.txt"])
    assert f["matrix"] and f["sunsfile"] == "suns.rad"
    text, mods = gen._sun_text(*sky.suns(), list(o.sun_color))
    assert text == META["suns_epw_file"]
    assert mods == META["mods_epw_file"]


@pytest.mark.parametrize("name", sorted(META["cases"]))
def test_matrix_shapes_match_reference(name):
    case = META["cases"][name]
    sky, f, o = _sky(case["tape"], case["args"])
    assert sky.npatch * sky.ncols * 3 == SKY[name].size
    m = int(case["args"][case["args"].index("-m") + 1])
    assert sky.npatch == 144 * m * m + 2          # ground, Reinhart patches, zenith


def test_header_fields_match_reference():
    head = META["ascii_wea1_m1"].split("\n\n", 1)[0].splitlines()
    sky, f, o = _sky("wea1", ["-m", "1"])
    lat, lon = sky.site()
    want = {ln.split("=")[0]: ln for ln in head if "=" in ln}
    assert want["LATLONG"] == "LATLONG= %.8f %.8f" % (lat, lon)
    assert want["NROWS"] == f"NROWS={sky.npatch}" and want["NCOLS"] == f"NCOLS={sky.ncols}"


def test_weather_errors():
    o, _, _ = gen._parse_args(["-m", "1"])
    with pytest.raises(RuntimeError, match="header format error"):
        _lib.Sky("place X\nlatitude 95\nlongitude 0\ntime_zone 0\nsite_elevation 0\nweather_data_file_units 1\n", o)
    with pytest.raises(RuntimeError, match="Illegal date"):
        _lib.Sky(META["tapes"]["wea1"].replace("\n1 15 0.500", "\n13 15 0.500", 1), o)
    with pytest.raises(RuntimeError, match="Usage"):
        gen._parse_args(["-q"])
