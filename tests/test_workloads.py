"""Seeded workload generators (pyradiance_amd/scenes.py) on the host: the
config-5 office's structure and the annual weather tape, no GPU."""

import numpy as np
import pytest

from pyradiance_amd import _lib, scenes


def _lib_or_skip():
    try:
        _lib.lib()
    except Exception as e:
        pytest.skip(f"libradgpu.so not loadable: {e}")


def _polygons(text: bytes):
    """(modifier, 12 vertex coordinates) of every 4-vertex polygon."""
    toks = text.split()
    out = []
    i = 0
    while i < len(toks):
        if toks[i] == b"polygon":
            mod = toks[i - 1].decode()
            assert toks[i + 2] == b"0" and toks[i + 3] == b"0" and toks[i + 4] == b"12"
            out.append((mod, np.array(toks[i + 5:i + 17], float)))
            i += 17
        else:
            i += 1
    return out


def test_config5_office_structure():
    _lib_or_skip()
    storeys, boxes, nu, nv = 2, 7, 3, 7
    text, sens = scenes.config5(3, storeys=storeys, boxes_per_storey=boxes, nu=nu, nv=nv, sensors_per_storey=(5, 4))
    polys = _polygons(text)
    # per storey: floor, ceiling, 4 x 3 facade bands, 2 x 10 corridor walls
    # and 2 x 9 x 2 partitions, all two-sided walls as two polygons
    per_storey = 2 + 12 + 2 * 20 + 2 * 18
    assert len(polys) == storeys * (per_storey + boxes * 6 * nu * nv)
    glass = [v for m, v in polys if m == "c5_glass"]
    assert len(glass) == storeys * 4
    # every quad planar with a non-zero area; box quads inside the plate
    for m, v in polys:
        p = v.reshape(4, 3)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        assert np.linalg.norm(n) > 1e-6
        assert abs(np.dot(p[3] - p[0], n)) < 1e-6 * np.linalg.norm(n)
        if m.startswith("c5_box"):
            assert (p[:, 0] > 0).all() and (p[:, 0] < 60).all() and (p[:, 1] > 0).all() and (p[:, 1] < 30).all()
    # box faces point outward: the normal leads away from the box centre
    boxq = np.stack([v.reshape(4, 3) for m, v in polys if m.startswith("c5_box")]).reshape(-1, 6 * nu * nv, 4, 3)
    for b in boxq:
        c = b.reshape(-1, 3).mean(0)
        for q in b:
            n = np.cross(q[1] - q[0], q[2] - q[0])
            assert np.dot(n, q.mean(0) - c) > 0
    # sensors: a grid per storey at 0.8 m, facing up
    assert sens.shape == (storeys * 20, 6)
    assert set(np.round(sens[:, 2], 6)) == {0.8, 4.3}
    assert (sens[:, 3:] == [0, 0, 1]).all()


def test_config5_full_size_counts():
    _lib_or_skip()
    text, sens = scenes.config5(0)
    assert text.count(b" polygon ") == 2_495_700
    assert sens.shape == (20_000, 6)


def test_annual_weather_tape_sets_up_every_hour():
    """8,760 hourly records that gendaymtx's host set-up (rg_sky_create) reads
    into 8,760 steps, about half of them daylit."""
    _lib_or_skip()
    from pyradiance_amd import gen
    wea = scenes.annual_wea(0)
    lines = wea.splitlines()
    assert len(lines) == 6 + 8760
    v = np.array([ln.split() for ln in lines[6:]], float)
    assert (v[:, 3:] >= 0).all()
    assert 0.4 < (v[:, 4] > 0).mean() < 0.6
    o, _, _ = gen._parse_args(["-m", "1"])
    sky = _lib.Sky(wea, o)
    try:
        assert (sky.npatch, sky.ncols, sky.ntsteps) == (146, 8760, 8760)
        # the placed suns are the hours with the sun above the horizon (as
        # gendaymtx -D lists them), nearly all with direct irradiance
        ids, good, alt, azi, ill = sky.suns()
        assert 0.4 < len(ids) / 8760 < 0.6
        assert good.mean() > 0.9 and (alt > -0.01).all()
        assert (np.diff(ids) > 0).all()
    finally:
        sky.close()
