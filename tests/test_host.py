"""Host-side logic: library exports, option parsing, bin specs, scene
generators, matrix I/O, error shapes.  CPU only (no GPU calls)."""

import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from pyradiance_amd import _lib, scenes
from pyradiance_amd import params as P
from pyradiance_amd import rt, util


def _header_symbols():
    text = (ROOT / "include" / "radgpu.h").read_text()
    return sorted(set(re.findall(r"\b(rg_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTS)


def test_library_carries_gfx950_code():
    """The fat binary embedded in our own build names its single target."""
    blob = _lib.LIBPATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_defaults_follow_the_reference():
    rt_ = _lib.default_params(False)     # rt/raycalls.c:118-153
    assert (rt_.ambounce, rt_.ambdiv, rt_.maxdepth) == (0, 1024, -10)
    assert math.isclose(rt_.minweight, 1e-4) and rt_.dstrsrc == 0.0 and math.isclose(rt_.specthresh, 0.15)
    rc = _lib.default_params(True)       # rt/rcontrib.c:24-58
    assert (rc.ambounce, rc.ambdiv) == (1, 350)
    assert math.isclose(rc.minweight, 2e-3) and math.isclose(rc.dstrsrc, 0.9) and math.isclose(rc.specthresh, 0.02)


def test_option_parsing():
    o = P.parse(["-I+", "-ab", "4", "-ad", "4096", "-lw", "1e-4", "-aa", "0", "-c", "2", "-av", "1", "2", "3",
                 "-dj", "0.5", "-u+", "-ffd", "-h"], rcontrib=True)
    p = o.params
    assert o.imm_irrad and o.uncorrelated and not o.header and o.accumulate == 2
    assert (p.ambounce, p.ambdiv) == (4, 4096) and math.isclose(p.minweight, 1e-4)
    assert list(p.ambval) == [1, 2, 3] and p.dstrsrc == 0.5
    assert (o.inform, o.outform) == ("f", "d")
    assert "-aa 0" in o.ignored
    o2 = P.parse(["-I", "-I"])           # bare boolean options toggle
    assert not o2.imm_irrad
    with pytest.raises(ValueError):
        P.parse(["-zz"])
    with pytest.raises(ValueError):
        P.parse(["-ab"])
    with pytest.raises(ValueError):
        P.parse(["-af", "amb.dat"])


def test_modifier_specs():
    m = rt._modifier_spec("skyglow", "reinhartb.cal", None, "Nrbins", "rbin",
                          "MF=4,rNx=0,rNy=0,rNz=-1,Ux=0,Uy=1,Uz=0,RHS=+1")
    assert (m.bin_kind, m.nbins, m.mf, m.rhs) == (1, 2305, 4.0, 1.0)
    assert list(m.nrm) == [0, 0, -1] and list(m.up) == [0, 1, 0]
    k = rt._modifier_spec("win", "klems_full.cal", None, "Nkbins", "kbinS", None)
    assert (k.bin_kind, k.nbins) == (2, 145) and list(k.nrm) == [0, 1, 0]
    k2 = rt._modifier_spec("win", "klems_half.cal", None, None, "khbin(0,0,-1,0,1,0)", None)
    assert (k2.bin_kind, k2.nbins) == (3, 77) and list(k2.up) == [0, 1, 0]
    c = rt._modifier_spec("groundglow", None, None, "1", "0", None)
    assert (c.bin_kind, c.nbins, c.cbin) == (0, 1, 0)
    with pytest.raises(ValueError):
        rt._modifier_spec("x", "tregenza.cal", None, None, "tbin", None)
    with pytest.raises(ValueError):
        rt.Rcontrib(b"", "x.oct").add_modifier()


def test_rtrace_type_error_and_error_shape(tmp_path):
    with pytest.raises(TypeError):
        rt.rtrace("0 0 0 0 0 1", "x.oct")
    with pytest.raises(RuntimeError, match=r"^An error occurred with exit code 1: "):
        rt.rtrace(b"0 0 0 0 0 1\n", tmp_path / "missing.oct")


def test_scene_generators_are_seeded():
    a = scenes.config2(3, nboxes=10)
    b = scenes.config2(3, nboxes=10)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
    rad, sens = scenes.config2(0)
    assert len(sens) == 2500
    assert rad.count(" polygon box") == 200 * 6 * 21
    r1, s1, sun = scenes.config1(0)
    assert s1.shape == (77, 6) and abs(np.linalg.norm(sun) - 1) < 1e-12
    alt = math.degrees(math.asin(sun[2]))
    assert 20 <= alt <= 60 and sun[1] < 0


def test_box_faces_point_outward():
    txt = scenes.box_quads_rad("m", "b", (0, 0, 0), (1, 2, 3), 1, 1)
    for block in txt.split("m polygon")[1:]:
        v = np.array(block.split()[4:], float).reshape(-1, 3)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        c = v.mean(0) - np.array([0.5, 1.0, 1.5])
        assert np.dot(n, c) > 0


def test_sky_matrix_night_hours_zero():
    s = scenes.sky_matrix(0, 146, 48)
    assert s.shape == (146, 48, 3) and s.dtype == np.float32
    night = (s == 0).all(axis=(0, 2))
    assert night[:4].all() and not night[11:14].any()


@pytest.mark.parametrize("form", ["a", "f", "d"])
def test_matrix_roundtrip(form):
    m = np.random.default_rng(0).random((5, 7, 3)).astype(np.float32)
    b = util.write_matrix(m, form)
    back = util.read_matrix(b)
    tol = 1e-6 if form == "a" else 0
    assert back.shape == m.shape and np.allclose(back, m, rtol=tol, atol=0)
    assert b"NROWS=5" in b and b"NCOLS=7" in b


def test_own_sky_cal_matches_reinhartb(refbin, tmp_path):
    """oracle/reinhart_sky.cal (used for the CPU baseline) bins like reinhartb.cal."""
    g = np.load(ROOT / "tests" / "golden" / "bins_reinhart.npz")
    sky = tmp_path / "sky.rad"
    sky.write_text("void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 180\n")
    octf = tmp_path / "sky.oct"
    with open(octf, "wb") as f:
        subprocess.run([str(refbin / "oconv"), str(sky)], stdout=f, check=True)
    d = g["dirs"][:500]
    inp = "\n".join(f"0 0 0 {x:.17g} {y:.17g} {z:.17g}" for x, y, z in d).encode()
    env = dict(os.environ, RAYPATH=f".:{refbin / 'lib'}")
    r = subprocess.run([str(refbin / "rcontrib"), "-h", "-ab", "0", "-f", str(ROOT / "oracle" / "reinhart_sky.cal"),
                        "-p", "MF=1", "-bn", "Nskybins", "-b", "skybin", "-m", "skyglow", str(octf)],
                       input=inp, capture_output=True, env=env, check=True)
    v = np.array(r.stdout.split(), float).reshape(len(d), -1)[:, 0::3]
    assert np.array_equal(np.where(v.max(1) > 0.5, v.argmax(1), -1), g["mf1"][:500])
