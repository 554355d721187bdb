"""Meshes on the GPU: compiled meshes from the reference obj2mesh, placed with
transforms (rotation, scale, mirror), traced by the HIP kernels and compared
with the reference rtrace: first hits (distance, unperturbed normal) and
deterministic radiance through smooth-shaded plastic, mirror metal, pure
trans and flat glass triangles."""

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    from pyradiance_amd import _lib
    _lib.lib()
    c = _lib.Context(0)
    c.load(str(GOLDEN / "mesh" / "scene.rad")).commit()
    p = _lib.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat, p.specthresh, p.maxdepth, p.minweight = 0, 0.0, 0.0, 1.0, 8, 1e-4
    p.ambval[:] = (0.0, 0.0, 0.0)
    c.set_params(p)
    yield c
    c.close()


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
def test_mesh_first_hits_match_reference(ctx, builder):
    from pyradiance_amd import _lib
    _say(f"mesh first hits {builder}")
    g = np.load(GOLDEN / "mesh_rtrace.npz")
    ctx.set_bvh_builder(_lib.BVH_SAH if builder == "sah" else _lib.BVH_LBVH).commit()
    t, nrm, prim = ctx.first_hit(g["rays"])
    hit_ref = g["dist"] < 1e9
    assert np.array_equal(prim >= 0, hit_ref)
    m = hit_ref
    assert np.allclose(t[m], g["dist"][m], rtol=1e-6, atol=1e-6)
    assert np.allclose(nrm[m], g["normal"][m], atol=1e-6)


def test_mesh_radiance_matches_reference(ctx):
    _say("mesh radiance")
    g = np.load(GOLDEN / "mesh_rtrace.npz")
    v = ctx.trace(g["rays"], 0)
    # colour coefficients are FP32 on both sides (the reference's SCOLOR)
    assert np.allclose(v, g["value"], rtol=1e-5, atol=1e-6)


def test_mesh_through_python_api(tmp_path):
    """rtrace() on the scene file: mesh paths resolve next to the scene."""
    from pyradiance_amd import rtrace
    _say("mesh rtrace api")
    g = np.load(GOLDEN / "mesh_rtrace.npz")
    rays = "".join(" ".join(f"{x:.17g}" for x in r) + "\n" for r in g["rays"][:200]).encode()
    out = rtrace(rays, GOLDEN / "mesh" / "scene.rad", header=False,
                 params=["-ab", "0", "-dj", "0", "-ds", "0", "-dt", "0", "-st", "1", "-lr", "8", "-lw", "1e-4",
                         "-av", "0", "0", "0"])
    v = np.array(out.split(), dtype=np.float64).reshape(-1, 3)
    assert np.allclose(v, g["value"][:200], rtol=1e-5, atol=1e-6)


def test_instance_scene_matches_reference():
    """Instanced octrees (frozen with every surface type, file-list, nested;
    rotated, scaled, mirrored, glass override): first hits and radiance
    against the reference rtrace."""
    from pyradiance_amd import _lib
    _say("instances")
    g = np.load(GOLDEN / "inst_rtrace.npz")
    c = _lib.Context(0)
    c.load(str(GOLDEN / "inst" / "scene.rad")).commit()
    p = _lib.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat, p.specthresh, p.maxdepth, p.minweight = 0, 0.0, 0.0, 1.0, 8, 1e-4
    p.ambval[:] = (0.0, 0.0, 0.0)
    c.set_params(p)
    t, nrm, prim = c.first_hit(g["rays"])
    hit_ref = g["dist"] < 1e9
    assert np.array_equal(prim >= 0, hit_ref)
    assert np.allclose(t[hit_ref], g["dist"][hit_ref], rtol=1e-6, atol=1e-6)
    assert np.allclose(nrm[hit_ref], g["normal"][hit_ref], atol=1e-6)
    v = c.trace(g["rays"], 0)
    assert np.allclose(v, g["value"], rtol=1e-5, atol=1e-6)
    c.close()
