"""GPU parity: the HIP path (through the C ABI) against the oracle, the
reference's golden vectors and size-independent properties."""

import math

import numpy as np
import pytest

from conftest import GOLDEN
from pyradiance_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def L():
    from pyradiance_amd import _lib
    _lib.lib()
    return _lib


def _ctx(L, rad, builder=None):
    c = L.Context(0)
    c.load_text(rad)
    if builder is not None:
        c.set_bvh_builder(builder)
    c.commit()
    return c


def _rc_params(L, **kw):
    p = L.default_params(True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _sky_mods(L, mf=1):
    m = L.RgModifier()
    m.name, m.bin_kind, m.nbins, m.mf, m.rhs = b"skyglow", 1, scenes.reinhart_nbins(mf), float(mf), 1.0
    m.nrm[:] = (0, 0, -1)
    m.up[:] = (0, 1, 0)
    g = L.RgModifier()
    g.name, g.bin_kind, g.nbins, g.cbin = b"groundglow", 0, 1, 0
    return [m, g]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_config1_irradiance_matches_reference(L, oracle, seed):
    _say(f"config1 seed {seed}")
    g = np.load(GOLDEN / f"c1_seed{seed}.npz")
    rad, sens, _ = scenes.config1(seed)
    c = _ctx(L, rad)
    p = L.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat = 0, 0.0, 0.0
    c.set_params(p)
    out = c.trace(sens, L.RT_IMM_IRRAD)
    ref = g["values"]
    assert np.abs(out - ref).max() <= 1e-4 * ref.max()
    orc, nr = oracle.Scene(rad).trace(sens, oracle.params(ambounce=0, dstrsrc=0.0, srcsizerat=0.0), 8)
    assert np.abs(out - orc).max() <= 1e-6 * orc.max()
    assert c.last_nrays() == nr          # ray counts: bit-exact
    assert c.last_nrays() == int(g["nrays"])   # the reference's own counter (rt/raytrace.c:606)


def _mixed_scene():
    return ("void plastic m 0 0 5 .5 .5 .5 0 0\n"
            "void glass g 0 0 3 .8 .8 .8\n"
            "m sphere s1 0 0 4 0 0 0 1\n"
            "m bubble s2 0 0 4 3 0 0 0.7\n"
            "m polygon floor 0 0 12 -6 -6 -2 6 -6 -2 6 6 -2 -6 6 -2\n"
            "m polygon tri 0 0 9 -3 -3 0 -1 -3 0 -2 -1 1\n"
            "g polygon pane 0 0 12 -4 2 -1 -2 2 -1 -2 2 1 -4 2 1\n"
            "m ring r1 0 0 8 0 3 0 0 1 0 0.2 1.2\n"
            "m cylinder c1 0 0 7 -3 0 -2 -3 0 2 0.5\n"
            "m tube t1 0 0 7 3 3 -2 3 3 2 0.6\n"
            "m cone k1 0 0 8 0 -3 -2 0 -3 1 1.0 0.2\n"
            "m cup k2 0 0 8 2 -3 -2 2 -3 0 0.1 0.8\n")


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
def test_closest_hit_matches_brute_force(L, oracle, builder):
    _say(f"closest hit, mixed primitives, {builder} tree")
    kind = {"sah": L.BVH_SAH, "lbvh": L.BVH_LBVH}[builder]
    rad = _mixed_scene()
    rng = np.random.default_rng(7)
    n = 4000
    org = rng.uniform(-5, 5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([org, d])
    prim, t = oracle.Scene(rad).closest(rays)
    # GPU: primary rays into an all-black scene carrying one tracked modifier
    # per primitive would need bins; instead compare through rtrace -i, whose
    # value is non-zero exactly when a primitive is hit under a bright sky.
    sky = "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 360\n"
    c = _ctx(L, rad + sky, kind)
    p = L.default_params(False)
    c.set_params(p)
    out = c.trace(rays, 0)
    o2, _ = oracle.Scene(rad + sky).trace(rays, oracle.params(), 0)
    same = np.isclose(out, o2, rtol=1e-6, atol=1e-9).all(1)
    assert same.mean() >= 0.999
    # rays that escape see the sky (value 1), hits on plastic with -ab 0 are black
    escaped = prim < 0
    assert np.allclose(out[escaped], 1.0)


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
def test_first_hit_matches_reference(L, builder):
    """rg_first_hit against rtrace -oLN / -os of the reference: distance,
    unflipped normal and surface for sphere, bubble, polygons, rings,
    cylinders, tubes, cones and cups."""
    _say(f"first hit vs reference, {builder} tree")
    g = np.load(GOLDEN / "prims.npz")
    c = _ctx(L, str(g["scene"]), {"sah": L.BVH_SAH, "lbvh": L.BVH_LBVH}[builder])
    t, nrm, prim = c.first_hit(g["rays"])
    hit = g["t"] < 1e9
    assert np.array_equal(prim >= 0, hit)
    assert np.allclose(t[hit], g["t"][hit], rtol=1e-5, atol=0)
    assert np.all(t[~hit] == 1e10)
    assert np.allclose(nrm, g["normal"], rtol=0, atol=2e-6)
    names = np.array([ln.split()[2] for ln in str(g["scene"]).splitlines() if ln.split()[1] not in
                      ("plastic", "glass")])
    assert np.array_equal(np.where(prim >= 0, names[np.maximum(prim, 0)], "*"), g["name"])
    c.close()


def test_local_sources_metal_trans_glass_match_reference(L):
    """Deterministic shading against the reference rtrace: polygon, sphere
    and ring light sources, a proximity glow, metal, trans, glass and
    mirror-smooth metal (-ab 0 -dj 0 -ds 0 -dt 0 -st 1), radiance and -I."""
    _say("shading vs reference")
    g = np.load(GOLDEN / "shading.npz")
    c = _ctx(L, str(g["scene"]))
    p = L.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat, p.specthresh, p.maxdepth, p.minweight = 0, 0.0, 0.0, 1.0, 8, 1e-4
    p.ambval[:] = (0.0, 0.0, 0.0)
    c.set_params(p)
    v = c.trace(g["rays"], 0)
    assert np.allclose(v, g["values"], rtol=1e-5, atol=1e-9)
    e = c.trace(g["sensors"], L.RT_IMM_IRRAD)
    assert np.allclose(e, g["irradiance"], rtol=1e-5, atol=1e-9)
    c.close()


def test_rcontrib_matches_oracle_same_samples(L, oracle):
    _say("rcontrib -ab 2 vs oracle")
    rng = np.random.default_rng(1)
    room = scenes.Room()
    rad = scenes.room_rad(room) + scenes.furniture_rad(rng, 5, room, 1, 1) + scenes.sky_rad((1, 1, 1), (1, 1, 1))
    pts = np.array([[1.0, 1.0, 0.8, 0, 0, 1], [3.0, 2.0, 0.8, 0, 0, 1], [5.0, 3.0, 1.5, 0, 0, 1],
                    [3.0, 0.5, 1.5, 0, -1, 0]])
    rays = np.repeat(pts, 8, axis=0)
    mods = _sky_mods(L)
    c = _ctx(L, rad)
    c.set_params(_rc_params(L, ambounce=2, ambdiv=256, minweight=1e-3, seed=3))
    out = c.contrib(rays, mods, 1, L.RT_IMM_IRRAD)
    omods = [oracle.mod("skyglow", 1, 145), oracle.mod("groundglow", 0, 1)]
    orc, nr = oracle.Scene(rad).contrib(rays, oracle.params(ambounce=2, ambdiv=256, minweight=1e-3, dstrsrc=0.9,
                                                            specthresh=0.02, seed=3), omods, 1, 8)
    # identical samples: totals agree tightly, entries agree except where a
    # sample's path diverged on an FP32-vs-FP64 edge decision
    assert np.allclose(out.sum(1), orc.sum(1), rtol=2e-3)
    close = np.isclose(out, orc, rtol=1e-4, atol=1e-7)
    assert close.mean() > 0.99
    assert abs(c.last_nrays() - nr) <= 0.002 * nr


def test_rtrace_ambient_matches_oracle(L, oracle):
    _say("rtrace -I -ab 1 with sun and sky vs oracle")
    rad, sens, _ = scenes.config1(0)
    c = _ctx(L, rad)
    p = L.default_params(False)
    p.ambounce, p.ambdiv, p.dstrsrc = 1, 128, 0.5
    p.seed = 9
    c.set_params(p)
    out = c.trace(sens, L.RT_IMM_IRRAD)
    orc, nr = oracle.Scene(rad).trace(sens, oracle.params(ambounce=1, ambdiv=128, dstrsrc=0.5, seed=9), 8)
    assert np.allclose(out.sum(1), orc.sum(1), rtol=1e-3, atol=1e-6)


def test_uniform_sky_pi(L):
    _say("uniform sky = pi")
    c = _ctx(L, scenes.sky_rad((1, 1, 1), ground=(1, 1, 1)))
    c.set_params(_rc_params(L, ambounce=1, ambdiv=4096, minweight=1e-4))
    rays = np.array([[0, 0, 1.0, 0, 0, 1], [5, 5, 0, 0, 0, 1], [0, 0, 0, 1, 0, 0]], float)
    out = c.contrib(rays, _sky_mods(L), 1, L.RT_IMM_IRRAD)
    assert np.allclose(out[:, 0::3].sum(1), math.pi, rtol=1e-6)
    assert np.allclose(out[2, -3], math.pi / 2, rtol=0.02)    # half the hemisphere is ground


@pytest.mark.parametrize("mf", [1, 2, 4])
def test_bins_match_reference(L, mf):
    _say(f"Reinhart MF {mf} bins")
    g = np.load(GOLDEN / "bins_reinhart.npz")
    d = g["dirs"]
    c = _ctx(L, "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 180\n")
    c.set_params(_rc_params(L, ambounce=0))
    out = c.contrib(np.hstack([np.zeros_like(d), d]), _sky_mods(L, mf)[:1], 1, 0)
    v = out[:, 0::3]
    got = np.where(v.max(1) > 0.5, v.argmax(1), -1)
    assert np.array_equal(got, g[f"mf{mf}"])


def test_klems_bins_match_reference(L):
    _say("Klems bins")
    g = np.load(GOLDEN / "bins_klems.npz")
    d = g["dirs"]
    c = _ctx(L, "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 360\n")
    c.set_params(_rc_params(L, ambounce=0))
    for name, nrm, up in (("kbinS", (0, 1, 0), (0, 0, 1)), ("kbinD", (0, 0, -1), (0, 1, 0))):
        m = L.RgModifier()
        m.name, m.bin_kind, m.nbins, m.rhs = b"skyglow", 2, 145, 1.0
        m.nrm[:], m.up[:] = nrm, up
        out = c.contrib(np.hstack([np.zeros_like(d), d]), [m], 1, 0)[:, 0::3]
        got = np.where(out.max(1) > 0.5, out.argmax(1), -1)
        assert np.array_equal(got, g[name]), name


def test_sah_tree_independent_of_host_threads(L, monkeypatch):
    """The host SAH build splits the top of the tree on one thread and hands
    the subtrees below to the host threads; every range is split and
    numbered as the one-thread build does it, so first hits (distance,
    normal, primitive) agree bit for bit whatever the thread count.  150k
    random triangles, above the size the build starts threads at."""
    _say("sah threads")
    rng = np.random.default_rng(5)
    n = 150_000
    c0 = rng.random((n, 3)) * 20.0
    parts = ["void plastic m 0 0 5 .5 .5 .5 0 0"]
    for i in range(n):
        v = c0[i] + rng.normal(0, 0.15, (3, 3))
        parts.append(f"m polygon t{i} 0 0 9 " + " ".join(f"{x:.9g}" for x in v.ravel()))
    rad = "\n".join(parts) + "\n"
    rays = np.hstack([rng.random((20_000, 3)) * 20.0, rng.normal(0, 1, (20_000, 3))])
    out = []
    for threads in ("1", "16"):
        monkeypatch.setenv("RG_BVH_THREADS", threads)
        c = _ctx(L, rad, L.BVH_SAH)
        out.append(c.first_hit(rays))
        c.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    assert (out[0][2] >= 0).mean() > 0.3


def test_dcsky_against_fp64(L):
    _say("dcsky")
    rng = np.random.default_rng(0)
    for (M, K, N) in [(37, 19, 51), (250, 146, 877), (64, 2305, 96), (300, 77, 130), (129, 33, 257),
                      (513, 160, 1203), (130, 64, 2048)]:
        A = rng.random((M, K, 3), dtype=np.float32)
        B = rng.lognormal(0, 1, (K, N, 3)).astype(np.float32)
        ref = np.einsum("mkc,knc->mnc", A.astype(np.float64), B.astype(np.float64), optimize=True)
        c32 = L.dcsky(A, B, 0)
        e32 = np.abs(c32 - ref).max() / np.abs(ref).max()
        assert e32 < 1e-5, (M, K, N, e32)
        # automatic, direct, LDS-staged and (K <= 160) short-K split-BF16 products
        for mode in ((1, 2, 3, 4) if K <= 160 else (1, 2, 3)):
            cbf = L.dcsky(A, B, mode)
            ebf = (np.abs(cbf - ref) / np.abs(ref)).max()
            assert ebf < 1e-4, (M, K, N, mode, ebf)


def test_config2_full_size_properties(L):
    """Full config-2 workload: every record is a valid irradiance split
    (non-negative, total <= pi per unit sky), repeatable for a fixed seed."""
    _say("config2 full size")
    rad, sens = scenes.config2(0)
    c = _ctx(L, rad)
    assert c.info()["primitives"] == 8 + 200 * 6 * 21 - 1
    c.set_params(_rc_params(L, ambounce=4, ambdiv=4096, minweight=1e-4, seed=1))
    out = c.contrib(sens, _sky_mods(L), 1, L.RT_IMM_IRRAD)
    n1 = c.last_nrays()
    assert out.shape == (2500, 438)
    assert (out >= 0).all()
    tot = out.sum(1) / 3
    assert (tot <= math.pi * 1.02).all() and tot.max() > 0.01
    out2 = c.contrib(sens, _sky_mods(L), 1, L.RT_IMM_IRRAD)
    assert c.last_nrays() == n1
    assert np.allclose(out, out2, rtol=1e-9, atol=1e-15)
    assert n1 > 2500 * 4096


def test_contrib_on_device_buffers_matches_host_buffers(L):
    """rg_contrib_device (rays and output resident on the GPU) gives the same
    records as rg_contrib for the same seed, also when records average
    several rays."""
    import torch
    _say("contrib on device buffers")
    rad, sens = scenes.config2(2)
    c = _ctx(L, rad)
    c.set_params(_rc_params(L, ambounce=2, ambdiv=256, minweight=1e-3, seed=5))
    mods = _sky_mods(L)
    for accum in (1, 3):
        rays = np.repeat(sens[:300], accum, axis=0)
        host = c.contrib(rays, mods, accum, L.RT_IMM_IRRAD)
        n_host = c.last_nrays()
        d_rays = torch.from_numpy(rays).cuda()
        d_out = torch.full(host.shape, -1.0, dtype=torch.float64, device="cuda")
        c.contrib_device(d_rays.data_ptr(), rays.shape[0] // accum, mods, d_out.data_ptr(), accum, L.RT_IMM_IRRAD)
        assert c.last_nrays() == n_host
        assert np.allclose(d_out.cpu().numpy(), host, rtol=1e-12, atol=1e-15)
    c.close()


def test_config5_full_size_properties(L):
    """Full config-5 scene (10-storey office, 2,495,700 quads = 5 M
    triangles) under a Reinhart MF:4 sky (2305 + 1 bins), -ab 5 -ad 8192
    -lw 5e-5, on every 100th of its 20,000 sensors: every record a valid
    irradiance split (non-negative, total <= pi per unit sky), the four-wide
    and the binary walks agree, and the rays exceed the first bounce."""
    _say("config5 full size")
    rad, sens = scenes.config5(0)
    c = _ctx(L, rad)
    del rad
    assert c.info()["primitives"] == 2_495_700
    sub = sens[::100]
    mods = _sky_mods(L, 4)
    c.set_params(_rc_params(L, ambounce=5, ambdiv=8192, minweight=5e-5, seed=3))
    out = c.contrib(sub, mods, 1, L.RT_IMM_IRRAD)
    n1 = c.last_nrays()
    assert out.shape == (200, 3 * 2306)
    assert (out >= 0).all()
    tot = out.sum(1) / 3
    assert (tot <= math.pi * 1.02).all() and tot.max() > 0.01
    assert n1 > 200 * 8192
    c.set_tuning(wide=0)
    out2 = c.contrib(sub, mods, 1, L.RT_IMM_IRRAD)
    assert c.last_nrays() == n1
    assert np.allclose(out, out2, rtol=1e-9, atol=1e-15)
    c.close()


def test_sah_and_lbvh_trees_give_the_same_contributions(L):
    """The two spatial indexes change only the traversal order: closest hits,
    and so every sample, agree (up to coincident-surface ties)."""
    _say("config2 subset, SAH vs LBVH")
    rad, sens = scenes.config2(0)
    sens = sens[::5]
    outs, nrs = [], []
    for kind in (L.BVH_SAH, L.BVH_LBVH):
        c = _ctx(L, rad, kind)
        c.set_params(_rc_params(L, ambounce=2, ambdiv=512, minweight=1e-3, seed=5))
        outs.append(c.contrib(sens, _sky_mods(L), 1, L.RT_IMM_IRRAD))
        nrs.append(c.last_nrays())
        c.close()
    assert abs(nrs[0] - nrs[1]) <= 1e-4 * nrs[0]
    close = np.isclose(outs[0], outs[1], rtol=1e-6, atol=1e-12).all(1)
    assert close.mean() >= 0.99
    assert np.allclose(outs[0].sum(), outs[1].sum(), rtol=1e-3)


def test_persistent_and_per_thread_traversal_agree(L):
    """The persistent dynamic-fetch kernel (four-wide and binary walks) and
    the one-ray-per-thread kernel trace the same rays: same ray count, same
    contributions (up to the order of the FP64 additions), also with small
    refill thresholds and windows that split the generations."""
    _say("persistent vs per-thread traversal")
    rad, sens = scenes.config2(1)
    sens = sens[::5]
    res = []
    for kw in ({"trace_kernel": 0}, {"trace_kernel": 1}, {"trace_kernel": 1, "wide": 0},
               {"trace_kernel": 1, "refill": 1, "window": 50000},
               {"trace_kernel": 1, "refill": 64, "blocks_per_cu": 1, "wide": 0}):
        c = _ctx(L, rad)
        c.set_tuning(**kw)
        c.set_params(_rc_params(L, ambounce=2, ambdiv=512, minweight=1e-3, seed=7))
        res.append((c.contrib(sens, _sky_mods(L), 1, L.RT_IMM_IRRAD), c.last_nrays()))
        c.close()
    for out, nr in res[1:]:
        assert nr == res[0][1]
        assert np.allclose(out, res[0][0], rtol=1e-12, atol=1e-15)


def test_python_api_bytes(L, tmp_path):
    _say("python API")
    from pyradiance_amd import Rcontrib, rtrace
    rad, sens, _ = scenes.config1(0)
    f = tmp_path / "room.rad"
    f.write_text(rad)
    inp = "\n".join(" ".join(f"{v:.10g}" for v in r) for r in sens[:5]).encode()
    out = rtrace(inp, f, irradiance=True, params=["-ab", "0", "-dj", "0", "-ds", "0"])
    head, _, body = out.partition(b"\n\n")
    assert head.startswith(b"#?RADIANCE") and b"FORMAT=ascii" in head
    lines = body.decode().strip("\n").split("\n")
    assert len(lines) == 5 and all(len(ln.split("\t")) == 4 for ln in lines)
    g = np.load(GOLDEN / "c1_seed0.npz")
    vals = np.array(body.split(), float).reshape(-1, 3)
    assert np.allclose(vals, g["values"][:5], rtol=1e-5, atol=1e-9)
    raw = rtrace(inp, f, header=False, irradiance=True, outform="f", params=["-ab", "0", "-dj", "0", "-ds", "0"])
    assert len(raw) == 5 * 3 * 4
    sky = tmp_path / "sky.rad"
    sky.write_text(scenes.sky_rad((1, 1, 1), (1, 1, 1)))
    rc = Rcontrib(b"0 0 0 0 0 1\n", sky, params=["-I+", "-ab", "1", "-ad", "1024", "-lw", "1e-4", "-h"])
    rc.add_modifier(modifier="skyglow", calfile="reinhartb.cal", param="MF=1", nbins="Nrbins", binv="rbin")
    rc.add_modifier(modifier="groundglow", nbins="1", binv="0")
    res = rc()
    v = np.array(res.split(), float)
    assert v.size == 146 * 3 and abs(v[0::3].sum() - math.pi) < 1e-4


def test_errors_are_status_codes(L):
    c = L.Context(0)
    with pytest.raises(RuntimeError, match="undefined modifier"):
        c.load_text("nosuch polygon p 0 0 9 0 0 0 1 0 0 0 1 0\n")
    c2 = L.Context(0)
    c2.load_text("void mirror mm 0 0 3 1 1 1\nmm polygon p 0 0 9 0 0 0 1 0 0 0 1 0\n")
    with pytest.raises(RuntimeError, match="not supported"):
        c2.commit()


def test_batched_scenes_match_individual_runs(L):
    """Several scenes in one wavefront give each scene's standalone result
    (scene k of a batch draws the samples of seed + k)."""
    _say("batched scenes vs individual")
    sc = scenes.config3(5, nscenes=4, prims=(600, 6000), sensors=(20, 200))
    ctxs = [_ctx(L, rad) for rad, _ in sc]
    p = _rc_params(L, ambounce=2, ambdiv=256, minweight=1e-3, seed=21)
    ctxs[0].set_params(p)
    mods = _sky_mods(L)
    res = L.contrib_batch(ctxs, [s for _, s in sc], mods, 1, L.RT_IMM_IRRAD)
    nb = ctxs[0].last_nrays()
    total = 0
    # fresh contexts: each first call counts its scene's shadow-cache rays,
    # as the batch did
    ctxs = [_ctx(L, rad) for rad, _ in sc]
    for k, (c, (_, sens)) in enumerate(zip(ctxs, sc)):
        p.seed = 21 + k
        c.set_params(p)
        single = c.contrib(sens, mods, 1, L.RT_IMM_IRRAD)
        total += c.last_nrays()
        assert single.shape == res[k].shape
        # coefficient bins accumulate in fixed point: bit for bit
        assert np.array_equal(single, res[k]), k
    assert nb == total


def test_bins_repeat_bit_for_bit_across_launch_shapes(L):
    """rcontrib coefficients add up in 64-bit fixed point, so the order in
    which deposits arrive (launch shape, persistent refill, window size,
    child order, windows in flight) leaves every record unchanged to the
    last bit."""
    _say("bin accumulation: identical across launch shapes")
    rad, sens = scenes.config2(0, nboxes=40)
    sens = sens[::10]
    c = _ctx(L, rad)
    c.set_params(_rc_params(L, ambounce=2, ambdiv=512, minweight=1e-3, seed=5))
    mods = _sky_mods(L)
    outs = []
    for shape in (dict(trace_kernel=1, refill=16, window=1 << 21), dict(trace_kernel=1, refill=48, window=1 << 12),
                  dict(trace_kernel=0, window=1 << 14, child_order=0),
                  dict(trace_kernel=1, refill=32, window=1 << 13, child_order=1, inflight=2)):
        c.set_tuning(**shape)
        outs.append(c.contrib(sens, mods, 1, L.RT_IMM_IRRAD))
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
    assert outs[0].sum() > 0
    c.close()


def test_batched_rtrace_config1_sun_positions(L):
    """Config 1 for many sun positions in one launch equals the reference
    fixtures scene by scene (deterministic at -ab 0 -dj 0)."""
    _say("batched rtrace, config-1 suns")
    ctxs, rays = [], []
    for seed in range(3):
        rad, sens, _ = scenes.config1(seed)
        ctxs.append(_ctx(L, rad))
        rays.append(sens)
    p = L.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat = 0, 0.0, 0.0
    ctxs[0].set_params(p)
    res = L.trace_batch(ctxs, rays, L.RT_IMM_IRRAD)
    for seed in range(3):
        ref = np.load(GOLDEN / f"c1_seed{seed}.npz")["values"]
        assert np.abs(res[seed] - ref).max() <= 1e-4 * ref.max()


def test_empty_and_smallest_inputs(L, tmp_path):
    """Zero rays, zero records and 1 x 1 x 1 products are valid calls: empty
    results, no error (the reference programs print only their header for an
    empty ray stream)."""
    import torch
    from pyradiance_amd import rt
    _say("empty and smallest inputs")
    rad, _ = scenes.config2(0, nboxes=10)
    c = _ctx(L, rad)
    assert c.trace(np.zeros((0, 6)), L.RT_IMM_IRRAD).shape == (0, 3)
    t, nrm, prim = c.first_hit(np.zeros((0, 6)))
    assert t.shape == (0,) and prim.shape == (0,)
    mods = _sky_mods(L)
    c.set_params(_rc_params(L, ambounce=1, ambdiv=64, seed=1))
    assert c.contrib(np.zeros((0, 6)), mods, 1, L.RT_IMM_IRRAD).shape == (0, 438)
    assert c.last_nrays() == 0
    c.contrib_device(0, 0, mods, 0, 1, L.RT_IMM_IRRAD)
    assert c.last_nrays() == 0
    # one ray, one record, records of several rays
    one = np.array([[3.0, 2.0, 0.8, 0.0, 0.0, 1.0]])
    assert c.trace(one, L.RT_IMM_IRRAD).shape == (1, 3)
    r3 = c.contrib(np.repeat(one, 3, axis=0), mods, 3, L.RT_IMM_IRRAD)
    assert r3.shape == (1, 438) and (r3 >= 0).all()
    c.close()
    # the Python entry points on an empty ray stream
    (tmp_path / "room.rad").write_text(rad)
    out = rt.rtrace(b"", tmp_path / "room.rad", params=["-ab", "0"])
    assert out.endswith(b"\n\n") and out.startswith(b"#?RADIANCE")
    assert rt.rtrace(b"", tmp_path / "room.rad", header=False) == b""
    # 1 x 1 x 1 and 1 x K x 1 products
    A = np.full((1, 1, 3), 2.0, dtype=np.float32)
    B = np.full((1, 1, 3), 3.0, dtype=np.float32)
    for mode in (0, 1, 2, 3, 4):
        assert np.allclose(L.dcsky(A, B, mode), 6.0, rtol=1e-5)
    A = np.ones((1, 300, 3), dtype=np.float32)
    B = np.ones((300, 1, 3), dtype=np.float32)
    for mode in (1, 3):
        assert np.allclose(L.dcsky(A, B, mode), 300.0, rtol=1e-5)


def test_child_order_changes_only_the_order(L):
    """Handing a spawn's hemisphere children out in 8 x 8 blocks of strata
    instead of by stratum row changes which rays share a wavefront, not
    which samples are drawn: same records (to summation order), same rays."""
    _say("child order: blocks of strata vs rows")
    rad, sens = scenes.config2(0)
    c = _ctx(L, rad)
    c.set_params(_rc_params(L, ambounce=3, ambdiv=1024, minweight=1e-3, seed=8))
    res = []
    for order in (0, 1):
        c.set_tuning(child_order=order)
        res.append((c.contrib(sens[::25], _sky_mods(L), 1, L.RT_IMM_IRRAD), c.last_nrays()))
    c.close()
    assert res[0][1] == res[1][1]
    assert np.allclose(res[0][0], res[1][0], rtol=1e-9, atol=1e-15)
