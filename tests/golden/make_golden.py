"""Writes the golden vectors under tests/golden/ from the reference tools
built into oracle/_ref (rtrace, rcontrib, oconv) from the reference sources.

    python tests/golden/make_golden.py [config1 bins stat_room primitives shading bins_flux rflux rflux_send mesh instance sky image]

Fixtures (inputs and the reference's outputs only):
  c1_seed{0,1,2}.npz     config-1 room: sensors, rtrace -I -ab 0 -dj 0 -ds 0 values
  c1_counts.npz          config 1's reference ray count for sun seeds 0 .. 999
  bins_reinhart.npz      sky directions and the rcontrib bin (MF 1, 2, 4)
  bins_klems.npz         directions and the klems_full.cal kbinS / kbinD bins
  stat_room.npz          48 repeats of 4 sensors, rcontrib -ab 2 -ad 256 bins
  prims.npz              mixed primitives (sphere, bubble, polygons, ring, cylinder,
                         tube, cone, cup): rays, rtrace -oLN distance + normal,
                         -os surface name
  shading.npz            local light sources (polygon, sphere, ring, proximity glow),
                         metal, trans, glass, mirror-smooth metal: rtrace radiance
                         and -I irradiance at -ab 0 -dj 0 -ds 0 -dt 0 -st 1
  bins_flux.npz          directions and the Shirley-Chiu, CIE, uniform-surface,
                         tilted left-handed Reinhart / Klems bins
  rflux_cases.json       receiver files, the rcontrib arguments rfluxmtx -v
                         forms from them, its exit status and messages
  rflux_box.npz          rays from the origin into receivers of every basis:
                         the non-zero output column and value of rfluxmtx -ab 0
  rflux_send.npz         sender surfaces of every basis: rfluxmtx's rcontrib
                         arguments and flux matrices (-ab 0, -c N samples)
  mesh/, mesh_rtrace.npz compiled meshes from the reference obj2mesh, a scene
                         placing them with transforms, and rtrace -oLNnv values
  inst/, inst_rtrace.npz octrees (frozen, file-list, nested) placed by instance
                         objects, and rtrace -oLNnv values
  vwrays.npz, image.json, image/*.hdr  reference vwrays rays and dimensions for
                         every view type; reference rpict pictures of the
                         mesh scene (-ps 1 -pj 0, deterministic shading)
  sky.npz, sky_cases.json  synthetic weather tapes (WEA of each kind, EPW) and
                         the reference gendaymtx matrices for a set of options,
                         one ASCII output with header, -D/-M sun descriptions
The config-1/2 scenes come from pyradiance_amd.scenes (seeded) and are not
stored; the two scenes above are stored as text with their vectors.
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from pyradiance_amd import scenes  # noqa: E402

REF = ROOT / "oracle" / "_ref"
REFSRC = Path(os.environ.get("RADIANCE_REF_SRC", "/root/reference/src/radiance"))
ENV = dict(os.environ, RAYPATH=f".:{REF / 'lib'}:{REFSRC / 'rt'}:{REFSRC / 'util'}:{REFSRC / 'gen'}")


def rays_text(rays: np.ndarray) -> bytes:
    return ("\n".join(" ".join(f"{v:.17g}" for v in r) for r in rays) + "\n").encode()


def oconv(rad: str, td: str, frozen: bool = False) -> str:
    """Octree of `rad` by the reference oconv.  A plain octree makes rtrace
    and rcontrib read the scene text itself, at full precision, as the GPU
    path does; a frozen one (-f) stores real arguments in Radiance's packed
    float form (about 31 bits of mantissa), which moves source directions
    and shadow-cache cells by ~1e-10."""
    radf, octf = Path(td) / "s.rad", Path(td) / "s.oct"
    radf.write_text(rad)
    with open(octf, "wb") as f:
        subprocess.run([str(REF / "oconv"), *(["-f"] if frozen else []), str(radf)], stdout=f, check=True)
    return str(octf)


def run(cmd, rays):
    r = subprocess.run(cmd, input=rays_text(rays), capture_output=True, env=ENV, check=True)
    return np.array(r.stdout.split(), dtype=np.float64)


def config1():
    """Config-1 irradiance and the reference's ray count (nrays of
    rt/raytrace.c:606, printed by the counting build) per sun position."""
    for seed in range(3):
        rad, sens, sun = scenes.config1(seed)
        with tempfile.TemporaryDirectory() as td:
            octf = oconv(rad, td)
            args = ["-h", "-I", "-ab", "0", "-dj", "0", "-ds", "0", octf]
            v = run([str(REF / "rtrace"), *args], sens).reshape(-1, 3)
            r = subprocess.run([str(REF / "rtrace_count"), *args], input=rays_text(sens), capture_output=True,
                               env=ENV, check=True)
            assert np.array_equal(np.array(r.stdout.split(), dtype=np.float64).reshape(-1, 3), v)
            nrays = [int(ln.split()[1]) for ln in r.stderr.decode().splitlines() if ln.startswith("NRAYS")][0]
        np.savez_compressed(HERE / f"c1_seed{seed}.npz", sensors=sens, sun=sun, values=v, nrays=nrays)


def config1_counts(n: int = 1000):
    """The reference's ray count of config 1 (rtrace_count NRAYS: the
    shadow-cache fill toward the sun plus the traced shadow rays and their
    continuations) for sun seeds 0 .. n-1.  Sensors on the room's centre
    lines project exactly onto shadow-cache cell boundaries, so these counts
    pin the cell arithmetic bit for bit."""
    from concurrent.futures import ThreadPoolExecutor

    def one(seed):
        rad, sens, _ = scenes.config1(seed)
        with tempfile.TemporaryDirectory() as td:
            octf = oconv(rad, td)
            r = subprocess.run([str(REF / "rtrace_count"), "-h", "-I", "-ab", "0", "-dj", "0", "-ds", "0", octf],
                               input=rays_text(sens), capture_output=True, env=ENV, check=True)
        return [int(ln.split()[1]) for ln in r.stderr.decode().splitlines() if ln.startswith("NRAYS")][0]
    with ThreadPoolExecutor(os.cpu_count() or 1) as ex:
        counts = list(ex.map(one, range(n)))
    np.savez_compressed(HERE / "c1_counts.npz", seeds=np.arange(n), nrays=np.array(counts, dtype=np.int64))


def _stat_runs(octf: str, pts: np.ndarray, R: int, args: list, nproc: int | None = None):
    """R in-stream repeats of every point through the counting rcontrib
    (-n 1), the points split over parallel processes.  The reference seeds
    its generator from time(0) under -u+ (rt/rcmain.c:331-333), so the
    processes are started in different seconds to draw independent streams.
    Returns the values (npts, R, ncols), the process boundaries and each
    process's ray count."""
    import time
    nproc = nproc or os.cpu_count() or 1
    chunks = [c for c in np.array_split(np.arange(len(pts)), nproc) if len(c)]
    procs = []
    with tempfile.TemporaryDirectory() as td:
        for k, ch in enumerate(chunks):
            pf = Path(td) / f"in{k}.txt"
            pf.write_bytes(rays_text(np.repeat(pts[ch], R, axis=0)))
            fin = open(pf, "rb")
            fout = open(Path(td) / f"out{k}.txt", "wb")
            procs.append((subprocess.Popen([str(REF / "rcontrib_count"), "-n", "1", *args, octf], stdin=fin,
                                           stdout=fout, stderr=subprocess.PIPE, env=ENV), fin, fout))
            time.sleep(1.05)
        vals, nrays = [], []
        for k, ((pr, fin, fout), ch) in enumerate(zip(procs, chunks)):
            _, err = pr.communicate()
            fin.close()
            fout.close()
            assert pr.returncode == 0, err.decode()
            out = (Path(td) / f"out{k}.txt").read_bytes()
            vals.append(np.array(out.split(), dtype=np.float64).reshape(len(ch), R, -1))
            nrays.append([int(ln.split()[1]) for ln in err.decode().splitlines() if ln.startswith("NRAYS")][0])
    bounds = np.array([0] + [int(c[-1]) + 1 for c in chunks])
    return np.concatenate(vals), bounds, np.array(nrays, dtype=np.int64)


def _save_stat(name: str, pts, R, v, bounds, nrays, args, **extra):
    """Per-(point, column) mean and variance (FP64) of R estimates, the
    per-estimate totals of each channel, and the ray counts."""
    np.savez_compressed(HERE / name, points=pts, repeats=R, mean=v.mean(1), var=v.var(1, ddof=1),
                        totals=np.stack([v[..., c::3].sum(2) for c in range(3)], -1), bounds=bounds, nrays=nrays,
                        args=np.array(args), **extra)


C2_ARGS = ["-h", "-I+", "-ab", "4", "-ad", "4096", "-lw", "1e-4", "-aa", "0", "-c", "1",
           "-f", "reinhartb.cal", "-p", "MF=1", "-bn", "Nrbins", "-b", "rbin", "-m", "skyglow",
           "-b", "0", "-bn", "1", "-m", "groundglow"]


def c2_stat():
    """Config 2 at the benchmark's options (rcontrib -I+ -ab 4 -ad 4096
    -lw 1e-4 -aa 0 -c 1, Reinhart MF:1 + ground): every 10th of the 2,500
    sensors, 16 in-stream estimates each."""
    rad, sens = scenes.config2(0)
    pts = sens[::10]
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(rad, td)
        v, bounds, nrays = _stat_runs(octf, pts, 16, C2_ARGS)
    _save_stat("c2_stat.npz", pts, 16, v, bounds, nrays, C2_ARGS, sensor_stride=10)


RC_SCENE = ("void plastic gray 0 0 5 .5 .5 .5 0 0\n"
            "gray polygon floor 0 0 12 -4 -4 0 4 -4 0 4 4 0 -4 4 0\n"
            "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 180\n"
            "void glow groundglow 0 0 4 .5 .5 .5 0\ngroundglow source ground 0 0 4 0 0 -1 180\n")


def rc_records():
    """rcontrib's record loop (rt/rcontrib.c:379-427) at -ab 0 (exact values):
    11 rays from above the floor, three of them with zero direction, under
    -c 1, -c 3 (a partial final record of two rays) and -c 0, Reinhart MF:1
    sky + ground bins; the outputs and the warnings the reference prints."""
    rng = np.random.default_rng(5)
    d = rng.normal(size=(11, 3))
    d[:, 2] = np.abs(d[:, 2])
    d[[2, 5, 9]] = 0.0
    d[7, 2] = -abs(d[7, 2])                 # one toward the floor
    rays = np.hstack([np.tile([0.3, -0.2, 1.0], (11, 1)), d])
    out = {"rays": rays, "scene": np.array(RC_SCENE)}
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(RC_SCENE, td)
        for c in ("1", "3", "0"):
            r = subprocess.run([str(REF / "rcontrib"), "-h", "-ab", "0", "-c", c, "-f", "reinhartb.cal", "-p", "MF=1",
                                "-bn", "Nrbins", "-b", "rbin", "-m", "skyglow", "-b", "0", "-bn", "1", "-m",
                                "groundglow", octf], input=rays_text(rays), capture_output=True, env=ENV, check=True)
            out[f"c{c}"] = np.array(r.stdout.split(), dtype=np.float64).reshape(-1, 438)
            out[f"c{c}_stderr"] = np.array(r.stderr.decode())
    np.savez_compressed(HERE / "rc_records.npz", **out)


ROUGH_SCENE = ("void plastic rplast 0 0 5 .6 .5 .4 .05 .15\n"
               "void metal rmetal 0 0 5 .8 .7 .6 .9 .08\n"
               "void trans rtrans 0 0 7 .7 .8 .9 .04 .1 .6 .5\n"
               "rplast polygon floor 0 0 12 -10 -10 0 10 -10 0 10 10 0 -10 10 0\n"
               "rmetal polygon wall 0 0 12 5 -3 0 5 3 0 5 3 3 5 -3 3\n"
               "rtrans polygon pane 0 0 12 -5 -3 0.5 -5 -3 3 -5 3 3 -5 3 0.5\n"
               "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 180\n"
               "void glow groundglow 0 0 4 1 1 1 0\ngroundglow source ground 0 0 4 0 0 -1 180\n")
ROUGH_ARGS = ["-h", "-ab", "0", "-st", "0", "-c", "100", "-f", "reinhartb.cal", "-p", "MF=2", "-bn", "Nrbins",
              "-b", "rbin", "-m", "skyglow", "-b", "0", "-bn", "1", "-m", "groundglow"]


def rough_stat():
    """Rough specular sampling (gaussamp, rt/normal.c:364-495): primary rays
    onto a rough plastic floor (specularity .05, roughness .15), a rough
    metal wall (.9, .08) and a rough trans pane (.04, .1; transmitted
    specular .5) at several incidence angles, rcontrib -ab 0 -st 0, so every
    value is a Gaussian lobe sample binned by direction (Reinhart MF:2 sky +
    ground).  48 records of -c 100 per ray."""
    rays = []
    for th in (0, 30, 60, 75):                      # onto the floor at (0, 0, 0)
        t = np.radians(th)
        d = np.array([np.sin(t), 0.3 * np.sin(t), -np.cos(t)])
        d /= np.linalg.norm(d)
        rays.append(np.r_[-3 * d, d])
    for th in (0, 45, 70):                          # onto the metal wall at (5, 0, 1.5)
        t = np.radians(th)
        d = np.array([np.cos(t), np.sin(t), -0.1])
        d /= np.linalg.norm(d)
        rays.append(np.r_[np.array([5, 0, 1.5]) - 2 * d, d])
    for th in (0, 40):                              # through the pane at (-5, 0, 2)
        t = np.radians(th)
        d = np.array([-np.cos(t), np.sin(t), 0.05])
        d /= np.linalg.norm(d)
        rays.append(np.r_[np.array([-5, 0, 2]) - 2 * d, d])
    pts = np.array(rays)
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(ROUGH_SCENE, td)
        v, bounds, nrays = _stat_runs_acc(octf, pts, 48, 100, ROUGH_ARGS)
    _save_stat("rough_stat.npz", pts, 48, v, bounds, nrays, ROUGH_ARGS, scene=np.array(ROUGH_SCENE), accum=100)


def _stat_runs_acc(octf, pts, R, acc, args):
    """R records of `acc` in-stream rays per point (rcontrib -c acc)."""
    import time
    chunks = [c for c in np.array_split(np.arange(len(pts)), min(len(pts), os.cpu_count() or 1)) if len(c)]
    procs = []
    with tempfile.TemporaryDirectory() as td:
        for k, ch in enumerate(chunks):
            pf = Path(td) / f"in{k}.txt"
            pf.write_bytes(rays_text(np.repeat(pts[ch], R * acc, axis=0)))
            fin = open(pf, "rb")
            fout = open(Path(td) / f"out{k}.txt", "wb")
            procs.append((subprocess.Popen([str(REF / "rcontrib_count"), "-n", "1", *args, octf], stdin=fin,
                                           stdout=fout, stderr=subprocess.PIPE, env=ENV), fin, fout))
            time.sleep(1.05)
        vals, nrays = [], []
        for k, ((pr, fin, fout), ch) in enumerate(zip(procs, chunks)):
            _, err = pr.communicate()
            fin.close()
            fout.close()
            assert pr.returncode == 0, err.decode()
            out = (Path(td) / f"out{k}.txt").read_bytes()
            vals.append(np.array(out.split(), dtype=np.float64).reshape(len(ch), R, -1))
            nrays.append([int(ln.split()[1]) for ln in err.decode().splitlines() if ln.startswith("NRAYS")][0])
    bounds = np.array([0] + [int(c[-1]) + 1 for c in chunks])
    return np.concatenate(vals), bounds, np.array(nrays, dtype=np.int64)


C4_ARGS = ["-h", "-I+", "-ab", "3", "-ad", "2048", "-aa", "0", "-c", "1",
           "-f", "reinhartb.cal", "-p", "MF=1", "-bn", "Nrbins", "-b", "rbin", "-m", "skyglow",
           "-b", "0", "-bn", "1", "-m", "groundglow"]


def c4_stat():
    """Config 4's forward pass (24 rough plastic groups, specularity up to
    .05 and roughness up to .2, glass window, binned sky; rcontrib -I+ -ab 3
    -ad 2048 with rcontrib's other defaults): every 10th of its 1000
    sensors, 16 in-stream estimates each."""
    rad, sens, _ = scenes.config4(seed=0)
    pts = sens[::10]
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(rad, td)
        v, bounds, nrays = _stat_runs(octf, pts, 16, C4_ARGS)
    _save_stat("c4_stat.npz", pts, 16, v, bounds, nrays, C4_ARGS, sensor_stride=10)


# the reference's own rcontrib test (tests/test_rcontrib.py:9-76): its
# octree (tests/Resources/contrib.oct, kept under refres/), rays, options and
# modifiers, the in-process manager's settings written as rcontrib options
CONTRIB_RAYS = np.array([[10.0, 10.0, 3.0, 0.0, 0.0, 1.0], [4.0, 5.0, 3.0, 0.0, 0.0, 1.0]])
CONTRIB_ARGS = ["-h", "-I+", "-u+", "-dj", "0.9", "-dr", "3", "-dp", "512", "-ds", "0.2", "-st", "0", "-ss", "1",
                "-lr", "-10", "-lw", "2e-3", "-ar", "256", "-ad", "350", "-ab", "6", "-aa", "0", "-as", "0",
                "-c", "1", "-f", "reinhartb.cal", "-p", "MF=1,rNx=0,rNy=0,rNz=-1,Ux=0,Uy=1,Uz=0,RHS=+1",
                "-bn", "1", "-b", "if(-Dx*0-Dy*0-Dz*-1,0,-1)", "-m", "groundglow",
                "-bn", "Nrbins", "-b", "rbin", "-m", "skyglow"]


def contrib_stat():
    """The reference's rcontrib test case: 48 in-stream estimates of its two
    rays (its assertion: the result sums to more than 5)."""
    octf = str(HERE / "refres" / "contrib.oct")
    v, bounds, nrays = _stat_runs(octf, CONTRIB_RAYS, 48, CONTRIB_ARGS, nproc=2)
    _save_stat("contrib_stat.npz", CONTRIB_RAYS, 48, v, bounds, nrays, CONTRIB_ARGS)


def bins():
    rng = np.random.default_rng(2024)
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d[:, 2] = np.abs(d[:, 2]) + 1e-9
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([np.zeros_like(d), d])
    sky = "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 180\n"
    out = {"dirs": d}
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(sky, td)
        for mf in (1, 2, 4):
            v = run([str(REF / "rcontrib"), "-h", "-ab", "0", "-f", "reinhartb.cal", "-p", f"MF={mf}",
                     "-bn", "Nrbins", "-b", "rbin", "-m", "skyglow", octf], rays)
            v = v.reshape(len(rays), -1)[:, 0::3]
            hit = v > 0.5
            b = np.where(hit.any(1), hit.argmax(1), -1)
            out[f"mf{mf}"] = b.astype(np.int32)
    np.savez_compressed(HERE / "bins_reinhart.npz", **out)
    # Klems: full directions over the sphere, sky source covering everything
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([np.zeros_like(d), d])
    sphere = "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 360\n"
    out = {"dirs": d}
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(sphere, td)
        for name in ("kbinS", "kbinD", "kbinE"):
            v = run([str(REF / "rcontrib"), "-h", "-ab", "0", "-f", "klems_full.cal", "-bn", "Nkbins",
                     "-b", name, "-m", "skyglow", octf], rays)
            v = v.reshape(len(rays), -1)[:, 0::3]
            hit = v > 0.5
            out[name] = np.where(hit.any(1), hit.argmax(1), -1).astype(np.int32)
    np.savez_compressed(HERE / "bins_klems.npz", **out)


MIXED = ("void plastic m 0 0 5 .5 .5 .5 0 0\n"
         "void glass g 0 0 3 .8 .8 .8\n"
         "m sphere s1 0 0 4 0 0 0 1\n"
         "m bubble s2 0 0 4 3 0 0 0.7\n"
         "m polygon floor 0 0 12 -6 -6 -2 6 -6 -2 6 6 -2 -6 6 -2\n"
         "m polygon tri 0 0 9 -3 -3 0 -1 -3 0 -2 -1 1\n"
         "g polygon pane 0 0 12 -4 2 -1 -2 2 -1 -2 2 1 -4 2 1\n"
         "m polygon notch 0 0 18 -5 -4 -1.5 -3 -4 -1.5 -3 -4 0.5 -4 -4 -0.5 -5 -4 0.5 -5 -4 -1.5\n"
         "m ring r1 0 0 8 0 3 0 0 1 0 0.2 1.2\n"
         "m ring disc 0 0 8 -3 3.5 1 0.3 0.2 0.9 0 0.8\n"
         "m cylinder c1 0 0 7 -3 0 -2 -3 0 2 0.5\n"
         "m tube t1 0 0 7 3 3 -2 3 3 2 0.6\n"
         "m cone k1 0 0 8 0 -3 -2 0 -3 1 1.0 0.2\n"
         "m cup k2 0 0 8 2 -3 -2 2 -3 0 0.1 0.8\n"
         "m cone k3 0 0 8 -1 1 3 1 2 4 0 0.7\n"
         "m cylinder c2 0 0 7 1 -1 3 3 -2 4.5 0.4\n")


def primitives():
    rng = np.random.default_rng(7)
    n = 4000
    org = rng.uniform(-5, 5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([org, d])
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(MIXED, td)
        r = subprocess.run([str(REF / "rtrace"), "-h", "-ab", "0", "-fad", "-oLN", octf], input=rays_text(rays),
                           capture_output=True, env=ENV, check=True)
        ln = np.frombuffer(r.stdout, dtype=np.float64).reshape(n, 4)
        r = subprocess.run([str(REF / "rtrace"), "-h", "-ab", "0", "-os", octf], input=rays_text(rays),
                           capture_output=True, env=ENV, check=True)
        names = r.stdout.decode().split()
    assert len(names) == n
    np.savez_compressed(HERE / "prims.npz", scene=np.array(MIXED), rays=rays, t=ln[:, 0], normal=ln[:, 1:],
                        name=np.array(names))


SHADING = ("void plastic floorm 0 0 5 .4 .35 .3 .03 .08\n"
           "void plastic plain 0 0 5 .6 .6 .6 0 0\n"
           "void metal brushed 0 0 5 .7 .6 .5 .9 .06\n"
           "void metal mirrorm 0 0 5 .8 .8 .8 1 0\n"
           "void trans frosted 0 0 7 .6 .7 .8 .04 .1 .5 .3\n"
           "void trans clear 0 0 7 .8 .8 .8 .02 0 .6 .9\n"
           "void glass gl 0 0 3 .85 .85 .85\n"
           "void light lamp 0 0 3 80 80 80\n"
           "void light warm 0 0 3 60 45 30\n"
           "void glow gp 0 0 4 4 4 4 6\n"
           "floorm polygon floor 0 0 12 -5 -5 0 5 -5 0 5 5 0 -5 5 0\n"
           "lamp polygon panel 0 0 12 -.5 -.5 3 -.5 .5 3 .5 .5 3 .5 -.5 3\n"
           "warm sphere bulb 0 0 4 2 1.5 2 .15\n"
           "warm ring halo 0 0 8 -2 2 2.5 0 0 -1 0 .25\n"
           "gp polygon glowpanel 0 0 12 4 -1 .5 4 -1 1.5 4 1 1.5 4 1 .5\n"
           "brushed polygon metaltop 0 0 12 .5 -2 .6 1.5 -2 .6 1.5 -1 .6 .5 -1 .6\n"
           "mirrorm polygon mirror 0 0 12 -1 -3 .2 1 -3 .2 1 -3 1.8 -1 -3 1.8\n"
           "frosted polygon pane1 0 0 12 -2.5 -.5 .1 -2.5 .5 .1 -2.5 .5 1.4 -2.5 -.5 1.4\n"
           "clear polygon pane2 0 0 12 2.5 -.5 .1 2.5 -.5 1.4 2.5 .5 1.4 2.5 .5 .1\n"
           "gl polygon window 0 0 12 -1 2.5 .1 1 2.5 .1 1 2.5 1.6 -1 2.5 1.6\n"
           "plain cylinder post 0 0 7 -1 -1 0 -1 -1 1.5 .2\n"
           "plain cone shade 0 0 8 1.5 1 1.2 1.5 1 1.6 .35 .15\n"
           "plain sphere ball 0 0 4 -1.2 1.2 .4 .4\n")

SHADING_ARGS = ["-ab", "0", "-dj", "0", "-ds", "0", "-dt", "0", "-dc", "1", "-dr", "0", "-st", "1", "-lr", "8",
                "-lw", "1e-4", "-av", "0", "0", "0"]


def shading():
    rng = np.random.default_rng(11)
    n = 1500
    org = np.column_stack([rng.uniform(-3.5, 3.5, n), rng.uniform(-3.5, 3.5, n), rng.uniform(0.3, 2.6, n)])
    d = rng.normal(size=(n, 3))
    d[:, 2] -= 0.8                       # mostly downward: floor, tops, panes, mirror
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([org, d])
    # irradiance sensors on the floor, the metal top and facing the panes
    ns = 400
    sens = np.column_stack([rng.uniform(-4, 4, ns), rng.uniform(-4, 4, ns), np.full(ns, 0.0)])
    sdir = np.tile([0.0, 0.0, 1.0], (ns, 1))
    side = rng.random(ns) < 0.25
    sens[side, 2] = rng.uniform(0.2, 1.3, side.sum())
    sdir[side] = rng.normal(size=(side.sum(), 3))
    sdir /= np.linalg.norm(sdir, axis=1)[:, None]
    sensors = np.hstack([sens, sdir])
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(SHADING, td)
        # the counting build prints the reference's nrays (rt/raytrace.c:606);
        # -u- because under -u+ a flat source's sampling axes are drawn at
        # random (rt/srcsupp.c:79), and they orient its shadow-cache cells,
        # so the count (never a value) changes from run to run
        r = subprocess.run([str(REF / "rtrace_count"), "-h", "-fad", "-u-", *SHADING_ARGS, octf], input=rays_text(rays),
                           capture_output=True, env=ENV, check=True)
        val = np.frombuffer(r.stdout, dtype=np.float64).reshape(n, 3)
        nr_val = [int(ln.split()[1]) for ln in r.stderr.decode().splitlines() if ln.startswith("NRAYS")][0]
        r = subprocess.run([str(REF / "rtrace_count"), "-h", "-fad", "-u-", "-I", *SHADING_ARGS, octf],
                           input=rays_text(sensors), capture_output=True, env=ENV, check=True)
        irr = np.frombuffer(r.stdout, dtype=np.float64).reshape(ns, 3)
        nr_irr = [int(ln.split()[1]) for ln in r.stderr.decode().splitlines() if ln.startswith("NRAYS")][0]
    np.savez_compressed(HERE / "shading.npz", scene=np.array(SHADING), args=np.array(SHADING_ARGS), rays=rays,
                        values=val, sensors=sensors, irradiance=irr, nrays_values=nr_val, nrays_irradiance=nr_irr)


def stat_room():
    rng = np.random.default_rng(1)
    room = scenes.Room()
    rad = scenes.room_rad(room) + scenes.furniture_rad(rng, 5, room, 1, 1) + scenes.sky_rad((1, 1, 1), (1, 1, 1))
    pts = np.array([[1.0, 1.0, 0.8, 0, 0, 1], [3.0, 2.0, 0.8, 0, 0, 1], [5.0, 3.0, 1.5, 0, 0, 1],
                    [3.0, 0.5, 1.5, 0, -1, 0]])
    R = 48
    rays = np.repeat(pts, R, axis=0)
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(rad, td)
        v = run([str(REF / "rcontrib"), "-h", "-n", "1", "-I+", "-ab", "2", "-ad", "256", "-lw", "1e-3", "-c", "1",
                 "-f", "reinhartb.cal", "-p", "MF=1", "-bn", "Nrbins", "-b", "rbin", "-m", "skyglow",
                 "-b", "0", "-bn", "1", "-m", "groundglow", octf], rays)
    np.savez_compressed(HERE / "stat_room.npz", points=pts, repeats=R, values=v.reshape(len(pts), R, -1))


def bins_flux():
    """Bins of the rfluxmtx receiver bases through the reference cal files:
    Shirley-Chiu (disk2square.cal), CIE sky scanner (cieskyscan.cal), the
    uniform-surface expression, and Reinhart / Klems with tilted frames and
    left-handed coordinates, over directions on the whole sphere."""
    rng = np.random.default_rng(77)
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([np.zeros_like(d), d])
    sphere = "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 360\n"
    tilt = "rNx=0.3,rNy=-0.4,rNz=-0.866025,Ux=0,Uy=0,Uz=1"
    specs = {
        "sc4": ("disk2square.cal", "SCdim=4,rNx=0,rNy=0,rNz=-1,Ux=0,Uy=1,Uz=0,RHS=+1", "SCdim*SCdim", "scbin"),
        "sc6_tilt_lh": ("disk2square.cal", "SCdim=6," + tilt + ",RHS=-1", "SCdim*SCdim", "scbin"),
        "cie": ("cieskyscan.cal", "rNx=0,rNy=0,rNz=-1,Ux=0,Uy=1,Uz=0,RHS=+1", "Ncbins", "cbin"),
        "cie_tilt_lh": ("cieskyscan.cal", tilt + ",RHS=-1", "Ncbins", "cbin"),
        "r2_tilt_lh": ("reinhartb.cal", "MF=2," + tilt + ",RHS=-1", "Nrbins", "rbin"),
        "kh_tilt_lh": ("klems_half.cal", "RHS=-1", "Nkhbins", "khbin(0.3,-0.4,-0.866025,0,0,1)"),
        "kq_tilt": ("klems_quarter.cal", "RHS=+1", "Nkqbins", "kqbin(-0.6,0,0.8,0,1,0)"),
        "half": (None, None, "1", "if(-Dx*0.3-Dy*-0.5-Dz*0.812404,0,-1)"),
    }
    out = {"dirs": d}
    with tempfile.TemporaryDirectory() as td:
        octf = oconv(sphere, td)
        for name, (cal, prm, nb, bv) in specs.items():
            cmd = [str(REF / "rcontrib"), "-h", "-ab", "0"]
            if cal:
                cmd += ["-f", cal]
            if prm:
                cmd += ["-p", prm]
            cmd += ["-bn", nb, "-b", bv, "-m", "skyglow", octf]
            v = run(cmd, rays).reshape(len(rays), -1)[:, 0::3]
            hit = v > 0.5
            assert (hit.sum(1) <= 1).all()
            out[name] = np.where(hit.any(1), hit.argmax(1), -1).astype(np.int32)
            out[name + "_spec"] = np.array([cal or "", prm or "", nb, bv])
    np.savez_compressed(HERE / "bins_flux.npz", **out)


# Receiver files for rfluxmtx (sender "-").  "box" surrounds the origin with
# glow receivers of every sampling type; the others exercise directive
# parsing and the reference's error messages.
RFLUX_BOX = """# receivers around the origin, one per sampling basis
#@rfluxmtx h=kf u=Z
void glow wall_w 0 0 4 1 1 1 0
wall_w polygon west 0 0 12 -3 -1 -1  -3 1 -1  -3 1 1  -3 -1 1

#@rfluxmtx h=-kh
#@rfluxmtx u=+Z
void glow wall_e 0 0 4 1 1 1 0
wall_e polygon east 0 0 12 3 -1 -1  3 -1 1  3 1 1  3 1 -1

#@rfluxmtx h=sc4 u=Z
void glow disc_n 0 0 4 1 1 1 0
disc_n ring north 0 0 8 0 3 0  0 -1 0  0.2 1.2

#@rfluxmtx h=u
void glow wall_s 0 0 4 1 1 1 0
wall_s polygon south 0 0 12 -1 -3 -1  -1 -3 1  1 -3 1  1 -3 -1
wall_s polygon south2 0 0 12 1.5 -2.5 -1  1.5 -2.5 1  2.5 -3.2 1  2.5 -3.2 -1

#@rfluxmtx h=cie u=Y
void glow ceil 0 0 4 1 1 1 0
ceil polygon ceiling 0 0 12 -1 -1 3  -1 1 3  1 1 3  1 -1 3

#@rfluxmtx h=t u=0.2,1,0.1
void glow flr 0 0 4 1 1 1 0
flr polygon floor 0 0 12 -1 -1 -3  1 -1 -3  1 1 -3  -1 1 -3

#@rfluxmtx h=-r2 u=-Y
void glow corner 0 0 4 1 1 1 0
corner polygon c1 0 0 9 2.5 2.5 -0.5  2.5 1.5 0.5  1.5 2.5 0.5

#@rfluxmtx h=u
void glow groundglow 0 0 4 1 1 1 0
groundglow source ground 0 0 4 0 0 -1 180

#@rfluxmtx h=r2 u=+Y
void glow skyglow 0 0 4 1 1 1 0
skyglow source sky 0 0 4 0 0 1 180
"""
RFLUX_BOX_SCENE = "void plastic grey 0 0 5 .5 .5 .5 0 0\ngrey sphere blocker 0 0 4 1.5 0.4 0.2 0.5\n"

RFLUX_PARSE = {
    "inherit": """#@rfluxmtx h=kq u=-X
void glow a 0 0 4 1 1 1 0
a polygon p1 0 0 9 0 0 0 1 0 0 0 1 0
# plain comment
void plastic m 0 0 5 .5 .5 .5 0 0
void glow b 0 0 4 1 1 1 0
b ring r1 0 0 8 0 0 1 0.3 0.2 -0.9 1.5 0.4
b polygon p2 0 0 12 0 0 2 0 1 2 1 1 2.5 1 0 2.5
void alias c b
""",
    "numeric_up": """#@rfluxmtx h=r4 u=--1,0,1
void glow sky 0 0 4 1 1 1 0
sky source s 0 0 4 0 0 1 180
#@rfluxmtx h=sc3 u=0,1,1
void glow w 0 0 4 1 1 1 0
w polygon q 0 0 12 0 0 0 0 2 0 0 2 2 0 0 2
#@rfluxmtx h=k2
void glow k 0 0 4 1 1 1 0
k polygon q2 0 0 9 0 0 0 1 1 0 0 1 1
""",
    "err_nohemis": "void glow a 0 0 4 1 1 1 0\na polygon p 0 0 9 0 0 0 1 0 0 0 1 0\n",
    "err_source90": "#@rfluxmtx h=kf\nvoid glow a 0 0 4 1 1 1 0\na source s 0 0 4 0 0 1 90\n",
    "err_badh": "#@rfluxmtx h=zz\nvoid glow a 0 0 4 1 1 1 0\na polygon p 0 0 9 0 0 0 1 0 0 0 1 0\n",
    "err_badparam": "#@rfluxmtx h=u q=1\nvoid glow a 0 0 4 1 1 1 0\na polygon p 0 0 9 0 0 0 1 0 0 0 1 0\n",
    "err_scsize": "#@rfluxmtx h=sc\nvoid glow a 0 0 4 1 1 1 0\na polygon p 0 0 9 0 0 0 1 0 0 0 1 0\n",
    "err_empty": "void glow a 0 0 4 1 1 1 0\n",
}


RFLUX_OPTS = [
    ["-I", "-ab", "2", "-ad", "64", "-c", "4", "-ffd"],
    ["-lw", "1e-4", "-x", "10", "-y", "5", "-ff", "-w"],
    ["-i", "-ld-", "-bj", "0.5", "-u+", "-dj", ".7", "-dv", "-av", "1", "1", "1", "-fad"],
    ["-h", "-V", "-lr", "8", "-aa", ".1", "-st", "0.1", "-ss", "0.5", "-ds", ".2", "-bv", "-ms", "0.2"],
    ["-fo"],
    ["-c", "0"],
    ["-q"],
]


# Sender surfaces (rfluxmtx sampling mode) with their receivers; -ab 0 and
# no other geometry, so every sample ray ends on a receiver with radiance 1.
RFLUX_SEND = {
    "window_kf": dict(
        sender="#@rfluxmtx h=kf u=Z\nvoid polygon win 0 0 12 -.5 0 0  .5 0 0  .5 0 1.5  -.5 0 1.5\n",
        receiver="#@rfluxmtx h=u\nvoid glow gg 0 0 4 1 1 1 0\ngg source ground 0 0 4 0 0 -1 180\n"
                 "#@rfluxmtx h=r1 u=Y\nvoid glow sg 0 0 4 1 1 1 0\nsg source sky 0 0 4 0 0 1 180\n",
        count=1000),
    "ring_r1": dict(
        sender="#@rfluxmtx h=r1 u=Y\nvoid ring lamp 0 0 8 0 0 2 0 0 1 0.1 0.4\n",
        receiver="#@rfluxmtx h=kq u=Y\nvoid glow fl 0 0 4 1 1 1 0\n"
                 "fl polygon floor 0 0 12 -10 -10 0 10 -10 0 10 10 0 -10 10 0\n"
                 "#@rfluxmtx h=u\nvoid glow gg 0 0 4 1 1 1 0\ngg source ground 0 0 4 0 0 -1 180\n"
                 "#@rfluxmtx h=u\nvoid glow sg 0 0 4 1 1 1 0\nsg source sky 0 0 4 0 0 1 180\n",
        count=1000),
    "pair_u": dict(
        sender="#@rfluxmtx h=u\nvoid polygon a 0 0 12 0 0 0 0 0 1 1 0 1 1 0 0\n"
               "void polygon b 0 0 12 1 0 0 1 0 1 1.8 0.6 1 1.8 0.6 0\n",
        receiver="#@rfluxmtx h=sc4 u=Z\nvoid glow w 0 0 4 1 1 1 0\nw source half 0 0 4 0 -1 0 180\n"
                 "#@rfluxmtx h=u\nvoid glow b 0 0 4 1 1 1 0\nb source back 0 0 4 0 1 0 180\n",
        count=40000),
    "sc3_cie": dict(
        sender="#@rfluxmtx h=-sc3 u=Y\nvoid polygon s 0 0 9 0 0 0 0 1 0 1 0 0\n",
        receiver="#@rfluxmtx h=cie u=Y\nvoid glow sg 0 0 4 1 1 1 0\nsg source sky 0 0 4 0 0 1 180\n"
                 "#@rfluxmtx h=u\nvoid glow gg 0 0 4 1 1 1 0\ngg source ground 0 0 4 0 0 -1 180\n",
        count=4000),
}


def rflux_send():
    """The reference rfluxmtx in sampling mode (sender surfaces): argument
    lists (-v) and the flux matrices, -ab 0, binary double output."""
    import shlex
    env = dict(ENV, PATH=f"{REF}:{os.environ.get('PATH', '')}")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, c in RFLUX_SEND.items():
            (Path(td) / "send.rad").write_text(c["sender"])
            (Path(td) / "recv.rad").write_text(c["receiver"])
            r = subprocess.run([str(REF / "rfluxmtx"), "-v", "-h", "-ab", "0", "-ffd", "-c", str(c["count"]),
                                "send.rad", "recv.rad"], capture_output=True, env=env, cwd=td, check=True)
            line = [ln for ln in r.stderr.decode().splitlines() if "opening pipe" in ln or ": running: " in ln]
            args = shlex.split(line[-1].split(": ", 2)[2])[1:-1]
            v = np.frombuffer(r.stdout, dtype="<f8")
            out[name + "_args"] = np.array(args)
            out[name] = v.reshape(int(args[args.index("-y") + 1]), -1)
            out[name + "_text"] = np.array([c["sender"], c["receiver"], str(c["count"])])
    np.savez_compressed(HERE / "rflux_send.npz", **out)


def rflux():
    """The reference rfluxmtx (sender "-"): the rcontrib arguments it forms
    for each receiver file (-v), its error messages, and the output of the
    box receivers at -ab 0 for rays from the origin."""
    import json
    import shlex
    env = dict(ENV, PATH=f"{REF}:{os.environ.get('PATH', '')}")
    cases = dict(RFLUX_PARSE, box=RFLUX_BOX)
    cases["skyr4"] = (REFSRC.parents[1] / "tests" / "Resources" / "skyr4.rad").read_text()
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, text in cases.items():
            rf = Path(td) / f"{name}.rad"
            rf.write_text(text)
            r = subprocess.run([str(REF / "rfluxmtx"), "-v", "-", str(rf)], input=b"", capture_output=True,
                               env=env, cwd=td)
            err = r.stderr.decode()
            run_line = [ln for ln in err.splitlines() if ": running: " in ln]
            args = shlex.split(run_line[0].split(": running: ", 1)[1])[1:-1] if run_line else None
            msgs = [ln.split(": ", 1)[1] for ln in err.splitlines()
                    if ln.startswith("rfluxmtx: ") and ": running: " not in ln]
            out[name] = {"text": text, "status": r.returncode, "args": args, "messages": msgs}
        # option screening: what rfluxmtx keeps, passes on and reorders
        rf = Path(td) / "skyr4.rad"
        rf.write_text(cases["skyr4"])
        opts = {}
        for i, o in enumerate(RFLUX_OPTS):
            r = subprocess.run([str(REF / "rfluxmtx"), "-v", *o, "-", str(rf)], input=b"", capture_output=True,
                               env=env, cwd=td)
            err = r.stderr.decode()
            run_line = [ln for ln in err.splitlines() if ": running: " in ln]
            args = shlex.split(run_line[0].split(": running: ", 1)[1])[1:-1] if run_line else None
            opts[str(i)] = {"options": o, "status": r.returncode, "args": args}
        out["_options"] = opts
        # the reference's own API test case (tests/test_api.py:351-361): its
        # receiver and scene files, one ray repeated 64 times, -c 64
        res = REFSRC.parents[1] / "tests" / "Resources"
        files = {n: (res / n).read_text() for n in ("materials.mat", "floor 1.rad", "ceiling.rad")}
        for n, t in files.items():
            (Path(td) / n).write_text(t)
        r = subprocess.run([str(REF / "rfluxmtx"), "-h", "-c", "64", "-", "skyr4.rad", *files],
                           input=b"1.0 1 1 0 0 1\n" * 64, capture_output=True, env=env, cwd=td, check=True)
        v = np.array(r.stdout.split(), dtype=np.float64).reshape(-1, 3)
        out["_api"] = {"files": files, "rays": "1.0 1 1 0 0 1", "repeat": 64, "ncols": len(v),
                       "nonzero": {str(i): list(v[i]) for i in np.flatnonzero(v.any(1))}}
        # values: box receivers + one occluder, rays from the origin, -ab 0
        rng = np.random.default_rng(5)
        d = rng.normal(size=(2000, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        rays = np.hstack([np.zeros_like(d), d])
        (Path(td) / "box.rad").write_text(RFLUX_BOX)
        (Path(td) / "blocker.rad").write_text(RFLUX_BOX_SCENE)
        r = subprocess.run([str(REF / "rfluxmtx"), "-h", "-ab", "0", "-", "box.rad", "blocker.rad"],
                           input=rays_text(rays), capture_output=True, env=env, cwd=td, check=True)
        v = np.array(r.stdout.split(), dtype=np.float64).reshape(len(rays), -1)
    nz = v[:, 0::3] != 0
    assert (nz.sum(1) <= 1).all()
    col = np.where(nz.any(1), nz.argmax(1), -1).astype(np.int32)
    val = np.where(col >= 0, v[np.arange(len(v)), 3 * np.maximum(col, 0)], 0.0)
    np.savez_compressed(HERE / "rflux_box.npz", rays=rays, col=col, value=val, ncols=v.shape[1] // 3,
                        receiver=np.array(RFLUX_BOX), scene=np.array(RFLUX_BOX_SCENE))
    (HERE / "rflux_cases.json").write_text(json.dumps(out, indent=1) + "\n")


def _uv_sphere_obj(nu, nv, r, center, mats, normals=True, uv=False):
    """OBJ text of a UV sphere; `mats` splits the latitude bands between
    usemtl names (None: no usemtl)."""
    cx, cy, cz = center
    lines, vid = [], lambda i, j: i * nu + (j % nu) + 1
    for i in range(nv + 1):
        th = np.pi * i / nv
        for j in range(nu):
            ph = 2 * np.pi * j / nu
            n = (np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th))
            lines.append("v %.9f %.9f %.9f" % (cx + r * n[0], cy + r * n[1], cz + r * n[2]))
            if normals:
                lines.append("vn %.9f %.9f %.9f" % n)
            if uv:
                lines.append("vt %.6f %.6f" % (j / nu, i / nv))
    bands = np.array_split(np.arange(nv), len(mats))
    for band, mat in zip(bands, mats):
        if mat:
            lines.append("usemtl " + mat)
        for i in band:
            for j in range(nu):
                a, b, c, d = vid(i, j), vid(i, j + 1), vid(i + 1, j + 1), vid(i + 1, j)
                def f(*k):
                    if normals and uv:
                        return " ".join(f"{x}/{x}/{x}" for x in k)
                    if normals:
                        return " ".join(f"{x}//{x}" for x in k)
                    return " ".join(str(x) for x in k)
                if i > 0:
                    lines.append("f " + f(a, d, b))
                if i < nv - 1:
                    lines.append("f " + f(b, d, c))
    return "\n".join(lines) + "\n"


def _torus_obj(R, r, nu, nv):
    lines = []
    for i in range(nu):
        u = 2 * np.pi * i / nu
        for j in range(nv):
            v = 2 * np.pi * j / nv
            lines.append("v %.9f %.9f %.9f" % ((R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u),
                                               r * np.sin(v)))
    idx = lambda i, j: (i % nu) * nv + (j % nv) + 1
    for i in range(nu):
        for j in range(nv):
            lines.append(f"f {idx(i, j)} {idx(i + 1, j)} {idx(i + 1, j + 1)}")
            lines.append(f"f {idx(i, j)} {idx(i + 1, j + 1)} {idx(i, j + 1)}")
    return "\n".join(lines) + "\n"


MESH_SCENE = """void plastic floor_m 0 0 5 .4 .4 .4 0 0
floor_m polygon floor 0 0 12 -5 -5 0  5 -5 0  5 5 0  -5 5 0
void light bright 0 0 3 50 50 50
bright sphere lamp 0 0 4 2 -3 5 0.5
void glass glassy 0 0 3 .9 .9 .9
void mesh ballA 7 ballA.rtm -rz 30 -t 0 0 1 0 0
glassy mesh torus 8 torus.rtm -s 0.5 -mx -t 2 0 0.5 0 0
void mesh bead 5 bead.rtm -t -2 0.5 1 0 0
"""


def mesh():
    """Compiled meshes made by the reference obj2mesh (smooth sphere with two
    materials over several patches; flat torus with the mesh object's glass,
    scaled and mirrored; smooth bead with (u,v) and a pure trans material),
    and the reference rtrace -oLNnv for rays
    aimed at them, with the deterministic settings of shading() (mesh/*.rtm,
    mesh/scene.rad, mesh_rtrace.npz)."""
    md = HERE / "mesh"
    md.mkdir(exist_ok=True)
    with tempfile.TemporaryDirectory() as td:
        t = Path(td)
        (t / "a.obj").write_text(_uv_sphere_obj(48, 32, 1.0, (0, 0, 0), ["red", "chrome"]))
        (t / "a.mat").write_text("void plastic red 0 0 5 .7 .2 .2 .03 .05\nvoid metal chrome 0 0 5 .8 .8 .8 .9 0\n")
        (t / "torus.obj").write_text(_torus_obj(0.6, 0.2, 24, 12))
        (t / "c.obj").write_text(_uv_sphere_obj(10, 8, 0.4, (0, 0, 0), ["frost"], uv=True))
        (t / "c.mat").write_text("void trans frost 0 0 7 .6 .6 .9 .04 0 .5 .8\n")
        subprocess.run([str(REF / "obj2mesh"), "-a", "a.mat", "a.obj", str(md / "ballA.rtm")], cwd=td, check=True)
        subprocess.run([str(REF / "obj2mesh"), "torus.obj", str(md / "torus.rtm")], cwd=td, check=True)
        subprocess.run([str(REF / "obj2mesh"), "-a", "c.mat", "c.obj", str(md / "bead.rtm")], cwd=td, check=True)
    (md / "scene.rad").write_text(MESH_SCENE)
    rng = np.random.default_rng(8)
    centers = np.array([[0, 0, 1], [2, 0, 0.5], [-2, 0.5, 1]])
    o = rng.normal(size=(3000, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * 6 + [0, 0, 1.5]
    o[:, 2] = np.abs(o[:, 2] - 1.5) + 0.2
    tg = centers[rng.integers(0, 3, len(o))] + rng.normal(size=(len(o), 3)) * 0.6
    d = tg - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([o, d])
    with tempfile.TemporaryDirectory() as td:
        octf = Path(td) / "scene.oct"
        with open(octf, "wb") as f:
            subprocess.run([str(REF / "oconv"), "-f", "scene.rad"], stdout=f, cwd=md, check=True)
        r = subprocess.run([str(REF / "rtrace"), "-h", *SHADING_ARGS, "-oLNnv", str(octf)], input=rays_text(rays),
                           capture_output=True, env=ENV, cwd=md, check=True)
    v = np.array(r.stdout.split(), dtype=np.float64).reshape(len(rays), 10)
    np.savez_compressed(HERE / "mesh_rtrace.npz", rays=rays, dist=v[:, 0], normal=v[:, 1:4], pnormal=v[:, 4:7],
                        value=v[:, 7:10])


INST_A = """void plastic wood 0 0 5 .5 .35 .2 0 0
void metal steel 0 0 5 .6 .6 .65 .8 .02
void trans frosted 0 0 7 .7 .8 .7 .03 0 .4 .3
wood polygon seat 0 0 12 -.4 -.4 .5  .4 -.4 .5  .4 .4 .5  -.4 .4 .5
wood polygon back 0 0 12 -.4 .4 .5  .4 .4 .5  .4 .4 1.2  -.4 .4 1.2
steel cylinder leg1 0 0 7 -.35 -.35 0  -.35 -.35 .5  .04
steel tube leg2 0 0 7 .35 -.35 0  .35 -.35 .5  .04
steel cone knob 0 0 8 0 .4 1.2  0 .4 1.5  .1 0
steel sphere ball 0 0 4 0 -.2 .8 .15
frosted ring disc 0 0 8 .2 -.1 .51  0 0 1  0 .15
"""
INST_B = """void plastic paint 0 0 5 .2 .6 .3 0 0
paint polygon panel 0 0 12 -.5 0 0  .5 0 0  .5 0 1  -.5 0 1
paint cup bowl 0 0 8 0 -.5 .2  0 -.5 .6  .1 .4
"""
INST_C = """void instance inner 5 a.oct -t 0 0 .1 0 0
void plastic base 0 0 5 .3 .3 .3 0 0
base polygon pad 0 0 12 -.6 -.6 .1  .6 -.6 .1  .6 .6 .1  -.6 .6 .1
"""
INST_SCENE = """void plastic floor_m 0 0 5 .4 .4 .4 0 0
floor_m polygon floor 0 0 12 -6 -6 0  6 -6 0  6 6 0  -6 6 0
void light bright 0 0 3 50 50 50
bright sphere lamp 0 0 4 2 -3 5 0.5
void glass glassy 0 0 3 .9 .9 .9
void instance a1 7 a.oct -rz 30 -t 1 1 0 0 0
glassy instance a2 8 a.oct -s 0.8 -mx -t -2 0 0 0 0
void instance c1 5 c.oct -t 0 -2.5 0 0 0
void instance b1 5 b.oct -t 2.5 -1 0 0 0
"""


def instance():
    """Octrees placed by instance objects: a frozen octree of every surface
    type with its own materials (placed rotated, and scaled, mirrored and
    overridden by a glass modifier), a file-list octree, and an octree that
    itself instances the first; the reference rtrace -oLNnv for rays aimed
    at them, with the deterministic settings of shading() (inst/*,
    inst_rtrace.npz)."""
    idir = HERE / "inst"
    idir.mkdir(exist_ok=True)
    for n, t in (("a.rad", INST_A), ("b.rad", INST_B), ("c.rad", INST_C), ("scene.rad", INST_SCENE)):
        (idir / n).write_text(t)
    for src, dst, frozen in (("a.rad", "a.oct", True), ("b.rad", "b.oct", False), ("c.rad", "c.oct", True)):
        with open(idir / dst, "wb") as f:
            subprocess.run([str(REF / "oconv"), *(["-f"] if frozen else []), src], stdout=f, cwd=idir, check=True)
    rng = np.random.default_rng(9)
    centers = np.array([[1, 1, 0.6], [-2, 0, 0.5], [0, -2.5, 0.7], [2.5, -1, 0.5]])
    o = rng.normal(size=(3000, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * 6 + [0, 0, 1.0]
    o[:, 2] = np.abs(o[:, 2] - 1.0) + 0.3
    tg = centers[rng.integers(0, 4, len(o))] + rng.normal(size=(len(o), 3)) * 0.4
    d = tg - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.hstack([o, d])
    with tempfile.TemporaryDirectory() as td:
        octf = Path(td) / "scene.oct"
        with open(octf, "wb") as f:
            subprocess.run([str(REF / "oconv"), "-f", "scene.rad"], stdout=f, cwd=idir, check=True)
        r = subprocess.run([str(REF / "rtrace"), "-h", *SHADING_ARGS, "-oLNnv", str(octf)], input=rays_text(rays),
                           capture_output=True, env=ENV, cwd=idir, check=True)
    v = np.array(r.stdout.split(), dtype=np.float64).reshape(len(rays), 10)
    np.savez_compressed(HERE / "inst_rtrace.npz", rays=rays, dist=v[:, 0], normal=v[:, 1:4], pnormal=v[:, 4:7],
                        value=v[:, 7:10])


def _weather(kind: str, rng) -> str:
    """Synthetic weather tapes (four days in four seasons, hourly): WEA with
    direct normal (1), direct horizontal (2) or photometric (3) columns, or an
    EPW file with a varying dew point."""
    days = [(1, 15), (3, 21), (6, 21), (9, 30)]
    lat, lon, tz = (37.62, 122.38, 120) if kind != "epw" else (-33.95, -151.18, -150)
    recs = []
    for mo, da in days:
        jd = [0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334][mo - 1] + da
        dec = 0.4093 * np.sin(2 * np.pi / 365 * (jd - 81))
        for h in range(24):
            hr = h + 0.5
            st = hr + (lon - tz) / 15.0 * (1 if kind != "epw" else -1) * -1
            alt = np.arcsin(np.sin(np.radians(lat)) * np.sin(dec) -
                            np.cos(np.radians(lat)) * np.cos(dec) * np.cos(st * np.pi / 12))
            sa = max(np.sin(alt), 0.0)
            cloud = rng.random()
            dn = 900 * sa ** 0.5 * cloud if sa > 0.02 else 0.0
            dh = (60 * sa + 90 * sa * (1 - cloud) + 5 * rng.random()) if sa > 0.0 else 0.0
            recs.append((mo, da, hr, dn, dh, sa))
    if kind == "epw":
        lines = ["LOCATION,Sydney,NSW,AUS,synthetic,947670,-33.95,151.18,10.0,3.0",
                 "DESIGN CONDITIONS,0", "TYPICAL/EXTREME PERIODS,0", "GROUND TEMPERATURES,0",
                 "HOLIDAYS/DAYLIGHT SAVINGS,No,0,0,0", "COMMENTS 1,synthetic", "COMMENTS 2,synthetic",
                 "DATA PERIODS,1,1,Data,Sunday, 1/ 1,12/31"]
        for k, (mo, da, hr, dn, dh, sa) in enumerate(recs):
            dew = 8 + 6 * np.sin(k / 7.0)
            lines.append(f"2001,{mo},{da},{int(hr + 0.5)},60,?9?9?9?9E0?9?9?9*9*9?9*9*9?9*9*9?9?9*9*9*9*9,"
                         f"20.0,{dew:.1f},60,101300,0,1415,350,{dn * sa + dh:.0f},{dn:.0f},{dh:.0f},"
                         "0,0,0,0,180,3.0,5,3,20.0,77777,9,999999999,20,0.1,0,88,0.2,0,0")
        return "\n".join(lines) + "\n"
    units = {"wea1": 1, "wea2": 2, "wea3": 3}[kind]
    lines = ["place San_Francisco_USA", f"latitude {lat}", f"longitude {lon}", f"time_zone {tz}",
             "site_elevation 5.0", f"weather_data_file_units {units}"]
    for mo, da, hr, dn, dh, sa in recs:
        a, b = dn, dh
        if units == 2:
            a = dn * sa
        elif units == 3:
            a, b = dn * 105.0, dh * 118.0
        lines.append(f"{mo} {da} {hr:.3f} {a:.1f} {b:.1f}")
    return "\n".join(lines) + "\n"


SKY_CASES = {
    "wea1_m1": ("wea1", ["-m", "1"]),
    "wea1_m2_opts": ("wea1", ["-m", "2", "-c", ".9", "1", "1.1", "-g", ".3", ".25", ".2", "-r", "30"]),
    "wea1_onesun": ("wea1", ["-m", "1", "-5", ".533"]),
    "wea1_sun_only": ("wea1", ["-m", "1", "-d"]),
    "wea1_sky_only": ("wea1", ["-m", "1", "-s"]),
    "wea1_sun_hours": ("wea1", ["-m", "1", "-u"]),
    "wea1_average": ("wea1", ["-m", "1", "-A"]),
    "wea1_solar": ("wea1", ["-m", "1", "-O1"]),
    "wea2_m1": ("wea2", ["-m", "1"]),
    "wea3_m1": ("wea3", ["-m", "1"]),
    "epw_m1": ("epw", ["-m", "1"]),
    "epw_m4_hours": ("epw", ["-m", "4", "-u", "-A"]),
}


def sky():
    """The reference gendaymtx on synthetic weather tapes: binary float
    matrices for every case of SKY_CASES, the ASCII output and header of one
    case, and the sun descriptions of -D (sky.npz, sky_cases.json)."""
    import json
    rng = np.random.default_rng(21)
    tapes = {k: _weather(k, rng) for k in ("wea1", "wea2", "wea3", "epw")}
    out, meta = {}, {"tapes": tapes, "cases": {}}
    with tempfile.TemporaryDirectory() as td:
        for k, t in tapes.items():
            (Path(td) / f"{k}.txt").write_text(t)
        for name, (tape, args) in SKY_CASES.items():
            r = subprocess.run([str(REF / "gendaymtx"), "-h", *args, "-of", f"{tape}.txt"], capture_output=True,
                               cwd=td, check=True)
            out[name] = np.frombuffer(r.stdout, dtype="<f4").copy()
            meta["cases"][name] = {"tape": tape, "args": args}
        r = subprocess.run([str(REF / "gendaymtx"), "-m", "1", "wea1.txt"], capture_output=True, cwd=td, check=True)
        meta["ascii_wea1_m1"] = r.stdout.decode()
        r = subprocess.run([str(REF / "gendaymtx"), "-D", "-", "wea1.txt"], capture_output=True, cwd=td, check=True)
        meta["suns_wea1"] = r.stdout.decode()
        r = subprocess.run([str(REF / "gendaymtx"), "-m", "1", "-D", "suns.rad", "-M", "mods.txt", "-of", "-h",
                            "epw.txt"], capture_output=True, cwd=td, check=True)
        meta["suns_epw_file"] = (Path(td) / "suns.rad").read_text()
        meta["mods_epw_file"] = (Path(td) / "mods.txt").read_text()
        out["epw_with_suns"] = np.frombuffer(r.stdout, dtype="<f4").copy()
    np.savez_compressed(HERE / "sky.npz", **out)
    (HERE / "sky_cases.json").write_text(json.dumps(meta, indent=1) + "\n")


VW_CASES = {
    "per_offset": ["-vtv", "-vp", "0", "-5", "1.5", "-vd", "0", "1", "-0.1", "-vh", "50", "-vv", "40", "-vs", ".1",
                   "-vl", "-.05", "-vo", ".5"],
    "parallel": ["-vtl", "-vp", "0", "-5", "1", "-vd", "0", "1", "0", "-vh", "6", "-vv", "4"],
    "cylinder": ["-vtc", "-vp", "0", "0", "1", "-vd", "1", "0", "0", "-vh", "180", "-vv", "60"],
    "angular": ["-vta", "-vp", "0", "0", "3", "-vd", "0", "0", "-1", "-vu", "0", "1", "0", "-vh", "180", "-vv",
                "180"],
    "hemi": ["-vth", "-vp", "0", "0", "3", "-vd", "0", ".2", "-1", "-vu", "0", "1", "0", "-vh", "180", "-vv",
             "180"],
    "planisphere": ["-vts", "-vp", "0", "-4", "1", "-vd", "0", "1", "0", "-vh", "150", "-vv", "120"],
}
RP_VIEWS = {
    "persp": ["-vtv", "-vp", "0", "-6", "2.5", "-vd", "0", "1", "-0.25", "-vh", "60", "-vv", "45"],
    "fisheye": ["-vta", "-vp", "0.2", "-0.3", "4.5", "-vd", "0", "0", "-1", "-vu", "0", "1", "0", "-vh", "150",
                "-vv", "150"],
}


def image():
    """The image path: reference vwrays rays (double output, 16 x 12 where the
    aspect allows) and -d dimensions for every view type, and reference rpict
    pictures of the mesh scene at -ps 1 -pj 0 with the deterministic shading
    settings (image/*.hdr, vwrays.npz, image.json)."""
    import json
    idir = HERE / "image"
    idir.mkdir(exist_ok=True)
    out, meta = {}, {"vwrays": VW_CASES, "rpict": RP_VIEWS, "dims": {}}
    for name, v in VW_CASES.items():
        r = subprocess.run([str(REF / "vwrays"), "-fd", "-x", "16", "-y", "12", *v], capture_output=True, check=True)
        out[name] = np.frombuffer(r.stdout, dtype="<f8").reshape(-1, 6)
        r = subprocess.run([str(REF / "vwrays"), "-d", "-x", "16", "-y", "12", *v], capture_output=True, check=True)
        meta["dims"][name] = r.stdout.decode()
    r = subprocess.run([str(REF / "vwrays"), "-fd", "-x", "16", "-y", "12", "-pa", "0", *VW_CASES["per_offset"]],
                       capture_output=True, check=True)
    out["per_offset_pa0"] = np.frombuffer(r.stdout, dtype="<f8").reshape(-1, 6)
    np.savez_compressed(HERE / "vwrays.npz", **out)
    md = HERE / "mesh"
    with tempfile.TemporaryDirectory() as td:
        octf = Path(td) / "scene.oct"
        with open(octf, "wb") as f:
            subprocess.run([str(REF / "oconv"), "-f", "scene.rad"], stdout=f, cwd=md, check=True)
        for name, v in RP_VIEWS.items():
            r = subprocess.run([str(REF / "rpict"), *v, "-x", "48", "-y", "48", "-ps", "1", "-pj", "0",
                                *SHADING_ARGS, str(octf)], capture_output=True, env=ENV, cwd=md, check=True)
            (idir / f"{name}.hdr").write_bytes(r.stdout)
    (HERE / "image.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    if not (REF / "rcontrib").exists():
        sys.exit("build the reference tools first: make -C oracle ref")
    jobs = {"config1": config1, "config1_counts": config1_counts, "c2_stat": c2_stat, "rc_records": rc_records, "rough_stat": rough_stat,
            "c4_stat": c4_stat, "contrib_stat": contrib_stat, "bins": bins, "stat_room": stat_room, "primitives": primitives,
            "shading": shading, "bins_flux": bins_flux, "rflux": rflux, "rflux_send": rflux_send,
            "mesh": mesh, "instance": instance, "sky": sky, "image": image}
    for name in sys.argv[1:] or list(jobs):
        jobs[name]()
    print("golden vectors written to", HERE)
