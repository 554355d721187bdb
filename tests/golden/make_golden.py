"""Writes the golden vectors under tests/golden/ from the reference tools
built into oracle/_ref (rtrace, rcontrib, oconv) from the reference sources.

    python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs only):
  c1_seed{0,1,2}.npz     config-1 room: sensors, rtrace -I -ab 0 -dj 0 -ds 0 values
  bins_reinhart.npz      sky directions and the rcontrib bin (MF 1, 2, 4)
  bins_klems.npz         directions and the klems_full.cal kbinS / kbinD bins
  stat_room.npz          48 repeats of 4 sensors, rcontrib -ab 2 -ad 256 bins
The scenes themselves come from pyradiance_amd.scenes (seeded) and are not
stored.
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


def oconv(rad: str, td: str) -> str:
    radf, octf = Path(td) / "s.rad", Path(td) / "s.oct"
    radf.write_text(rad)
    with open(octf, "wb") as f:
        subprocess.run([str(REF / "oconv"), "-f", str(radf)], stdout=f, check=True)
    return str(octf)


def run(cmd, rays):
    r = subprocess.run(cmd, input=rays_text(rays), capture_output=True, env=ENV, check=True)
    return np.array(r.stdout.split(), dtype=np.float64)


def config1():
    for seed in range(3):
        rad, sens, sun = scenes.config1(seed)
        with tempfile.TemporaryDirectory() as td:
            octf = oconv(rad, td)
            v = run([str(REF / "rtrace"), "-h", "-I", "-ab", "0", "-dj", "0", "-ds", "0", octf], sens).reshape(-1, 3)
        np.savez_compressed(HERE / f"c1_seed{seed}.npz", sensors=sens, sun=sun, values=v)


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


if __name__ == "__main__":
    if not (REF / "rcontrib").exists():
        sys.exit("build the reference tools first: make -C oracle ref")
    config1()
    bins()
    stat_room()
    print("golden vectors written to", HERE)
