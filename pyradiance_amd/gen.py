"""gendaymtx with the signature of pyradiance (src/pyradiance/gen.py:236-318).

The annual Perez sky matrix of gen/gendaymtx.c: options are put on a
command line exactly as the reference wrapper does and read back with the
reference program's option rules (gen/gendaymtx.c:268-450), the weather data
is read and every time step set up on the host, and the sky patches are
evaluated on the GPU (libradgpu.so, rg_sky_*).  The output bytes follow the
reference: header, then one row per sky patch (ground first) and one column
per time step, ASCII "%.3g" triplets or binary float / double.
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

from . import _lib

SUN_ANG_DEG = 0.533
WHTEFFICACY = 179.0


def _err(msg: str, code: int = 1) -> RuntimeError:
    return RuntimeError(f"An error occurred with exit code {code}: {msg}")


def _parse_args(argv: list[str]):
    """gendaymtx's option loop: returns (options, flags) or raises."""
    o = _lib.RgSkyOpts()
    o.mfactor = 1
    o.sky_color[:] = (.96, 1.004, 1.118)
    o.ground_color[:] = (.2, .2, .2)
    o.sun_color[:] = (1., 1., 1.)
    o.rotation = 0.0
    o.nsuns = 4
    o.fixed_sun_sa = -1.0
    f = {"header": True, "outfmt": "a", "sunsfile": None, "modsfile": None, "matrix": True, "verbose": 0}
    i = 0

    def usage():
        return _err("Usage: gendaymtx [-v][-h][-A][-d|-s|-n][-u][-D file [-M modfile]][-r deg][-m N][-g r g b]"
                    "[-c r g b][-o{f|d}][-O{0|1}] [tape.wea]")

    def nums(k):
        nonlocal i
        if i + k >= len(argv):
            raise usage()
        v = [float(x) for x in argv[i + 1:i + 1 + k]]
        i += k
        return v

    while i < len(argv) and argv[i].startswith("-"):
        a = argv[i]
        c = a[1:2]
        if c == "g":
            o.ground_color[:] = nums(3)
        elif c == "v":
            f["verbose"] += 1
        elif c == "h":
            f["header"] = False
        elif c == "o":
            if a[2:3] not in ("f", "d", "a"):
                raise usage()
            f["outfmt"] = a[2]
        elif c == "O":
            if a[2:] not in ("0", "1"):
                raise usage()
            o.output = int(a[2])
        elif c == "m":
            o.mfactor = int(nums(1)[0])
        elif c == "c":
            o.sky_color[:] = nums(3)
        elif c in ("D", "n", "d"):
            if c == "D":
                i += 1
                if i >= len(argv):
                    raise usage()
                if argv[i] != "-":
                    f["sunsfile"] = argv[i]
                    i += 1
                    continue
                f["sunsfile"] = "-"
            if c in ("D", "n"):
                f["matrix"] = False
                o.mfactor = 1
            o.sky_color[:] = (0., 0., 0.)
            o.ground_color[:] = (0., 0., 0.)
        elif c == "M":
            i += 1
            if i >= len(argv):
                raise usage()
            f["modsfile"] = argv[i]
        elif c == "s":
            o.sun_color[:] = (0., 0., 0.)
        elif c == "u":
            o.sun_hours_only = 1
        elif c == "r":
            if a[2:] not in ("", "z"):
                raise usage()
            o.rotation = nums(1)[0]
        elif c == "5":
            o.nsuns = 1
            sa = math.pi / 360. * nums(1)[0]
            if sa <= 0:
                raise _err("missing solar disk size argument for '-5' option")
            o.fixed_sun_sa = sa * sa * math.pi
        elif c == "A":
            o.average = 1
        else:
            raise usage()
        i += 1
    if len(argv) - i > 1:
        raise usage()
    return o, f, (argv[i] if i < len(argv) else None)


def _sun_text(ids, good, alt, azi, illum, suncolor) -> tuple[str, str]:
    """OutputSun() of gen/gendaymtx.c:874-893 for every placed sun."""
    srad0 = math.radians(SUN_ANG_DEG / 2.)
    out, mods = [], []
    for i, g, a, z, d in zip(ids, good, alt, azi, illum):
        srad = d / (WHTEFFICACY * math.pi * srad0 * srad0) if g else 0.0
        v1 = math.cos(a)
        sv = (v1 * math.sin(z), v1 * math.cos(z), math.sin(a))
        out.append(f"\nvoid light solar{i}\n0\n0\n")
        out.append("3 %.3e %.3e %.3e\n" % (srad * suncolor[0], srad * suncolor[1], srad * suncolor[2]))
        out.append(f"\nsolar{i} source sun{i}\n0\n0\n")
        out.append("4 %.6f %.6f %.6f %.4f\n" % (sv[0], sv[1], sv[2], SUN_ANG_DEG))
        mods.append(f"solar{i}\n")
    return "".join(out), "".join(mods)


def sky_matrix(weather: str, argv: list[str]):
    """The sky matrix as (npatch, ncols, 3) float32 plus run information,
    for gendaymtx arguments `argv` (no program name, no weather file)."""
    o, f, _ = _parse_args(list(argv))
    sky = _lib.Sky(weather, o)
    try:
        mtx = sky.matrix() if f["matrix"] else None
        return mtx, sky, f, o
    except Exception:
        sky.close()
        raise


def gendaymtx(
    weather_data: str | Path | bytes,
    verbose: bool = False,
    header: bool = True,
    average: bool = False,
    sun_only: bool = False,
    sky_only: bool = False,
    sun_file: None | str = None,
    sun_mods: None | str = None,
    daylight_hours_only: bool = False,
    dryrun: bool = False,
    sky_color: None | list[float] = None,
    ground_color: None | list[float] = None,
    rotate: None | float = None,
    outform: None | str = None,
    onesun: bool = False,
    solar_radiance: bool = False,
    mfactor: int = 1,
) -> bytes:
    """Generate an annual Perez sky matrix from a weather tape (WEA or EPW)."""
    cmd = ["-m", str(mfactor)]
    if verbose:
        cmd.append("-v")
    if not header:
        cmd.append("-h")
    if average:
        cmd.append("-A")
    if sun_only:
        cmd.append("-d")
    elif sky_only:
        cmd.append("-s")
    if onesun:
        cmd.extend(["-5", ".533"])
    if sky_color:
        cmd.extend(["-c", *[str(i) for i in sky_color]])
    if ground_color:
        cmd.extend(["-g", *[str(i) for i in ground_color]])
    if dryrun:
        cmd.append("-n")
    if daylight_hours_only:
        cmd.append("-u")
    if solar_radiance:
        cmd.append("-O1")
    if sun_file is not None:
        cmd.extend(["-D", sun_file])
    if sun_mods is not None:
        cmd.extend(["-M", sun_mods])
    if rotate is not None:
        cmd.extend(["-r", str(rotate)])
    if outform is not None:
        cmd.append(f"-o{outform}")
    if isinstance(weather_data, bytes):
        text = weather_data.decode(errors="replace")
        argv = cmd
    elif isinstance(weather_data, (str, Path)):
        try:
            text = Path(weather_data).read_text(errors="replace")
        except OSError as e:
            raise _err(f"{weather_data}: {e.strerror}")
        argv = cmd + [str(weather_data)]
    else:
        raise TypeError("weather_data must be a string, Path, or bytes")
    try:
        mtx, sky, f, o = sky_matrix(text, cmd)
    except RuntimeError as e:
        if str(e).startswith("An error occurred"):
            raise
        raise _err(f"gendaymtx: {e}")
    try:
        return _output(mtx, sky, f, o, argv)
    finally:
        sky.close()


def _output(mtx, sky, f, o, argv) -> bytes:
    out = []
    if f["sunsfile"] is not None:
        ids, good, alt, azi, ill = sky.suns()
        stext, mtext = _sun_text(ids, good, alt, azi, ill, list(o.sun_color))
        if f["sunsfile"] == "-":
            out.append(stext.encode())
        else:
            Path(f["sunsfile"]).write_text(stext)
        if f["modsfile"] is not None:
            Path(f["modsfile"]).write_text(mtext)
    elif f["modsfile"] is not None:
        Path(f["modsfile"]).write_text("")
    if mtx is None:
        return b"".join(out)
    fmt = f["outfmt"]
    if f["header"]:
        lat, lon = sky.site()
        h = ["#?RADIANCE", " ".join(["gendaymtx", *argv]), "LATLONG= %.8f %.8f" % (lat, lon),
             f"NROWS={sky.npatch}", f"NCOLS={sky.ncols}"]
        if o.average:
            h.append(f"NAVERAGED={sky.ntsteps}")
        h.append("NCOMP=3")
        if fmt in "fd":
            h.append(f"BigEndian={int(sys.byteorder == 'big')}")
        h.append("FORMAT=" + {"a": "ascii", "f": "float", "d": "double"}[fmt])
        out.append(("\n".join(h) + "\n\n").encode())
    if fmt == "f":
        out.append(np.ascontiguousarray(mtx, dtype="=f4").tobytes())
    elif fmt == "d":
        out.append(mtx.astype("=f8").tobytes())
    else:
        # "%.3g %.3g %.3g\n" per column, a blank line after each patch row
        # when there is more than one column (gen/gendaymtx.c:652-663)
        out.append(_lib.format_ascii(mtx, 3, "%.3g", " ", "", group=mtx.shape[1] if mtx.shape[1] > 1 else 0))
    return b"".join(out)
