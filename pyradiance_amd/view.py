"""Views, view rays and Radiance pictures (reference: common/image.c view
handling, util/vwrays.c, the RGBE picture format of common/color.c and
common/resolu.c).

`vwrays` has the signature of pyradiance (src/pyradiance/util.py:1025-1067);
the rays are computed here with the reference's viewray() formulas.  For
rendering, `rpict` (in rt.py) generates the same rays on the GPU instead.
"""

from __future__ import annotations

import math
import re
import sys
from dataclasses import dataclass, field
from pathlib import Path
from typing import Sequence

import numpy as np

FTINY = 1e-6
VIEW_TYPES = "vlcahs"


def _err(msg: str, code: int = 1) -> RuntimeError:
    return RuntimeError(f"An error occurred with exit code {code}: {msg}")


def _normalize(v) -> float:
    """normalize() of common/fvect.c:132-155."""
    d = float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    if d == 0.0:
        return 0.0
    if 1.0 - FTINY <= d <= 1.0 + FTINY:
        ln = 0.5 + 0.5 * d
        d = 2.0 - ln
    else:
        ln = math.sqrt(d)
        d = 1.0 / ln
    v[0] *= d
    v[1] *= d
    v[2] *= d
    return ln


@dataclass
class View:
    """VIEW of common/view.h (STDVIEW defaults) plus its set-up vectors."""
    type: str = "v"
    vp: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    vdir: list = field(default_factory=lambda: [0.0, 1.0, 0.0])
    vup: list = field(default_factory=lambda: [0.0, 0.0, 1.0])
    vdist: float = 1.0
    horiz: float = 45.0
    vert: float = 45.0
    hoff: float = 0.0
    voff: float = 0.0
    vfore: float = 0.0
    vaft: float = 0.0
    hvec: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    vvec: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    hn2: float = 0.0
    vn2: float = 0.0

    def setup(self) -> None:
        """setview() of common/image.c:23-118; raises on a bad view."""
        if (self.vfore < -FTINY) or (self.vaft < -FTINY) or (self.vaft > FTINY and self.vaft <= self.vfore):
            raise ValueError("illegal fore/aft clipping plane")
        if self.vdist <= FTINY:
            raise ValueError("illegal view distance")
        self.vdir = list(self.vdir)
        self.vup = list(self.vup)
        self.vdist *= _normalize(self.vdir)
        if self.vdist == 0.0:
            raise ValueError("zero view direction")
        if _normalize(self.vup) == 0.0:
            raise ValueError("zero view up vector")
        d, u = self.vdir, self.vup
        h = [d[1] * u[2] - d[2] * u[1], d[2] * u[0] - d[0] * u[2], d[0] * u[1] - d[1] * u[0]]
        if _normalize(h) == 0.0:
            raise ValueError("view up parallel to view direction")
        vv = [h[1] * d[2] - h[2] * d[1], h[2] * d[0] - h[0] * d[2], h[0] * d[1] - h[1] * d[0]]
        if self.horiz <= FTINY:
            raise ValueError("illegal horizontal view size")
        if self.vert <= FTINY:
            raise ValueError("illegal vertical view size")
        t = self.type
        if t == "l":
            hn, vn = self.horiz, self.vert
        elif t == "v":
            if self.horiz >= 180.0 - FTINY:
                raise ValueError("illegal horizontal view size")
            if self.vert >= 180.0 - FTINY:
                raise ValueError("illegal vertical view size")
            hn, vn = 2.0 * math.tan(self.horiz * (math.pi / 360.)), 2.0 * math.tan(self.vert * (math.pi / 360.))
        elif t == "c":
            if self.horiz > 360.0 + FTINY:
                raise ValueError("illegal horizontal view size")
            if self.vert >= 180.0 - FTINY:
                raise ValueError("illegal vertical view size")
            hn, vn = self.horiz * (math.pi / 180.0), 2.0 * math.tan(self.vert * (math.pi / 360.))
        elif t == "a":
            if self.horiz > 360.0 + FTINY:
                raise ValueError("illegal horizontal view size")
            if self.vert > 360.0 + FTINY:
                raise ValueError("illegal vertical view size")
            hn, vn = self.horiz * (math.pi / 180.0), self.vert * (math.pi / 180.0)
        elif t == "h":
            if self.horiz > 180.0 + FTINY:
                raise ValueError("illegal horizontal view size")
            if self.vert > 180.0 + FTINY:
                raise ValueError("illegal vertical view size")
            hn, vn = 2.0 * math.sin(self.horiz * (math.pi / 360.)), 2.0 * math.sin(self.vert * (math.pi / 360.))
        elif t == "s":
            if self.horiz >= 360.0 - FTINY:
                raise ValueError("illegal horizontal view size")
            if self.vert >= 360.0 - FTINY:
                raise ValueError("illegal vertical view size")
            hn = 2. * math.sin(self.horiz * (math.pi / 360.)) / (1.0 + math.cos(self.horiz * (math.pi / 360.)))
            vn = 2. * math.sin(self.vert * (math.pi / 360.)) / (1.0 + math.cos(self.vert * (math.pi / 360.)))
        else:
            raise ValueError("unknown view type")
        if t not in "as":
            if t != "c":
                h = [x * hn for x in h]
            vv = [x * vn for x in vv]
        self.hvec, self.vvec = h, vv
        self.hn2, self.vn2 = hn * hn, vn * vn

    def aspect(self) -> float:
        """viewaspect(): vertical over horizontal extent."""
        return math.sqrt(self.vn2 / self.hn2)

    def header(self) -> str:
        """fprintview() of common/image.c:555-569."""
        g7 = lambda x: "%.7g" % x
        g5 = lambda x: "%.5g" % x
        return (f" -vt{self.type} -vp {' '.join(g7(x) for x in self.vp)}"
                f" -vd {' '.join(g7(x * self.vdist) for x in self.vdir)}"
                f" -vu {' '.join(g5(x) for x in self.vup)} -vh {g5(self.horiz)} -vv {g5(self.vert)}"
                f" -vo {g5(self.vfore)} -va {g5(self.vaft)} -vs {g5(self.hoff)} -vl {g5(self.voff)}")

    def rays(self, x: np.ndarray, y: np.ndarray):
        """viewray() of common/image.c:213-301 for image positions in
        [0,1)^2: (origins, directions, rmax); rmax < -FTINY where the view
        has no ray."""
        x = np.asarray(x, dtype=np.float64) + self.hoff - 0.5
        y = np.asarray(y, dtype=np.float64) + self.voff - 0.5
        vd, hv, vv, vp = (np.asarray(a, dtype=np.float64) for a in (self.vdir, self.hvec, self.vvec, self.vp))
        n = x.shape[0]
        rmax = np.zeros(n)
        t = self.type
        if t == "l":
            org = vp + self.vfore * vd + x[:, None] * hv + y[:, None] * vv
            d = np.broadcast_to(vd, (n, 3)).copy()
            if self.vaft > FTINY:
                rmax[:] = self.vaft - self.vfore
            return org, d, rmax
        if t in "vc":
            if t == "v":
                d = vd + x[:, None] * hv + y[:, None] * vv
            else:
                ang = x * self.horiz * (math.pi / 180.0)
                d = np.cos(ang)[:, None] * vd + np.sin(ang)[:, None] * hv + y[:, None] * vv
            org = vp + d * self.vfore
            ln = np.linalg.norm(d, axis=1)
            d = d / np.where(ln > 0, ln, 1.0)[:, None]
            if self.vaft > FTINY:
                rmax = (self.vaft - self.vfore) * ln
            return org, d, rmax
        if t == "h":
            z = 1.0 - x * x * self.hn2 - y * y * self.vn2
            bad = z < 0.0
            z = np.sqrt(np.maximum(z, 0.0))
        elif t == "a":
            x = x * (1.0 / 180.0) * self.horiz
            y = y * (1.0 / 180.0) * self.vert
            dd = x * x + y * y
            bad = dd > 1.0
            dd = np.sqrt(np.minimum(dd, 1.0))
            z = np.cos(math.pi * dd)
            with np.errstate(divide="ignore", invalid="ignore"):
                f = np.where(dd <= FTINY, math.pi, np.sqrt(np.maximum(1.0 - z * z, 0.0)) / dd)
            x, y = x * f, y * f
        else:
            x = x * math.sqrt(self.hn2)
            y = y * math.sqrt(self.vn2)
            dd = x * x + y * y
            z = (1. - dd) / (1. + dd)
            x, y = x * (1. + z), y * (1. + z)
            bad = np.zeros(n, dtype=bool)
        d = z[:, None] * vd + x[:, None] * hv + y[:, None] * vv
        org = vp + d * self.vfore
        if self.vaft > FTINY:
            rmax[:] = self.vaft - self.vfore
        rmax[bad] = -1.0
        return org, d, rmax


def _isfloat(s: str) -> bool:
    try:
        float(s)
        return True
    except ValueError:
        return False


def view_option(v: View, av: Sequence[str]) -> int:
    """getviewopt() of common/image.c:450-518: applies one -v option,
    returns the number of arguments it used after the option, -1 if none."""
    if not av or not av[0].startswith("-v") or len(av[0]) < 3:
        return -1
    o = av[0]
    c = o[2]

    def need(k):
        if len(o) != 3 or len(av) <= k or not all(_isfloat(a) for a in av[1:1 + k]):
            return False
        return True

    if c == "t":
        if len(o) != 4 or o[3] not in VIEW_TYPES:
            return -1
        v.type = o[3]
        return 0
    if c in "pdu":
        if not need(3):
            return -1
        vals = [float(a) for a in av[1:4]]
        if c == "p":
            v.vp = vals
        elif c == "d":
            v.vdir = vals
            v.vdist = 1.0
        else:
            v.vup = vals
        return 3
    if c in "hvoasl":
        if not need(1):
            return -1
        val = float(av[1])
        setattr(v, {"h": "horiz", "v": "vert", "o": "vfore", "a": "vaft", "s": "hoff", "l": "voff"}[c], val)
        return 1
    return -1


def scan_view(v: View, s: str) -> int:
    """sscanview(): every view option in a string; returns how many."""
    toks = s.split()
    n, i = 0, 0
    while i < len(toks):
        k = view_option(v, toks[i:i + 4])
        if k >= 0:
            n += 1
            i += k + 1
        else:
            i += 1
    return n


def view_file(path: str | Path, v: View) -> int:
    """viewfile(): view options from the header (or lines) of a view file or
    picture, from lines naming a viewing program or VIEW=."""
    data = Path(path).read_bytes()
    head = data.split(b"\n\n", 1)[0] if data.startswith(b"#?RADIANCE") else data
    n = 0
    for line in head.decode(errors="replace").splitlines():
        word = line.split()[0] if line.split() else ""
        base = word.rsplit("/", 1)[-1]
        if line.startswith("VIEW=") or any(base.startswith(p) for p in ("rpict", "rvu", "rpiece", "rxpiece",
                                                                          "pinterp", "rview", "vwrays")):
            n += scan_view(v, line[5:] if line.startswith("VIEW=") else line)
    return n


def parse_view(args: Sequence[str]) -> tuple[View, list[str]]:
    """View options (and -vf files) out of an argument list; returns the
    view and the remaining arguments."""
    v = View()
    rest = []
    i = 0
    args = list(args)
    while i < len(args):
        a = args[i]
        if a == "-vf":
            if i + 1 >= len(args) or view_file(args[i + 1], v) <= 0:
                raise _err(f"{args[i + 1] if i + 1 < len(args) else ''}: no view in file")
            i += 2
            continue
        k = view_option(v, args[i:i + 4]) if a.startswith("-v") else -1
        if k >= 0:
            i += k + 1
            continue
        rest.append(a)
        i += 1
    return v, rest


def normaspect(va: float, pa: float, x: int, y: int) -> tuple[float, int, int]:
    """normaspect() of common/image.c:196-210."""
    if pa <= FTINY:
        pa = va * x / y
    elif va * x > pa * y:
        x = int(y / va * pa + .5)
    else:
        y = int(x * va / pa + .5)
    return pa, x, y


def pixel_positions(xres: int, yres: int) -> tuple[np.ndarray, np.ndarray]:
    """Image positions of every pixel centre, scanlines top first (pix2loc()
    for the standard -Y +X orientation)."""
    sc, si = np.divmod(np.arange(xres * yres), xres)
    return (si + .5) / xres, (yres - 1 - sc + .5) / yres


def vwrays(
    pixpos: None | bytes = None,
    unbuf: bool = False,
    outform: str = "a",
    ray_count: int = 1,
    pixel_jitter: float = 0,
    pixel_diameter: float = 0,
    pixel_aspect: float = 1,
    xres: int = 512,
    yres: int = 512,
    dimensions: bool = False,
    view: None | Sequence[str] = None,
    pic: None | Path = None,
    zbuf: None | Path = None,
) -> bytes:
    """vwrays (util/vwrays.c): the rays of a view, one per pixel (or
    `ray_count` each), scanlines top first.  Jitter uses numpy's generator,
    so jittered rays agree with the reference in distribution."""
    if zbuf is not None:
        raise _err("depth buffers (zbuf) are not supported on the GPU path")
    if pixel_diameter > FTINY:
        raise _err("pixel diameter (-pd) is not supported on the GPU path")
    v = View()
    pa = float(pixel_aspect)
    if view is not None:
        v, rest = parse_view(view)
        if rest:
            raise _err(f"unsupported option '{rest[0]}'")
    elif pic is not None:
        if view_file(pic, v) <= 0:
            raise _err(f"{pic}: no view in picture")
        m = re.search(rb"\n([-+][XY]) +(\d+) +([-+][XY]) +(\d+)\n", Path(pic).read_bytes())
        if m:
            a, b = int(m.group(2)), int(m.group(4))
            xres, yres = (b, a) if m.group(1)[1:] == b"Y" else (a, b)
    else:
        raise ValueError("Either view or pic should be provided.")
    try:
        v.setup()
    except ValueError as e:
        raise _err(f"vwrays: {e}")
    if pic is None:
        pa, xres, yres = normaspect(v.aspect(), pa, xres, yres)
    if dimensions:
        return f"-x {xres} -y {yres}{' -ld+' if v.vaft > FTINY else ''}\n".encode()
    rng = np.random.default_rng()
    if pixpos is not None:
        pp = np.array(pixpos.split(), dtype=np.float64).reshape(-1, 2)
        lx, ly = (pp[:, 0] + .5) / xres, (pp[:, 1] + .5) / yres
    else:
        lx, ly = pixel_positions(xres, yres)
    c = max(int(ray_count), 1)
    lx, ly = np.repeat(lx, c), np.repeat(ly, c)
    if pixel_jitter > FTINY:
        lx = lx + pixel_jitter * (.5 - rng.random(lx.shape)) / xres
        ly = ly + pixel_jitter * (.5 - rng.random(ly.shape)) / yres
    org, d, rmax = v.rays(lx, ly)
    bad = rmax < -FTINY
    d = np.where((rmax > FTINY)[:, None], d * rmax[:, None], d)
    org[bad] = 0.0
    d[bad] = 0.0
    rays = np.hstack([org, d])
    if outform == "f":
        return rays.astype("=f4").tobytes()
    if outform == "d":
        return rays.astype("=f8").tobytes()
    return "".join("%.5e %.5e %.5e %.5e %.5e %.5e\n" % tuple(r) for r in rays).encode()


# ---- Radiance pictures ----------------------------------------------------

def rgbe(rgb: np.ndarray) -> np.ndarray:
    """setcolr() of common/color.c:798-823 on an (..., 3) array -> uint8 RGBE."""
    rgb = np.asarray(rgb, dtype=np.float64)
    d = rgb.max(axis=-1)
    m, e = np.frexp(d)
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.where(d > 1e-32, m * 256.0 / d, 0.0)
    out = np.zeros(rgb.shape[:-1] + (4,), dtype=np.uint8)
    v = (rgb * scale[..., None]).astype(np.int64)
    v = np.where(rgb > 0, v, 0)
    ok = d > 1e-32
    out[..., :3] = np.where(ok[..., None], np.clip(v, 0, 255), 0)
    out[..., 3] = np.where(ok, e + 128, 0)
    return out


def from_rgbe(c: np.ndarray) -> np.ndarray:
    """colr_color(): (m + 1/2) 2^(e-136), zero for exponent 0."""
    c = np.asarray(c)
    e = c[..., 3].astype(np.int64)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (c[..., :3].astype(np.float64) + .5) * f[..., None]


def write_picture(rgb: np.ndarray, header_lines: Sequence[str]) -> bytes:
    """A Radiance picture: header, -Y yres +X xres, flat RGBE scanlines (top
    first), which every reader of the format accepts."""
    yres, xres = rgb.shape[:2]
    head = "#?RADIANCE\n" + "".join(ln + "\n" for ln in header_lines) + "FORMAT=32-bit_rle_rgbe\n\n"
    res = "-Y %8d +X %8d\n" % (yres, xres)
    return head.encode() + res.encode() + rgbe(rgb).tobytes()


def read_picture(data: bytes) -> tuple[np.ndarray, list[str]]:
    """Header lines and the (yres, xres, 3) float image of a Radiance picture
    (flat or run-length encoded scanlines, standard orientation)."""
    end = data.index(b"\n\n")
    header = data[:end].decode(errors="replace").splitlines()[1:]
    pos = end + 2
    nl = data.index(b"\n", pos)
    res = data[pos:nl].split()
    pos = nl + 1
    if res[0] != b"-Y" or res[2] != b"+X":
        raise ValueError("unsupported picture orientation")
    yres, xres = int(res[1]), int(res[3])
    img = np.zeros((yres, xres, 4), dtype=np.uint8)
    buf = np.frombuffer(data, dtype=np.uint8)
    for y in range(yres):
        if 8 <= xres < 0x8000 and buf[pos] == 2 and buf[pos + 1] == 2 and not (buf[pos + 2] & 0x80):
            pos += 4
            for ch in range(4):
                x = 0
                while x < xres:
                    c = int(buf[pos])
                    pos += 1
                    if c > 128:
                        img[y, x:x + c - 128, ch] = buf[pos]
                        pos += 1
                        x += c - 128
                    else:
                        img[y, x:x + c, ch] = buf[pos:pos + c]
                        pos += c
                        x += c
        else:
            img[y] = buf[pos:pos + 4 * xres].reshape(xres, 4)
            pos += 4 * xres
    return from_rgbe(img), header
