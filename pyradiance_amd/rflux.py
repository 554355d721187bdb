"""rfluxmtx on the GPU rcontrib path (reference: util/rfluxmtx.c; Python
wrapper src/pyradiance/util.py:833-871).

rfluxmtx reads receiver surfaces annotated with `#@rfluxmtx` directives and
turns each receiver modifier into rcontrib bin arguments
(util/rfluxmtx.c:426-583), then runs rcontrib on the system scene with the
receiver appended (util/rfluxmtx.c:138-181).  With sender "-" the caller's
rays pass straight through to rcontrib (util/rfluxmtx.c:1401-1419,
1441-1446).

Here the receiver file is read with the same rules, the rcontrib arguments
are formed exactly as the reference forms them (`receiver_args`), and those
arguments map onto the GPU bin functions (Reinhart, Klems, Shirley-Chiu, CIE
sky scanner, uniform) of libradgpu.so, which runs the record loop.
"""

from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Sequence

from . import rt as _rt

FTINY = 1e-6
PARAMSTART = "@rfluxmtx"            # util/rfluxmtx.c:24 (after the '#')
_TOKEN = re.compile(r"\S+")
_FLOAT = re.compile(r"\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)")
_SURF_TYPES = ("polygon", "ring", "source")   # surf_type(), util/rfluxmtx.c:125-135


def _err(msg: str) -> RuntimeError:
    return _rt._rterr(msg)


def _normalize(v: list) -> float:
    """normalize() of common/fvect.c:132-155, first-order branch included."""
    d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
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


def _g(x: float) -> str:
    """C's %g, the format rfluxmtx prints bin parameters with."""
    return "%g" % x


@dataclass
class Surface:
    """One receiver (or sender) surface element (SURF, util/rfluxmtx.c:68-76)."""
    styp: str            # polygon / ring / source
    name: str
    farg: list
    nrm: list
    area: float


@dataclass
class _Params:
    """Directive state of the current modifier (PARAMS, util/rfluxmtx.c:78-89)."""
    hemis: str = ""
    hsiz: int = 0
    sign: str = "+"
    nrm: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    vup: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    slist: list = field(default_factory=list)   # newest first, as the reference's linked list
    outfn: str | None = None

    def reset(self) -> None:
        """clear_params(p, 1): the sampling type and size carry over."""
        self.slist = []
        self.nrm = [0.0, 0.0, 0.0]
        self.vup = [0.0, 0.0, 0.0]
        self.outfn = None


@dataclass
class Receiver:
    """A finished receiver modifier: its rcontrib arguments as the reference
    emits them and the modifier tuple the GPU path bins with."""
    modifier: str
    hemis: str
    nrm: tuple
    vup: tuple
    surfaces: list
    args: list                  # -f/-p/-bn/-b/-m (+ -o) exactly as rfluxmtx passes them
    spec: tuple                 # (name, calfile, expression, nbins, binv, param)


def _get_direction(s: str):
    """get_direction() of util/rfluxmtx.c:295-332: an axis letter with signs
    or three comma-separated numbers (the sign applies to x only, as there)."""
    sign, i = 1, 0
    while i < len(s) and s[i] in "+-":
        if s[i] == "-":
            sign = -sign
        i += 1
    c = s[i:i + 1]
    if c and c in "xyzXYZ":
        dv = [0.0, 0.0, 0.0]
        dv["xyz".index(c.lower())] = float(sign)
        return dv, (i + 1 >= len(s) or s[i + 1].isspace())
    vals, j = [], i
    for k in range(3):
        m = _FLOAT.match(s, j)
        if not m:
            return [0.0, 0.0, 0.0], False
        vals.append(float(m.group(1)))
        j = m.end()
        if k < 2:
            if j >= len(s) or s[j] != ",":
                return [0.0, 0.0, 0.0], False
            j += 1
    vals[0] *= sign
    return vals, _normalize(vals) > 0


def _parse_params(p: _Params, pargs: str) -> int:
    """parse_params() of util/rfluxmtx.c:334-411: h=, u=, o= directives."""
    n, cp, nparams = len(pargs), 0, 0

    def bad():
        return _err(f"bad parameter string:{pargs.rstrip(chr(10))}")

    while True:
        c = pargs[cp] if cp < n else "\0"
        cp += 1
        if c == "h":
            if cp >= n or pargs[cp] != "=":
                raise bad()
            cp += 1
            if cp < n and pargs[cp] in "+-":
                p.sign = pargs[cp]
                cp += 1
            else:
                p.sign = "+"
            p.hsiz, start = 0, cp
            while cp < n and not pargs[cp].isspace():
                if pargs[cp].isdigit():
                    p.hsiz = 10 * p.hsiz + int(pargs[cp])
                cp += 1
            if cp == start:
                raise bad()
            p.hemis = pargs[start:cp]
            p.hsiz += not p.hsiz
            nparams += 1
        elif c == "u":
            if cp >= n or pargs[cp] != "=":
                raise bad()
            cp += 1
            dv, ok = _get_direction(pargs[cp:])
            p.vup = dv
            if not ok:
                raise bad()
            while cp < n:            # past the token and one blank
                ch = pargs[cp]
                cp += 1
                if ch.isspace():
                    break
            nparams += 1
        elif c == "o":
            if cp >= n or pargs[cp] != "=":
                raise bad()
            cp += 1
            quot = pargs[cp] if cp < n and pargs[cp] in "\"'" else ""
            cp += bool(quot)
            start = cp
            while cp < n and (pargs[cp] != quot if quot else not pargs[cp].isspace()):
                cp += 1
            if cp == start or (quot and cp >= n):
                raise bad()
            p.outfn = pargs[start:cp]
            cp += 1
            nparams += 1
        elif c in " \t\r\n\v\f":
            continue
        elif c == "\0":
            return nparams
        else:
            raise bad()


def _finish_polygon(farg: list):
    """finish_polygon() of util/rfluxmtx.c:1012-1031: vector area."""
    nv = len(farg) // 3
    v = [farg[3 * i:3 * i + 3] for i in range(nv)]
    s = [0.0, 0.0, 0.0]
    e1 = [v[1][k] - v[0][k] for k in range(3)]
    for i in range(2, nv):
        e2 = [v[i][k] - v[0][k] for k in range(3)]
        s[0] += e1[1] * e2[2] - e1[2] * e2[1]
        s[1] += e1[2] * e2[0] - e1[0] * e2[2]
        s[2] += e1[0] * e2[1] - e1[1] * e2[0]
        e1 = e2
    area = _normalize(s) * 0.5
    return s, area


class _SceneReader:
    """load_scene() of util/rfluxmtx.c:1197-1252 over text: directives after
    '#@rfluxmtx' accumulate until the next modifier starts; other comments are
    skipped; objects go to the callback."""

    def __init__(self, text: str, name: str):
        self.t = text
        self.i = 0
        self.name = name
        self.newparams = ""

    def token(self):
        m = _TOKEN.search(self.t, self.i)
        if not m:
            return None
        self.i = m.end()
        return m.group(0)

    def int_token(self, what: str):
        tok = self.token()
        try:
            return int(tok) if tok is not None else None
        except ValueError:
            raise _err(f"bad {what} count in '{self.name}'")

    def skip_args(self) -> bool:
        for kind in ("string", "integer", "real"):
            n = self.int_token(kind)
            if n is None:
                return False
            for _ in range(max(n, 0)):
                self.token()
        return True

    def run(self, on_object) -> None:
        t, n = self.t, len(self.t)
        while self.i < n:
            c = t[self.i]
            if c.isspace():
                self.i += 1
                continue
            if c == "!":
                raise _err(f"inline commands are not supported in '{self.name}'")
            if c == "#":
                self.i += 1
                if self.i >= n:
                    break
                if not t[self.i].isspace():
                    m = _TOKEN.match(t, self.i)
                    self.i = m.end()
                    if m.group(0) == PARAMSTART:
                        j = t.find("\n", self.i)
                        j = n if j < 0 else j + 1
                        self.newparams += t[self.i:j]
                        self.i = j
                        continue
                j = t.find("\n", self.i)
                self.i = n if j < 0 else j + 1
                continue
            if not on_object(self):
                continue


class _ReceiverParser:
    """add_recv_object() / add_surface() / finish_receiver() of
    util/rfluxmtx.c:1034-1151 and :426-583."""

    def __init__(self, nowarn: bool = False, binjitter: str | None = None):
        self.cur = _Params()
        self.curmod = ""
        self.nowarn = nowarn
        self.binjitter = binjitter
        self.receivers: list[Receiver] = []
        self.warnings: list[str] = []
        # cal files are named once, where first needed (the variables stay loaded)
        self.calfirst = {"reinhartb.cal": True, "disk2square.cal": True, "klems_full.cal": True,
                         "klems_half.cal": True, "klems_quarter.cal": True, "cieskyscan.cal": True}

    def on_object(self, rd: _SceneReader) -> bool:
        thismod, otype, oname = rd.token(), rd.token(), rd.token()
        if oname is None:
            rd.i = len(rd.t)
            return False
        if otype == "alias":
            rd.token()
            return False
        if otype in _SURF_TYPES:
            if thismod != self.curmod:
                if self.curmod:
                    self.finish()
                    self.cur.reset()
                _parse_params(self.cur, rd.newparams)
                rd.newparams = ""
                self.curmod = thismod
            self.add_surface(otype, oname, rd)
            return True
        rd.skip_args()
        return False

    def add_surface(self, st: str, oname: str, rd: _SceneReader) -> None:
        for kind in ("string", "integer"):
            k = rd.int_token(kind)
            if k is None:
                return
            for _ in range(max(k, 0)):
                rd.token()
        k = rd.int_token("real")
        if k is None or k <= 0:
            return
        farg = []
        for _ in range(k):
            tok = rd.token()
            try:
                farg.append(float(tok))
            except (TypeError, ValueError):
                raise _err(f"error reading arguments for '{oname}'")
        if st == "ring":
            if len(farg) != 8:
                raise _err(f"bad argument count for surface element '{oname}'")
            nrm = farg[3:6]
            if _normalize(nrm) == 0:
                raise _err(f"bad orientation for surface element '{oname}'")
            if farg[7] < farg[6]:
                farg[6], farg[7] = farg[7], farg[6]
            area = math.pi * (farg[7] * farg[7] - farg[6] * farg[6])
        elif st == "polygon":
            if len(farg) < 9 or len(farg) % 3:
                raise _err(f"bad argument count for surface element '{oname}'")
            nrm, area = _finish_polygon(farg)
        else:
            if len(farg) != 4:
                raise _err(f"bad argument count for surface element '{oname}'")
            nrm = [-farg[0], -farg[1], -farg[2]]      # a source's "normal" points back
            if _normalize(nrm) == 0:
                raise _err(f"bad orientation for surface element '{oname}'")
            area = math.sin((math.pi / 180.0 / 2.0) * farg[3])
            area *= math.pi * area
        if area <= FTINY * FTINY and not self.nowarn:
            self.warnings.append(f"warning - zero area for surface '{oname}'")
            return
        for a in range(3):
            self.cur.nrm[a] += nrm[a] * area
        self.cur.slist.insert(0, Surface(st, oname, farg, nrm, area))

    def _cal(self, name: str):
        first = self.calfirst[name]
        self.calfirst[name] = False
        return name if first else None

    def _jitter(self, s: str) -> str:
        return s + ",JTR=" + self.binjitter if self.binjitter is not None else s

    def finish(self) -> None:
        p = self.cur
        if not self.curmod:
            raise _err("missing receiver surface!")
        args = []
        if p.outfn is not None:
            args += ["-o", p.outfn]
        if not p.hemis:
            raise _err("missing hemisphere sampling type!")
        if _normalize(p.nrm) == 0:
            raise _err("undefined normal for hemisphere sampling")
        if _normalize(p.vup) == 0:
            p.vup = [0.0, 0.0, 1.0] if abs(p.nrm[2]) < .7 else [0.0, 1.0, 0.0]
        h = p.hemis
        h0, h1 = h[0].lower(), h[1:2].lower()
        N, U = [_g(x) for x in p.nrm], [_g(x) for x in p.vup]
        orient = f"rNx={N[0]},rNy={N[1]},rNz={N[2]},Ux={U[0]},Uy={U[1]},Uz={U[2]},RHS={p.sign}1"
        uniform = False
        calfn = params = binv = binf = nbins = None
        calfile = None                     # bin file the GPU bin function follows
        if h0 == "u" or h[0] == "1":
            binv = "0" if p.slist[0].styp == "source" else f"if(-Dx*{N[0]}-Dy*{N[1]}-Dz*{N[2]},0,-1)"
            nbins, uniform = "1", True
        elif h0 == "s" and h1 == "c":
            if p.hsiz <= 1:
                raise _err("missing size for Shirley-Chiu sampling!")
            calfile = "disk2square.cal"
            params = self._jitter(f"SCdim={p.hsiz}," + orient)
            binv, nbins = "scbin", "SCdim*SCdim"
        elif h0 in ("r", "t"):
            calfile = "reinhartb.cal"
            params = self._jitter(f"MF={p.hsiz}," + orient)
            binv, nbins = "rbin", "Nrbins"
        elif h0 == "k" and (h1 == "" or h1 == "f" or h[1:2] == "1"):
            calfile, binf, nbins = "klems_full.cal", "kbin", "Nkbins"
        elif h0 == "k" and (h1 == "h" or h[1:2] == "2"):
            calfile, binf, nbins = "klems_half.cal", "khbin", "Nkhbins"
        elif h0 == "k" and (h1 == "q" or h[1:2] == "4"):
            calfile, binf, nbins = "klems_quarter.cal", "kqbin", "Nkqbins"
        elif h.lower() == "cie":
            calfile = "cieskyscan.cal"
            params = self._jitter(orient)
            binv, nbins = "cbin", "Ncbins"
        else:
            raise _err(f"unrecognized hemisphere sampling: h={h}")
        if h0 == "k":
            params = self._jitter(f"RHS={p.sign}1")
        if not uniform:
            for sp in p.slist:
                if sp.styp == "source" and abs(sp.area - math.pi) > 1e-3:
                    raise _err(f"source '{sp.name}' must be 180-degrees")
        if calfile is not None:
            calfn = self._cal(calfile)
        if calfn is not None:
            args += ["-f", calfn]
        if params is not None:
            args += ["-p", params]
        if nbins is not None:
            args += ["-bn", nbins]
        if binv is None and binf is not None:
            binv = f"{binf}({N[0]},{N[1]},{N[2]},{U[0]},{U[1]},{U[2]})"
        args += ["-b", binv, "-m", self.curmod]
        self.receivers.append(Receiver(self.curmod, h, tuple(p.nrm), tuple(p.vup), list(p.slist), args,
                                       (self.curmod, calfile, None, nbins, binv, params)))


def parse_receiver(receiver: str | Path, text: str | None = None, nowarn: bool = False,
                   binjitter: str | None = None) -> list[Receiver]:
    """Receiver modifiers of a receiver file (or of `text`), in file order."""
    if text is None:
        try:
            text = Path(receiver).read_text()
        except OSError:
            raise _err(f"cannot load '{receiver}'")
    pr = _ReceiverParser(nowarn, binjitter)
    _SceneReader(text, str(receiver)).run(pr.on_object)
    pr.finish()
    return pr.receivers


def receiver_args(receiver: str | Path, text: str | None = None, nowarn: bool = False) -> list[str]:
    """The rcontrib bin arguments rfluxmtx derives from a receiver file."""
    return [a for r in parse_receiver(receiver, text, nowarn) for a in r.args]


@dataclass
class _Screened:
    passthru: list
    fmt: str = "-faa"
    xrs: str | None = None
    yrs: str | None = None
    ldopt: str | None = None
    iropt: str | None = None
    sampcnt: int = 0
    verbose: bool = False
    nowarn: bool = False
    binjitter: str | None = None


def _screen(params: Sequence[str]) -> _Screened:
    """rfluxmtx's option screen (util/rfluxmtx.c:1268-1398): options it keeps
    for itself, options passed on to rcontrib with their argument counts, and
    the ones it refuses."""
    av = list(params)
    s = _Screened(passthru=[])
    fmt = list(s.fmt)
    a = 0

    def usage(arg):
        return _err(f"unsupported option '{arg}'")

    while a < len(av):
        arg = av[a]
        if not arg.startswith("-") or len(arg) < 2:
            raise usage(arg)
        c, c2 = arg[1], arg[2:3]
        na = 1
        if c == "v":
            s.verbose = not s.verbose
            a += 1
            continue
        if c == "f":
            if c2 == "":
                na = 2
            elif c2 in "afdc":
                fmt[2] = c2
                fmt[3] = arg[3] if len(arg) > 3 else c2
                a += 1
                continue
            else:
                raise usage(arg)
        elif c in "xy":
            if a + 1 >= len(av):
                raise usage(arg)
            if c == "x":
                s.xrs = av[a + 1]
            else:
                s.yrs = av[a + 1]
            a += 2
            continue
        elif c == "c":
            if c2 == "s":
                na = 2
            elif c2 == "w":
                na = 3
            elif c2 == "":
                if a + 1 >= len(av):
                    raise usage(arg)
                s.sampcnt = _atoi(av[a + 1])
                if s.sampcnt <= 0:
                    raise usage(av[a + 1])
                a += 2
                continue
        elif c in "Ii":
            s.iropt = arg
            a += 1
            continue
        elif c == "w":
            if c2 == "":
                s.nowarn = not s.nowarn
            else:
                s.nowarn = c2 not in "+1tTyY"
        elif c in "Vuhr":
            pass
        elif c in "nsote":
            na = 2
        elif c == "b":
            if c2 == "j":
                if a + 1 >= len(av):
                    raise usage(arg)
                s.binjitter = av[a + 1]
                a += 2
                continue
            if c2 != "v":
                raise usage(arg)
        elif c == "l":
            if c2 == "d":
                s.ldopt = arg
                a += 1
                continue
            na = 2
        elif c == "d":
            if c2 != "v":
                na = 2
        elif c == "a":
            if c2 == "p":
                na = 2
                if a + 1 < len(av) and _atoi(av[a + 1]) > 0:
                    na += 1 + (a + 2 < len(av) and _atoi(av[a + 2]) > 0)
            else:
                na = 4 if c2 == "v" else 2
        elif c == "m":
            if c2 == "":
                raise usage(arg)
            na = 4 if c2 in "ea" else 2
        else:
            raise usage(arg)
        if a + na > len(av):
            raise _err(f"missing argument for '{arg}'")
        s.passthru += av[a:a + na]
        a += na
    s.fmt = "".join(fmt)
    return s


def _atoi(x: str) -> int:
    m = re.match(r"\s*([-+]?\d+)", x)
    return int(m.group(1)) if m else 0


def _oconv_line(files: list[str], receiver: str, nowarn: bool) -> str:
    """oconv_command() of util/rfluxmtx.c:138-181: system files, receiver last."""
    def q(x):
        return f"'{x}'" if re.search(r"[\s\"'$&*;<>?\\`|()]", x) else x
    return "oconv -f " + ("-w " if nowarn else "") + " ".join([*(q(f) for f in files), q(receiver)])


class _SenderParser:
    """add_send_object() of util/rfluxmtx.c:1153-1195: every polygon or ring
    of the sender file is one sender element; the directives seen so far
    apply to the whole sender."""

    def __init__(self, nowarn: bool):
        self.recv = _ReceiverParser(nowarn)     # add_surface() with the same rules
        self.recv.cur = _Params(sign="")
        self.curmod = ""
        self.warnings: list[str] = []

    def on_object(self, rd: _SceneReader) -> bool:
        thismod, otype, oname = rd.token(), rd.token(), rd.token()
        if oname is None:
            rd.i = len(rd.t)
            return False
        if otype == "alias":
            rd.token()
            return False
        if otype in _SURF_TYPES:
            if otype == "source":
                raise _err("cannot use source as a sender!")
            if thismod != self.curmod:
                if self.curmod:
                    self.warnings.append("warning - multiple modifiers in sender")
                self.curmod = thismod
            _parse_params(self.recv.cur, rd.newparams)
            rd.newparams = ""
            self.recv.add_surface(otype, oname, rd)
            return False
        rd.skip_args()
        return False


@dataclass
class Sender:
    """Prepared sender basis (prepare_sampler(), util/rfluxmtx.c:928-1010)."""
    hemis: str
    hsiz: int
    sign: str
    nrm: list
    vup: list
    udir: list
    vdir: list
    surfaces: list
    nbins: int


def parse_sender(surface: str | Path, text: str | None = None, nowarn: bool = False) -> Sender:
    """The sender surfaces and their hemisphere basis, checked and oriented
    as prepare_sampler() does."""
    if text is None:
        try:
            text = Path(surface).read_text()
        except OSError:
            raise _err(f"cannot load '{surface}'")
    sp = _SenderParser(nowarn)
    _SceneReader(text, str(surface)).run(sp.on_object)
    p = sp.recv.cur
    if not p.slist:
        raise _err("no sender surface!")
    if not p.hemis:
        raise _err("missing sender sampling type!")
    if _normalize(p.nrm) == 0:
        raise _err("undefined normal for sender sampling")
    if _normalize(p.vup) == 0:
        p.vup = [0.0, 0.0, 1.0] if abs(p.nrm[2]) < .7 else [0.0, 1.0, 0.0]
    n, up = p.nrm, p.vup
    udir = [up[1] * n[2] - up[2] * n[1], up[2] * n[0] - up[0] * n[2], up[0] * n[1] - up[1] * n[0]]
    if _normalize(udir) == 0:
        raise _err("up vector coincides with sender normal")
    vdir = [n[1] * udir[2] - n[2] * udir[1], n[2] * udir[0] - n[0] * udir[2], n[0] * udir[1] - n[1] * udir[0]]
    if p.sign == "-":
        udir = [-x for x in udir]
    h = p.hemis
    h0 = h[0].lower()
    if h0 == "u" or h[0] == "1" or (h0 == "s" and h[1:2].lower() == "c") or h0 in ("r", "t"):
        pass
    elif h0 == "k":
        k = {"1": "1", "2": "2", "4": "4", "f": "1", "F": "1", "": "1", "h": "2", "H": "2",
             "q": "4", "Q": "4"}.get(h[1:2])
        if k is None:
            raise _err(f"unrecognized sender sampling: h={h}")
        h = "k" + k
    else:
        raise _err(f"unrecognized sender sampling: h={h}")
    from . import sender as _snd
    return Sender(h, p.hsiz, p.sign, list(n), list(p.vup), udir, vdir, list(p.slist), _snd.nbins(h, p.hsiz))


def rfluxmtx(
    receiver: str | Path,
    surface: None | str | Path = None,
    rays: None | bytes = None,
    params: None | Sequence[str] = None,
    octree: None | Path | str = None,
    scene: None | Sequence[Path | str] = None,
    seed: int | None = None,
) -> bytes:
    """rfluxmtx with the signature of pyradiance (src/pyradiance/util.py:833-871).

    Sender "-" (rays given, no `surface`): each input ray (or each -c group
    of rays) is one output record over the receivers' bins, as
    `rcontrib -fo+ [params] -faa -c N <receiver bins> <system + receiver>`.

    Sender surface: one record per bin of the sender's basis, each the mean
    of -c N (default 10000) rays sampled over that bin and the sender
    surfaces (util/rfluxmtx.c:1447-1474); `seed` seeds the sampling (fixed by
    default, as the reference's unseeded generator)."""
    argv, recs, s, snd = command(receiver, params, sender=surface)
    if any(r.args[0] == "-o" for r in recs):
        raise _err("receiver output files (o=) are not supported on the GPU path")
    files = (["-i", str(octree)] if octree is not None else []) + [str(f) for f in (scene or [])]
    oconv = _oconv_line(files, str(receiver), s.nowarn)
    paths = [f for f in files if f != "-i"] + [str(receiver)]
    nrecv = sum(len(r.args) for r in recs)
    if snd is None:
        opts = argv[:len(argv) - nrecv]
        inp = rays if rays is not None else b""
    else:
        import numpy as np
        from . import sender as _snd
        opts = argv[:len(argv) - nrecv - 2] + argv[-2:]
        rng = np.random.default_rng(0 if seed is None else seed)
        try:
            inp = _snd.sender_rays(snd, s.sampcnt, rng).astype("<f8").tobytes()
        except ValueError as e:
            raise _err(str(e))
    out, opt, mods = _rt._contrib_run(opts, [r.spec for r in recs], inp, lambda: _rt._scene_files(paths))
    # rcontrib's octree is the "!oconv ..." command: its header is that line
    # (rt/rc2.c:101-105)
    return _rt._contrib_bytes(out, opt, mods, "rcontrib", argv, None, scene_lines=[oconv])


def command(receiver: str | Path, params: None | Sequence[str] = None, text: str | None = None,
            sender: str | Path | None = None, sender_text: str | None = None):
    """The rcontrib argument list rfluxmtx forms (util/rfluxmtx.c:1401-1459),
    without the program name and the octree: (argv, receivers, screened
    options, prepared sender or None for pass-through)."""
    s = _screen(params or [])
    passthru = sender is None and sender_text is None
    if not passthru and s.iropt:
        raise _err("-i, -I supported for pass-through only")
    argv = ["-fo+", *s.passthru]
    if passthru:
        if s.iropt:
            argv.append(s.iropt)
        if s.xrs:
            argv += ["-x", s.xrs]
        if s.yrs:
            argv += ["-y", s.yrs]
        if s.ldopt:
            argv.append(s.ldopt)
        s.sampcnt = s.sampcnt if s.sampcnt > 0 else 1
    else:
        s.fmt = s.fmt[:2] + "d" + s.fmt[3]      # rays go to rcontrib as binary doubles
        s.sampcnt = s.sampcnt if s.sampcnt > 0 else 10000
    argv += [s.fmt, "-c", str(s.sampcnt)]
    recs = parse_receiver(receiver, text, nowarn=s.nowarn, binjitter=s.binjitter)
    argv += [a for r in recs for a in r.args]
    snd = None
    if not passthru:
        snd = parse_sender(sender, sender_text, nowarn=s.nowarn)
        argv += ["-y", str(snd.nbins)]
    return argv, recs, s, snd
