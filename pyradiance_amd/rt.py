"""Drop-in rtrace / Rcontrib with the signatures of pyradiance (src/pyradiance/rt.py).

The argument lists and byte formats follow the reference wrappers
(rt.py:165-235 for Rcontrib, rt.py:278-362 for rtrace); the work runs on the
GPU through libradgpu.so instead of a subprocess.  Errors surface as
RuntimeError("An error occurred with exit code N: <message>"), the shape the
reference's handle_called_process_error produces (anci.py:13-30).
"""

from __future__ import annotations

import re
import sys
import time
from pathlib import Path
from typing import Sequence

import numpy as np

from . import _lib
from . import params as _params
from .scenes import reinhart_nbins

VERSION = "RADIANCE 6.0a pyradiance_amd (MI355X)"

_contexts: dict = {}


def _rterr(msg: str, code: int = 1) -> RuntimeError:
    return RuntimeError(f"An error occurred with exit code {code}: {msg}")


def _warn(prog: str, msg: str) -> None:
    """A warning on stderr in the reference's form (error(WARNING, ...),
    common/rterror.c), which the reference's wrappers let through."""
    sys.stderr.write(f"{prog}: warning - {msg}\n")


def _scene(octree: Path | str) -> "_lib.Context":
    """One committed context per scene file, reused while the file is unchanged."""
    return _scene_files([octree])


def _scene_files(paths: Sequence[Path | str]) -> "_lib.Context":
    """One committed context per ordered list of octree / scene files (loaded
    in order, as oconv reads its arguments), reused while none of them
    changes."""
    key = []
    for f in paths:
        p = Path(f)
        try:
            st = p.stat()
        except OSError:
            raise _rterr(f"cannot open octree file \"{f}\"" if len(paths) == 1 else f"cannot open scene file \"{f}\"")
        key.append((str(p.resolve()), st.st_mtime_ns, st.st_size))
    name = "\0".join(k[0] for k in key)
    ctx = _contexts.get(name)
    if ctx is not None and ctx[0] == key:
        return ctx[1]
    c = _lib.Context(0)
    try:
        for f in paths:
            c.load(str(f))
        c.commit()
    except RuntimeError as e:
        c.close()
        raise _rterr(str(e))
    _contexts[name] = (key, c)
    return c


def _read_rays(rays: bytes, inform: str) -> np.ndarray:
    if inform == "a":
        vals = np.array(rays.split(), dtype=np.float64)
    elif inform == "f":
        vals = np.frombuffer(rays, dtype=np.float32).astype(np.float64)
    elif inform == "d":
        vals = np.frombuffer(rays, dtype=np.float64)
    else:
        raise _rterr(f"unsupported input format -f{inform}")
    if vals.size % 6:
        raise _rterr("unexpected EOF on input")
    return vals.reshape(-1, 6)


def _octree_header(octree: Path | str) -> list[str]:
    """Header lines of the scene file (checkheader() copy, common/header.c)."""
    try:
        with open(octree, "rb") as f:
            head = f.read(65536)
    except OSError:
        return []
    if not head.startswith(b"#?RADIANCE"):
        return []
    end = head.find(b"\n\n")
    lines = head[:end].decode(errors="replace").split("\n")[1:] if end > 0 else []
    return [ln for ln in lines if ln and not ln.startswith("FORMAT=")]


def _header(prog: str, argv: list[str], octree, outform: str, extra: str = "",
            scene_lines: list[str] | None = None) -> bytes:
    """Output header as rtrace / rcontrib write it (rt/rtmain.c:418-424,
    rt/rc2.c:97-125): the scene's own header lines (or `scene_lines`), the
    command without the octree argument, SOFTWARE, CAPDATE/GMT, NCOMP, `extra`,
    byte order for binary output, FORMAT."""
    fmt = {"a": "ascii", "f": "float", "d": "double", "c": "32-bit_rle_rgbe"}[outform]
    now = time.time()
    lines = ["#?RADIANCE", *(_octree_header(octree) if scene_lines is None else scene_lines),
             " ".join([prog, *argv]), f"SOFTWARE= {VERSION}",
             time.strftime("CAPDATE= %Y:%m:%d %H:%M:%S", time.localtime(now)),
             time.strftime("GMT= %Y:%m:%d %H:%M:%S", time.gmtime(now)), "NCOMP=3"]
    if extra:
        lines.append(extra.rstrip("\n"))
    if outform in "fd":
        lines.append(f"BigEndian={int(np.little_endian is False)}")
    lines.append(f"FORMAT={fmt}")
    return ("\n".join(lines) + "\n\n").encode()


def _format(values: np.ndarray, outform: str, fmt: str, per_line: int) -> bytes:
    """Records of `per_line` values: ASCII with `fmt` + TAB, or binary."""
    if outform == "a":
        if values.size == 0:
            return b""
        return _lib.format_ascii(values, per_line, fmt, "\t", "\t")
    if outform == "f":
        return values.astype("<f4").tobytes()
    if outform == "d":
        return values.astype("<f8").tobytes()
    raise _rterr(f"unsupported output format -f?{outform}")


def _seed(uncorrelated: bool) -> int:
    return int(time.time_ns() & 0xFFFFFFFF) if uncorrelated else 0


def rtrace(
    rays: bytes,
    octree: Path | str,
    header: bool = True,
    inform: str = "a",
    outform: str = "a",
    irradiance: bool = False,
    irradiance_lambertian: bool = False,
    outspec: None | str = None,
    trace_exclude: str = "",
    trace_include: str = "",
    trace_exclude_file: None | str | Path = None,
    trace_include_file: None | str | Path = None,
    uncorrelated: bool = False,
    xres: None | int = None,
    yres: None | int = None,
    nproc: None | int = None,
    params: None | Sequence[str] = None,
    report: bool = False,
    version: bool = False,
) -> bytes:
    """rtrace on the GPU (reference: src/pyradiance/rt.py:278-362).

    The result bytes are laid out as bin/rtrace writes them: a header, then
    one "%e\\t%e\\t%e\\t\\n" line (or 3 float/double values) per input ray.
    Only the value output (-ov) is produced on the GPU path.
    """
    if version:
        return (VERSION + "\n").encode()
    if not isinstance(rays, bytes):
        raise TypeError("Rays must be bytes")
    argv: list[str] = []
    if not header:
        argv.append("-h")
    if irradiance:
        argv.append("-I")
    elif irradiance_lambertian:
        argv.append("-i")
    argv.append(f"-f{inform}{outform}")
    if outspec:
        argv.append(f"-o{outspec}")
    if trace_exclude or trace_include or trace_exclude_file or trace_include_file:
        raise _rterr("trace include/exclude lists are not supported on the GPU path")
    if uncorrelated:
        argv.append("-u+")
    if params is not None:
        argv.extend(params)
    try:
        opt = _params.parse(argv, rcontrib=False)
    except ValueError as e:
        raise _rterr(str(e))
    if opt.outspec != "v":
        raise _rterr(f"output specification -o{opt.outspec} is not supported on the GPU path (only -ov)")
    ctx = _scene(octree)
    opt.params.seed = _seed(opt.uncorrelated)
    vals = _read_rays(rays, opt.inform)
    flags = (_lib.RT_IMM_IRRAD if opt.imm_irrad else 0) | (_lib.RT_LAMB_IRRAD if opt.lamb_irrad else 0)
    try:
        ctx.set_params(opt.params)
        dirs = np.linalg.norm(vals[:, 3:], axis=1) if len(vals) else np.zeros(0)
        good = dirs > 0
        out = np.zeros((len(vals), 3))
        if good.any():
            out[good] = ctx.trace(vals[good], flags)
    except RuntimeError as e:
        raise _rterr(str(e))
    body = _format(out, opt.outform, "%e", 3)
    if not opt.header:
        return body
    return _header("rtrace", argv, octree, opt.outform) + body


def _parse_assign(s: str | None) -> dict:
    out = {}
    if not s:
        return out
    for part in re.split(r"[,;]", s):
        part = part.strip()
        if not part:
            continue
        k, _, v = part.partition("=")
        out[k.strip()] = float(v)
    return out


_KLEMS = {"klems_full.cal": (2, "kbin", 145, "Nkbins"), "klems_half.cal": (3, "khbin", 77, "Nkhbins"),
          "klems_quarter.cal": (4, "kqbin", 41, "Nkqbins")}
_KDIR = {"N": ((0, -1, 0), (0, 0, 1)), "E": ((-1, 0, 0), (0, 0, 1)), "S": ((0, 1, 0), (0, 0, 1)),
         "W": ((1, 0, 0), (0, 0, 1)), "D": ((0, 0, -1), (0, 1, 0))}


_FLT = r"([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)"
# rfluxmtx's uniform surface receiver bin (util/rfluxmtx.c:465-470)
_HALFSPACE = re.compile(r"if\(-Dx\*" + _FLT + r"-Dy\*" + _FLT + r"-Dz\*" + _FLT + r",0,-1\)")


def _oriented(m, vars_: dict, defaults: tuple | None, cal: str) -> None:
    """rNx..rNz, Ux..Uz of a binning cal file; without defaults they must be set."""
    keys = ("rNx", "rNy", "rNz", "Ux", "Uy", "Uz")
    if defaults is None:
        miss = [k for k in keys if k not in vars_]
        if miss:
            raise ValueError(f"{cal}: undefined variable {miss[0]}")
        defaults = (0.0,) * 6
    v = [vars_.get(k, d) for k, d in zip(keys, defaults)]
    m.nrm[:], m.up[:] = v[:3], v[3:]


def _modifier_spec(name: str, calfile, expression, nbins, binv, param) -> "_lib.RgModifier":
    """Map a modifier's -f/-e/-p/-bn/-b arguments onto a GPU bin function."""
    m = _lib.RgModifier()
    m.name = name.encode()
    m.rhs = 1.0
    vars_ = _parse_assign(param)
    vars_.update(_parse_assign(expression))
    cal = Path(str(calfile)).name if calfile else None
    if "RHS" in vars_:
        m.rhs = vars_["RHS"]
    if cal == "reinhartb.cal":
        mf = vars_.get("MF", 1.0)
        m.bin_kind, m.mf = 1, mf
        m.nrm[:] = (vars_.get("rNx", 0.0), vars_.get("rNy", 0.0), vars_.get("rNz", -1.0))
        m.up[:] = (vars_.get("Ux", 0.0), vars_.get("Uy", 1.0), vars_.get("Uz", 0.0))
        if binv not in (None, "rbin"):
            raise ValueError(f"bin expression {binv!r} is not supported with reinhartb.cal")
        n = reinhart_nbins(int(mf))
        m.nbins = n if nbins in (None, "Nrbins") else int(float(nbins))
        return m
    if cal in _KLEMS:
        kind, fname, n, nname = _KLEMS[cal]
        m.bin_kind = kind
        b = (binv or "").replace(" ", "")
        mm = re.fullmatch(fname + r"([NESWD])", b)
        if mm:
            m.nrm[:], m.up[:] = _KDIR[mm.group(1)]
        else:
            mm = re.fullmatch(fname + r"\(([^)]*)\)", b)
            if not mm:
                raise ValueError(f"bin expression {binv!r} is not supported with {cal}")
            v = [float(x) for x in mm.group(1).split(",")]
            if len(v) != 6:
                raise ValueError(f"bad {fname} arguments")
            m.nrm[:], m.up[:] = v[:3], v[3:]
        m.nbins = n if nbins in (None, nname) else int(float(nbins))
        return m
    if cal == "disk2square.cal":   # Shirley-Chiu square bins, util/disk2square.cal:67-90
        if binv not in (None, "scbin"):
            raise ValueError(f"bin expression {binv!r} is not supported with {cal}")
        if "SCdim" not in vars_:
            raise ValueError(f"{cal}: undefined variable SCdim")
        m.bin_kind, m.mf = 6, vars_["SCdim"]
        _oriented(m, vars_, None, cal)
        n = int(m.mf) * int(m.mf)
        m.nbins = n if nbins in (None, "SCdim*SCdim") else int(float(nbins))
        return m
    if cal == "cieskyscan.cal":   # CIE sky-scanner patches, util/cieskyscan.cal:21-53
        if binv not in (None, "cbin"):
            raise ValueError(f"bin expression {binv!r} is not supported with {cal}")
        m.bin_kind = 7
        _oriented(m, vars_, (0.0, 0.0, -1.0, 0.0, 1.0, 0.0), cal)
        m.nbins = 145 if nbins in (None, "Ncbins") else int(float(nbins))
        return m
    if cal is not None:
        raise ValueError(f"bin file {cal} is not supported on the GPU path")
    m.nbins = int(float(nbins)) if nbins is not None else 1
    b = (binv or "0").replace(" ", "")
    mm = _HALFSPACE.fullmatch(b)
    if mm:
        m.bin_kind = 5
        m.nrm[:] = [float(x) for x in mm.groups()]
        return m
    m.bin_kind = 0
    try:
        m.cbin = int(float(b))
    except ValueError:
        raise ValueError(f"bin expression {binv!r} is not supported on the GPU path")
    return m


class Rcontrib:
    """rcontrib on the GPU (reference: src/pyradiance/rt.py:165-235)."""

    def __init__(
        self,
        inp: bytes,
        octree: Path | str,
        nproc: int = 1,
        yres: None | int = None,
        inform: None | str = None,
        outform: None | str = None,
        report: int = 0,
        params: None | Sequence[str] = None,
    ):
        self.octree = octree
        self.inp = inp
        self.argv: list[str] = ["-n", str(nproc)]
        if params is not None:
            self.argv.extend(params)
        if None not in (inform, outform):
            self.argv.append(f"-f{inform}{outform}")
        if yres is not None:
            self.argv.extend(["-y", str(yres)])
        if report:
            self.argv.extend(["-t", str(report)])
        self.modifiers: list = []

    def add_modifier(
        self,
        modifier: None | str = None,
        modifier_path: None | str = None,
        calfile: None | str = None,
        expression: None | str = None,
        nbins: None | str = None,
        binv: None | str = None,
        param: None | str = None,
        xres: None | int = None,
        yres: None | int = None,
        output: None | str = None,
    ):
        if modifier is not None:
            names = [modifier]
        elif modifier_path is not None:
            names = Path(modifier_path).read_text().split()
        else:
            raise ValueError("Modifier or modifier path must be provided.")
        if output is not None:
            raise _rterr("per-modifier output files (-o) are not supported on the GPU path")
        for nm in names:
            self.modifiers.append((nm, calfile, expression, nbins, binv, param))
        return self

    def matrix(self) -> np.ndarray:
        """Contribution matrix (records x 3*sum(nbins)) as float64, no formatting."""
        return self._run()[0]

    def _run(self):
        return _contrib_run(self.argv, self.modifiers, self.inp, lambda: _scene(self.octree))

    def __call__(self) -> bytes:
        out, opt, mods = self._run()
        return _contrib_bytes(out, opt, mods, "rcontrib", self.argv, self.octree)


def _contrib_run(argv: list[str], modifiers: list, inp, scene):
    """rcontrib's record loop on the GPU: parse `argv`, map each modifier
    tuple (name, calfile, expression, nbins, binv, param) onto a bin function,
    read the rays and accumulate -c records.  `scene()` returns the committed
    context.  Returns (matrix, options, modifier specs).

    Records follow rcontrib() and done_contrib() (rt/rcontrib.c:358-427):
    with -c N (N >= 1) every input ray, a zero-direction one included, is
    one of its record's N rays and contributes nothing when its direction is
    zero; a final record with k < N rays is written as the mean of those k,
    with the reference's warning; with -c 0 zero-direction rays are ignored
    (with its warning) and all the others sum into one record."""
    try:
        opt = _params.parse(argv, rcontrib=True)
        mods = [_modifier_spec(*m) for m in modifiers]
    except ValueError as e:
        raise _rterr(str(e))
    if not mods:
        raise _rterr("missing required modifier argument")
    ctx = scene()
    opt.params.seed = _seed(opt.uncorrelated)
    inp = inp if isinstance(inp, bytes) else bytes(inp)
    vals = _read_rays(inp, opt.inform)
    ncols = 3 * sum(m.nbins for m in mods)
    flags = (_lib.RT_IMM_IRRAD if opt.imm_irrad else 0) | (_lib.RT_LAMB_IRRAD if opt.lamb_irrad else 0)
    good = (np.linalg.norm(vals[:, 3:], axis=1) > 0) if len(vals) else np.zeros(0, bool)
    acc = opt.accumulate

    def trace(rays, accum):
        try:
            ctx.set_params(opt.params)
            return ctx.contrib(rays, mods, accum, flags)
        except RuntimeError as e:
            raise _rterr(str(e))

    if acc <= 0:
        if not good.all():
            _warn("rcontrib", "dummy ray(s) ignored during accumulation")
        rays = vals[good]
        if len(rays) == 0:
            return np.zeros((1, ncols)), opt, mods
        # one record: the sum of every ray (no division, rt/rc2.c:301-302)
        return trace(rays, len(rays)) * len(rays), opt, mods
    n = len(vals)
    nrec = -(-n // acc)
    if nrec == 0:
        return np.zeros((0, ncols)), opt, mods
    partial = n % acc
    if partial:
        _warn("rcontrib", "partial accumulation in final record")
    if good.all() and not partial:
        return trace(vals, acc), opt, mods
    # per-ray values, then the records on the host: zero-direction rays add
    # nothing, the last record is the mean of the rays it has
    per = np.zeros((n, ncols))
    if good.any():
        per[good] = trace(vals[good], 1)
    out = np.zeros((nrec, ncols))
    full = (n // acc) * acc
    if full:
        out[:n // acc] = per[:full].reshape(n // acc, acc, ncols).mean(1)
    if partial:
        out[-1] = per[full:].mean(0)
    return out, opt, mods


def _contrib_bytes(out: np.ndarray, opt, mods: list, prog: str, argv: list[str], octree,
                   scene_lines: list[str] | None = None) -> bytes:
    """Formatted rcontrib output to stdout with its header (rt/rc2.c:150-170):
    NROWS when -y is set, NCOLS = total bins over the modifiers."""
    body = _format(out.ravel(), opt.outform, "%.6e", out.shape[1] if out.size else 3)
    if not opt.header:
        return body
    reclen = sum(m.nbins for m in mods)
    info = []
    if opt.yres > 0:
        info.append(f"NROWS={opt.yres * max(opt.xres, 1)}")
    if opt.xres <= 0 or reclen > 1:
        info.append(f"NCOLS={reclen}")
    return _header(prog, argv, octree, opt.outform, "\n".join(info), scene_lines) + body


def rpict(
    view: Sequence[str],
    octree: Path | str,
    xres: None | int = None,
    yres: None | int = None,
    report: float = 0,
    report_file: None | Path = None,
    params: None | Sequence[str] = None,
) -> bytes:
    """rpict on the GPU (reference wrapper: src/pyradiance/rt.py:238-275).

    One primary ray per pixel, generated on the GPU from the view
    (rt/rpict.c:668-712 with -pj jitter), shaded as rtrace does, written as a
    Radiance picture.  Every pixel is sampled: -ps / -pt (adaptive sampling
    and interpolation) are accepted and have no effect, as with -ps 1."""
    from . import view as _view
    argv = list(view)
    if report:
        argv += ["-t", str(report)]
    if report_file:
        argv += ["-e", str(report_file)]
    if xres is not None:
        argv += ["-x", str(xres)]
    if yres is not None:
        argv += ["-y", str(yres)]
    if params:
        argv += list(params)
    v, rest = _view.parse_view(argv)
    xr, yr, pa, pj = 512, 512, 1.0, 0.67
    ropts = []
    i = 0
    try:
        while i < len(rest):
            a = rest[i]
            if a in ("-x", "-y", "-pa", "-pj", "-ps", "-pt", "-pd", "-pm", "-t", "-e"):
                if i + 1 >= len(rest):
                    raise ValueError(f"missing argument for {a}")
                val = rest[i + 1]
                if a == "-x":
                    xr = int(val)
                elif a == "-y":
                    yr = int(val)
                elif a == "-pa":
                    pa = float(val)
                elif a == "-pj":
                    pj = float(val)
                elif a in ("-pd", "-pm") and float(val) > 1e-6:
                    raise ValueError(f"{a}: depth of field and motion blur are not supported on the GPU path")
                i += 2
                continue
            if a in ("-o", "-z", "-r", "-ro", "-S", "-P", "-PP", "-i"):
                raise ValueError(f"rpict option {a} is not supported on the GPU path")
            ropts.append(a)
            i += 1
        if xr <= 0 or yr <= 0:
            raise ValueError("bad resolution")
        opt = _params.parse(ropts, rcontrib=False)
        # rpict's own defaults (rt/rpict.c:60-106) where they differ from rtrace's
        given = set(ropts)
        p = opt.params
        if "-ds" not in given:
            p.srcsizerat = 0.25
        if "-lr" not in given:
            p.maxdepth = 7
        if "-lw" not in given:
            p.minweight = 1e-4
        if "-ad" not in given:
            p.ambdiv = 512
        if "-dj" not in given:
            p.dstrsrc = 0.0
        v.setup()
    except ValueError as e:
        raise _rterr(str(e))
    pa, xr, yr = _view.normaspect(v.aspect(), pa, xr, yr)
    ctx = _scene(octree)
    opt.params.seed = _seed(opt.uncorrelated)
    flags = (_lib.RT_IMM_IRRAD if opt.imm_irrad else 0) | (_lib.RT_LAMB_IRRAD if opt.lamb_irrad else 0)
    try:
        ctx.set_params(opt.params)
        img = ctx.trace_view(v, xr, yr, pj, flags)
    except RuntimeError as e:
        raise _rterr(str(e))
    now = time.time()
    head = [*_octree_header(octree), " ".join(["rpict", *argv]), f"SOFTWARE= {VERSION}", "VIEW=" + v.header(),
            time.strftime("CAPDATE= %Y:%m:%d %H:%M:%S", time.localtime(now)),
            time.strftime("GMT= %Y:%m:%d %H:%M:%S", time.gmtime(now))]
    if pa < .99 or pa > 1.01:
        head.append("PIXASPECT=%f" % pa)
    head.append("PRIMARIES= 0.6400 0.3300 0.2900 0.6000 0.1500 0.0600 0.3333 0.3333")
    return _view.write_picture(img, head)
