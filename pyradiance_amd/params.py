"""Render-option parsing with the reference's semantics (rt/renderopts.c:124-350).

Maps an rtrace / rcontrib argument list onto the GPU path's rg_params.
Options the GPU path honours are applied; options that only tune features
the GPU path does not run (irradiance cache, photon maps, shadow-test
thresholds) are accepted and ignored, as documented per option below.
Anything unknown raises ValueError instead of being silently dropped.
"""

from __future__ import annotations

from dataclasses import dataclass, field

from . import _lib

# options taking N numeric arguments that the GPU path accepts but does not use:
#   -aa -ar -as -aw: irradiance cache (forced off: -aa 0, -as 0)
#   -dt -dc -dr -dp: shadow threshold / certainty / virtual relays (every source is tested)
#   -ps -pt -pj -pa -pm -pd: image sampling (rpict only)
#   -ms -me -ma -mg: participating media (not part of the GPU material set)
_IGNORED = {"-aa": 1, "-ar": 1, "-as": 1, "-aw": 1, "-dt": 1, "-dc": 1, "-dr": 1, "-dp": 1,
            "-ps": 1, "-pt": 1, "-pj": 1, "-pa": 1, "-pm": 1, "-pd": 1, "-ms": 1,
            "-me": 3, "-ma": 3, "-mg": 1, "-t": 1, "-n": 1, "-lt": 0}


def _bool_opt(arg: str, cur: bool) -> bool:
    if arg.endswith("+"):
        return True
    if arg.endswith("-"):
        return False
    return not cur


@dataclass
class Options:
    params: _lib.RgParams
    imm_irrad: bool = False      # -I
    lamb_irrad: bool = False     # -i
    accumulate: int = 1          # rcontrib -c
    contrib: bool = False        # rcontrib -V
    header: bool = True          # -h
    inform: str = "a"
    outform: str = "a"
    uncorrelated: bool = False   # -u
    outspec: str = "v"           # rtrace -o
    xres: int = 0
    yres: int = 0
    ignored: list = field(default_factory=list)


def parse(argv: list[str], rcontrib: bool = False) -> Options:
    """Parse render options; starts from the defaults of the emulated program."""
    p = _lib.default_params(rcontrib)
    o = Options(params=p)
    i = 0
    n = len(argv)

    def nums(k):
        nonlocal i
        if i + k >= n:
            raise ValueError(f"missing argument for {argv[i]}")
        vals = [float(v) for v in argv[i + 1:i + 1 + k]]
        i += k
        return vals

    while i < n:
        a = argv[i]
        if a in ("-ab",):
            p.ambounce = int(nums(1)[0])
        elif a == "-ad":
            p.ambdiv = int(nums(1)[0])
        elif a == "-av":
            p.ambval[:] = nums(3)
        elif a == "-lw":
            p.minweight = nums(1)[0]
        elif a == "-lr":
            p.maxdepth = int(nums(1)[0])
        elif a == "-dj":
            p.dstrsrc = nums(1)[0]
        elif a == "-ds":
            p.srcsizerat = nums(1)[0]
        elif a == "-st":
            p.specthresh = nums(1)[0]
        elif a == "-ss":
            p.specjitter = nums(1)[0]
        elif a.startswith("-dv"):
            p.directvis = int(_bool_opt(a, bool(p.directvis)))
        elif a.startswith("-bv"):
            p.backvis = int(_bool_opt(a, bool(p.backvis)))
        elif a.startswith("-u"):
            o.uncorrelated = _bool_opt(a, o.uncorrelated)
        elif a.startswith("-I"):
            o.imm_irrad = _bool_opt(a, o.imm_irrad)
        elif a.startswith("-i") and a[2:] in ("", "+", "-"):
            o.lamb_irrad = _bool_opt(a, o.lamb_irrad)
        elif a.startswith("-V"):
            o.contrib = _bool_opt(a, o.contrib)
            p.contrib = int(o.contrib)
        elif a.startswith("-h") and a[2:] in ("", "+", "-"):
            o.header = _bool_opt(a, o.header)
        elif a == "-c":
            o.accumulate = int(nums(1)[0])
        elif a == "-x":
            o.xres = int(nums(1)[0])
        elif a == "-y":
            o.yres = int(nums(1)[0])
        elif a.startswith("-f") and len(a) == 4 and a[2] in "afdc" and a[3] in "afdc":
            o.inform, o.outform = a[2], a[3]
        elif a.startswith("-f") and len(a) == 3 and a[2] in "afd":
            o.inform = o.outform = a[2]          # one letter: same format both ways (rt/rcmain.c:141-144)
        elif rcontrib and a in ("-fo", "-fo+", "-fo-"):
            o.ignored.append(a)                  # force overwrite of -o files; output goes to stdout only
        elif a.startswith("-w") and a[2:] in ("", "+", "-"):
            o.ignored.append(a)                  # warning messages on/off
        elif a in ("-ld", "-ld+", "-ld-"):
            if a != "-ld-":
                raise ValueError(f"{a}: limited ray distance is not supported on the GPU path")
        elif a.startswith("-o") and len(a) > 2 and not rcontrib:
            o.outspec = a[2:]
        elif a in _IGNORED:
            k = _IGNORED[a]
            o.ignored.append(" ".join(argv[i:i + 1 + k]))
            i += k
        elif a in ("-ae", "-ai", "-aE", "-aI", "-af"):
            raise ValueError(f"{a}: ambient include/exclude lists and ambient files are not supported on the GPU path")
        else:
            raise ValueError(f"unsupported option {a!r} on the GPU path")
        i += 1
    if o.accumulate < 0:
        raise ValueError("bad -c value")
    return o
