"""ASCII record formatting on the host cores (rg_format_ascii): no GPU.

Pinned against the reference gendaymtx's own ASCII output of a matrix whose
binary form the golden fixtures also hold (gen/gendaymtx.c:652-663), and
against CPython's own float formatting (an implementation independent of the
C library's printf) for rcontrib's "%.6e\\t" records (rt/rc2.c:306) and
rtrace's "%e\\t" values (rt/rtrace.c:912-916).
"""

import json
from pathlib import Path

import numpy as np
import pytest

from pyradiance_amd import _lib

GOLD = Path(__file__).parent / "golden"


def _lib_or_skip():
    try:
        _lib.lib()
    except Exception as e:  # library not built on this host
        pytest.skip(f"libradgpu.so not loadable: {e}")


def test_gendaymtx_ascii_matches_reference_text():
    _lib_or_skip()
    meta = json.loads((GOLD / "sky_cases.json").read_text())
    head, body = meta["ascii_wea1_m1"].split("\n\n", 1)
    ncols = int(next(h for h in head.split("\n") if h.startswith("NCOLS="))[6:])
    mtx = np.load(GOLD / "sky.npz")["wea1_m1"].reshape(-1, ncols, 3)
    out = _lib.format_ascii(mtx, 3, "%.3g", " ", "", group=ncols)
    assert out.decode() == body


@pytest.mark.parametrize("fmt", ["%.6e", "%e", "%.3g", "%g", "%.9g"])
def test_records_match_python_formatting(fmt):
    _lib_or_skip()
    rng = np.random.default_rng(7)
    v = rng.lognormal(0, 6, (20_011, 5)) * rng.choice([-1.0, 1.0, 0.0], (20_011, 5))
    v[0, :] = [0.0, -0.0, np.inf, -np.inf, 5e-324]
    v[1, :] = [1e308, 9.9995e-5, 0.5, 1.5, 2.5]
    out = _lib.format_ascii(v, 5, fmt)
    row = (fmt + "\t") * 5 + "\n"
    assert out.decode() == "".join(row % tuple(r) for r in v)


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_groups_and_thread_split(nthreads):
    _lib_or_skip()
    v = np.arange(3 * 50_000, dtype=np.float64) / 7.0
    out = _lib.format_ascii(v, 3, "%.3g", " ", "", group=7, nthreads=nthreads).decode()
    recs = ["%.3g %.3g %.3g\n" % tuple(r) for r in v.reshape(-1, 3)]
    want = "".join(r + ("\n" if (i + 1) % 7 == 0 else "") for i, r in enumerate(recs))
    assert out == want


def test_empty_and_bad_arguments():
    _lib_or_skip()
    assert _lib.format_ascii(np.zeros(0), 3, "%e") == b""
    v = np.ones(6)
    for bad in ["%s", "%d", "%f", "%.18e", "%.3gx", "e", "%%"]:
        with pytest.raises(ValueError):
            _lib.format_ascii(v, 3, bad)
    with pytest.raises(ValueError):
        _lib.format_ascii(np.ones(5), 3, "%e")
