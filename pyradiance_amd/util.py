"""dctimestep with the pyradiance signature (src/pyradiance/util.py:140-192),
computing the DC x sky product on the GPU (MFMA, BF16 hi/lo split).

Matrix files use Radiance's matrix format: a header with NROWS / NCOLS /
NCOMP / FORMAT lines, a blank line, then ascii or binary data in row-major
[row][col][component] order (util/cmatrix.c:421-545).
"""

from __future__ import annotations

import re
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib


def _rterr(msg: str) -> RuntimeError:
    return RuntimeError(f"An error occurred with exit code 1: {msg}")


def read_matrix(src: str | Path | bytes) -> np.ndarray:
    """Radiance matrix (header + ascii/float/double data) -> (nrows, ncols, 3) float32."""
    data = src if isinstance(src, bytes) else Path(src).read_bytes()
    nrows = ncols = None
    ncomp = 3
    fmt = "ascii"
    big = sys.byteorder == "big"
    body = data
    if data.startswith(b"#?"):
        end = data.find(b"\n\n")
        if end < 0:
            raise _rterr("bad matrix header")
        for line in data[:end].decode(errors="replace").split("\n"):
            m = re.match(r"\s*(NROWS|NCOLS|NCOMP|FORMAT|BigEndian)=\s*(\S+)", line)
            if not m:
                continue
            k, v = m.groups()
            if k == "NROWS":
                nrows = int(v)
            elif k == "NCOLS":
                ncols = int(v)
            elif k == "NCOMP":
                ncomp = int(v)
            elif k == "FORMAT":
                fmt = v
            else:
                big = v != "0"
        body = data[end + 2:]
    if ncomp != 3:
        raise _rterr(f"NCOMP={ncomp} is not supported (RGB matrices only)")
    if fmt == "ascii":
        vals = np.array(body.split(), dtype=np.float64)
    elif fmt == "float":
        vals = np.frombuffer(body, dtype=">f4" if big else "<f4").astype(np.float64)
    elif fmt == "double":
        vals = np.frombuffer(body, dtype=">f8" if big else "<f8")
    else:
        raise _rterr(f"matrix format {fmt} is not supported")
    if nrows is None or ncols is None:
        if ncols is None and nrows is not None:
            ncols = vals.size // (3 * nrows)
        elif nrows is None and ncols is not None:
            nrows = vals.size // (3 * ncols)
        else:
            raise _rterr("matrix without NROWS/NCOLS")
    need = nrows * ncols * 3
    if vals.size < need:
        raise _rterr("unexpected EOF reading matrix")
    return vals[:need].reshape(nrows, ncols, 3).astype(np.float32)


def write_matrix(m: np.ndarray, outform: str = "a", header: bool = True, argv: list[str] | None = None) -> bytes:
    """(nrows, ncols, 3) -> Radiance matrix bytes as cm_write() lays them out."""
    m = np.asarray(m)
    nrows, ncols = m.shape[:2]
    out = []
    if header:
        fmt = {"a": "ascii", "f": "float", "d": "double"}[outform]
        lines = ["#?RADIANCE", " ".join(argv or ["dctimestep"]), time.strftime("CAPDATE= %Y:%m:%d %H:%M:%S"),
                 f"NROWS={nrows}", f"NCOLS={ncols}", "NCOMP=3"]
        if outform in "fd":
            lines.append(f"BigEndian={int(sys.byteorder == 'big')}")
        lines.append(f"FORMAT={fmt}")
        out.append(("\n".join(lines) + "\n\n").encode())
    if outform == "a":
        flat = m.reshape(nrows, ncols * 3)
        row = "\t".join(["%.6e %.6e %.6e"] * ncols) + "\n"
        out.append("".join(row % tuple(r) for r in flat).encode())
    elif outform == "f":
        out.append(m.astype(np.float32).tobytes())
    elif outform == "d":
        out.append(m.astype(np.float64).tobytes())
    else:
        raise _rterr(f"output format {outform} is not supported")
    return b"".join(out)


def dcsky(dc: np.ndarray, sky: np.ndarray, mode: int = 1) -> np.ndarray:
    """C = DC x sky per channel on the GPU: (M,K,3) x (K,N,3) -> (M,N,3) float32.

    mode 1: BF16 hi/lo split MFMA with FP32 accumulation; mode 0: FP32 FMA.
    """
    return _lib.dcsky(dc, sky, mode)


def dctimestep(
    *mtx: str | bytes,
    nstep: None | int = None,
    header: bool = True,
    xres: None | int = None,
    yres: None | int = None,
    inform: None | str = None,
    outform: None | str = None,
    ospec: None | str = None,
) -> None | bytes:
    """Daylight-coefficient time-step product (reference: util.py:140-192).

    Supports the two-matrix form (DC matrix, sky matrix); the sky matrix may
    be given as bytes (the reference feeds it on stdin).
    """
    if len(mtx) not in (2, 4):
        raise ValueError("mtx must be a list of 2 or 4 items")
    if len(mtx) == 4:
        raise _rterr("the view/BSDF/daylight (three-phase) form is not supported on the GPU path")
    if ospec:
        raise _rterr("per-step output files (-o spec) are not supported on the GPU path")
    dc = read_matrix(mtx[0])
    sky = read_matrix(mtx[1])
    if nstep:
        sky = sky[:, :nstep]
    if dc.shape[1] != sky.shape[0]:
        raise _rterr(f"matrix dimension mismatch: DC has {dc.shape[1]} columns, sky has {sky.shape[0]} rows")
    res = dcsky(dc, sky)
    argv = ["dctimestep"] + (["-n", str(nstep)] if nstep else []) + [str(m) for m in mtx if not isinstance(m, bytes)]
    return write_matrix(res, outform or "a", header, argv)
