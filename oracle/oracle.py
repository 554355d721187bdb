"""ctypes wrapper of oracle/_ref/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker of the GPU product.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE / "_ref"
LIB = REF / "liboracle.so"
REFSRC = Path(os.environ.get("RADIANCE_REF_SRC", "/root/reference/src/radiance"))


class Params(C.Structure):
    _fields_ = [
        ("ambounce", C.c_int), ("ambdiv", C.c_int), ("maxdepth", C.c_int), ("backvis", C.c_int),
        ("minweight", C.c_double), ("dstrsrc", C.c_double), ("srcsizerat", C.c_double),
        ("specthresh", C.c_double), ("specjitter", C.c_double), ("ambval", C.c_double * 3),
        ("directvis", C.c_int), ("contrib", C.c_int), ("seed", C.c_ulonglong),
    ]


class Mod(C.Structure):
    _fields_ = [
        ("name", C.c_char_p), ("bin_kind", C.c_int), ("nbins", C.c_int), ("cbin", C.c_int),
        ("mf", C.c_double), ("rhs", C.c_double), ("nrm", C.c_double * 3), ("up", C.c_double * 3),
    ]


def build(quiet: bool = True) -> Path:
    """Compile liboracle.so from oracle/radoracle.c (no reference sources needed)."""
    r = subprocess.run(["make", "-C", str(HERE), "oracle"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    return LIB


def build_reference(jobs: int = 8) -> bool:
    """Compile the reference tools into oracle/_ref when the reference tree is present."""
    if not (REFSRC / "rt" / "raytrace.c").exists():
        return False
    r = subprocess.run(["make", "-C", str(HERE), f"-j{jobs}", "ref", f"REF={REFSRC}"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-4000:] + r.stderr[-4000:])
    return True


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        dp = C.POINTER(C.c_double)
        L.ro_scene_parse.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.ro_scene_parse.restype = C.c_void_p
        L.ro_scene_free.argtypes = [C.c_void_p]
        L.ro_scene_counts.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.ro_closest.argtypes = [C.c_void_p, dp, C.c_long, C.POINTER(C.c_int), dp]
        L.ro_trace.argtypes = [C.c_void_p, C.POINTER(Params), dp, C.c_long, C.c_int, dp,
                               C.POINTER(C.c_ulonglong), C.c_char_p, C.c_int]
        L.ro_contrib.argtypes = [C.c_void_p, C.POINTER(Params), dp, C.c_long, C.c_int, C.c_int,
                                 C.POINTER(Mod), C.c_int, dp, C.POINTER(C.c_ulonglong), C.c_char_p, C.c_int]
        L.ro_bins.argtypes = [C.POINTER(Mod), dp, C.c_long, C.POINTER(C.c_int)]
        L.ro_child_key.argtypes = [C.c_ulonglong, C.c_ulonglong]
        L.ro_child_key.restype = C.c_ulonglong
        L.ro_sensor_key.argtypes = [C.c_ulonglong, C.c_ulonglong]
        L.ro_sensor_key.restype = C.c_ulonglong
        L.ro_rnd.argtypes = [C.c_ulonglong, C.c_uint]
        L.ro_rnd.restype = C.c_double
        _lib = L
    return _lib


def params(**kw) -> Params:
    """Library defaults of rt/raycalls.c:118-153, overridden by keywords."""
    p = Params(ambounce=0, ambdiv=1024, maxdepth=-10, backvis=1, minweight=1e-4, dstrsrc=0.0,
               srcsizerat=0.2, specthresh=0.15, specjitter=1.0, directvis=1, contrib=0, seed=0)
    for k, v in kw.items():
        if k == "ambval":
            p.ambval[:] = v
        else:
            setattr(p, k, v)
    return p


def mod(name: str, bin_kind: int = 0, nbins: int = 1, cbin: int = 0, mf: float = 1.0, rhs: float = 1.0,
        nrm=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0)) -> Mod:
    m = Mod()
    m.name = name.encode()
    m.bin_kind, m.nbins, m.cbin, m.mf, m.rhs = bin_kind, nbins, cbin, mf, rhs
    m.nrm[:] = nrm
    m.up[:] = up
    return m


class Scene:
    def __init__(self, text: str):
        L = lib()
        err = C.create_string_buffer(512)
        self._h = L.ro_scene_parse(text.encode(), err, 512)
        if not self._h:
            raise RuntimeError(err.value.decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ro_scene_free(self._h)
            self._h = None

    def counts(self):
        v = (C.c_longlong * 4)()
        lib().ro_scene_counts(self._h, v)
        return dict(zip(("objects", "primitives", "materials", "sources"), list(v)))

    def closest(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        prim = np.empty(len(rays), dtype=np.int32)
        t = np.empty(len(rays))
        lib().ro_closest(self._h, rays.ctypes.data_as(C.POINTER(C.c_double)), len(rays),
                         prim.ctypes.data_as(C.POINTER(C.c_int)), t.ctypes.data_as(C.POINTER(C.c_double)))
        return prim, t

    def trace(self, rays: np.ndarray, p: Params, flags: int = 0):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        out = np.zeros((len(rays), 3))
        nr = C.c_ulonglong()
        err = C.create_string_buffer(512)
        if lib().ro_trace(self._h, C.byref(p), rays.ctypes.data_as(C.POINTER(C.c_double)), len(rays), flags,
                          out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nr), err, 512) != 0:
            raise RuntimeError(err.value.decode())
        return out, int(nr.value)

    def contrib(self, rays: np.ndarray, p: Params, mods: list, accum: int = 1, flags: int = 0):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        nrec = len(rays) // accum
        ncol = 3 * sum(m.nbins for m in mods)
        out = np.zeros((nrec, ncol))
        arr = (Mod * len(mods))(*mods)
        nr = C.c_ulonglong()
        err = C.create_string_buffer(512)
        if lib().ro_contrib(self._h, C.byref(p), rays.ctypes.data_as(C.POINTER(C.c_double)), nrec, accum, flags,
                            arr, len(mods), out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nr), err, 512) != 0:
            raise RuntimeError(err.value.decode())
        return out, int(nr.value)


def bins(m: Mod, dirs: np.ndarray) -> np.ndarray:
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(dirs), dtype=np.int32)
    lib().ro_bins(C.byref(m), dirs.ctypes.data_as(C.POINTER(C.c_double)), len(dirs),
                  out.ctypes.data_as(C.POINTER(C.c_int)))
    return out
