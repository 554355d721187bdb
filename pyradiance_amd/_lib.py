"""ctypes binding of libradgpu.so (declarations in include/radgpu.h).

The library is the product path: there is no CPU fallback.  If the shared
library is missing or no HIP device is present, calls raise RuntimeError.
"""

from __future__ import annotations

import ctypes as C
import importlib.util
import os
from pathlib import Path

import numpy as np

LIBPATH = Path(os.environ.get("RADGPU_LIB", Path(__file__).resolve().parent / "libradgpu.so"))

# symbols declared in include/radgpu.h
EXPORTS = (
    "rg_ctx_create", "rg_ctx_destroy", "rg_last_error", "rg_scene_load", "rg_scene_load_text",
    "rg_scene_commit", "rg_scene_info", "rg_set_params", "rg_default_params",
    "rg_default_params_rcontrib", "rg_trace", "rg_contrib", "rg_contrib_batch", "rg_trace_batch", "rg_last_nrays",
    "rg_set_profiling", "rg_set_bvh_builder", "rg_set_tuning", "rg_last_stats", "rg_dcsky", "rg_dcsky_device",
    "rg_material_index", "rg_material_get", "rg_material_set", "rg_contrib_adjoint", "rg_first_hit",
    "rg_sky_create", "rg_sky_destroy", "rg_sky_error", "rg_sky_shape", "rg_sky_site", "rg_sky_suns",
    "rg_sky_matrix", "rg_sky_matrix_device", "rg_trace_view", "rg_format_ascii", "rg_contrib_device",
    "rg_set_stream", "rg_sky_matrix_device_noalloc",
)

BVH_AUTO, BVH_SAH, BVH_LBVH = 0, 1, 2

RT_IMM_IRRAD = 8      # -I  (RTimmIrrad, rt/RtraceSimulManager.h:86)
RT_LAMB_IRRAD = 16    # -i


class RgParams(C.Structure):
    _fields_ = [
        ("ambounce", C.c_int), ("ambdiv", C.c_int), ("maxdepth", C.c_int), ("backvis", C.c_int),
        ("minweight", C.c_double), ("dstrsrc", C.c_double), ("srcsizerat", C.c_double),
        ("specthresh", C.c_double), ("specjitter", C.c_double), ("ambval", C.c_double * 3),
        ("directvis", C.c_int), ("contrib", C.c_int), ("seed", C.c_ulonglong),
    ]


class RgSkyOpts(C.Structure):
    _fields_ = [
        ("mfactor", C.c_int), ("sky_color", C.c_double * 3), ("ground_color", C.c_double * 3),
        ("sun_color", C.c_double * 3), ("rotation", C.c_double), ("nsuns", C.c_int),
        ("fixed_sun_sa", C.c_double), ("output", C.c_int), ("sun_hours_only", C.c_int), ("average", C.c_int),
    ]


class RgView(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("vp", C.c_double * 3), ("vdir", C.c_double * 3), ("hvec", C.c_double * 3),
        ("vvec", C.c_double * 3), ("hn2", C.c_double), ("vn2", C.c_double), ("horiz", C.c_double),
        ("vert", C.c_double), ("hoff", C.c_double), ("voff", C.c_double), ("vfore", C.c_double),
        ("vaft", C.c_double),
    ]


class RgModifier(C.Structure):
    _fields_ = [
        ("name", C.c_char_p), ("bin_kind", C.c_int), ("nbins", C.c_int), ("cbin", C.c_int),
        ("mf", C.c_double), ("rhs", C.c_double), ("nrm", C.c_double * 3), ("up", C.c_double * 3),
    ]


_lib = None
_hip = None


def _share_hip_runtime() -> None:
    """Load the HIP runtime that PyTorch-ROCm ships, when PyTorch is
    installed, before this library.  Both carry the soname libamdhip64.so.7,
    but torch's libraries ask for "libamdhip64.so": loaded the other way
    round, the process would map two runtimes and the second cannot use the
    device ("No HIP GPUs are available").  Preloaded here, this library binds
    to it by soname and torch's later import finds the same file.  torch
    itself is not imported."""
    global _hip
    if os.environ.get("RADGPU_HIP_RUNTIME", "") == "system":
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    for d in spec.submodule_search_locations:
        rt = Path(d) / "lib" / "libamdhip64.so"
        if rt.exists():
            _hip = C.CDLL(str(rt), mode=C.RTLD_GLOBAL)
            return


def hip_runtime() -> C.CDLL:
    """The HIP runtime this library is bound to (the one shared with torch
    when torch is installed).  Opening libamdhip64 by bare name instead can
    map a second runtime into the process."""
    lib()
    if _hip is not None:
        return _hip
    # already mapped as libradgpu.so's dependency: dlopen returns that object
    return C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL | os.RTLD_NOLOAD)


def lib() -> C.CDLL:
    """Load libradgpu.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise RuntimeError(f"{LIBPATH} is not built (run `python -m pyradiance_amd.build`)")
    _share_hip_runtime()
    L = C.CDLL(str(LIBPATH))
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    L.rg_ctx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.rg_ctx_destroy.argtypes = [vp]
    L.rg_ctx_destroy.restype = None
    L.rg_last_error.argtypes = [vp]
    L.rg_last_error.restype = C.c_char_p
    L.rg_scene_load.argtypes = [vp, C.c_char_p, C.c_int]
    L.rg_scene_load_text.argtypes = [vp, C.c_char_p, C.c_int]
    L.rg_scene_commit.argtypes = [vp]
    L.rg_scene_info.argtypes = [vp, C.POINTER(C.c_longlong), C.c_int]
    L.rg_set_params.argtypes = [vp, C.POINTER(RgParams)]
    L.rg_default_params.argtypes = [C.POINTER(RgParams)]
    L.rg_default_params.restype = None
    L.rg_default_params_rcontrib.argtypes = [C.POINTER(RgParams)]
    L.rg_default_params_rcontrib.restype = None
    L.rg_trace.argtypes = [vp, dp, C.c_longlong, C.c_int, dp]
    L.rg_contrib.argtypes = [vp, dp, C.c_longlong, C.c_int, C.c_int, C.POINTER(RgModifier), C.c_int, dp]
    L.rg_contrib_device.argtypes = [vp, vp, C.c_longlong, C.c_int, C.c_int, C.POINTER(RgModifier), C.c_int, vp]
    L.rg_set_stream.argtypes = [vp, vp]
    L.rg_contrib_batch.argtypes = [C.POINTER(vp), C.c_int, dp, C.POINTER(C.c_longlong), C.c_int, C.c_int,
                                   C.POINTER(RgModifier), C.c_int, dp]
    L.rg_trace_batch.argtypes = [C.POINTER(vp), C.c_int, dp, C.POINTER(C.c_longlong), C.c_int, dp]
    L.rg_last_nrays.argtypes = [vp, C.POINTER(C.c_ulonglong)]
    L.rg_set_profiling.argtypes = [vp, C.c_int]
    L.rg_set_bvh_builder.argtypes = [vp, C.c_int]
    L.rg_set_tuning.argtypes = [vp, C.c_char_p, C.c_longlong]
    L.rg_last_stats.argtypes = [vp, dp, C.c_int]
    L.rg_dcsky.argtypes = [fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.rg_dcsky_device.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.rg_first_hit.argtypes = [vp, dp, C.c_longlong, dp, dp, C.POINTER(C.c_int)]
    L.rg_material_index.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
    L.rg_material_get.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp, C.c_int]
    L.rg_material_set.argtypes = [vp, C.c_int, dp, C.c_int]
    L.rg_contrib_adjoint.argtypes = [vp, dp, C.c_longlong, C.c_int, C.c_int, C.POINTER(RgModifier), C.c_int, dp, dp,
                                     C.POINTER(C.c_int), C.c_int, dp]
    L.rg_trace_view.argtypes = [vp, C.POINTER(RgView), C.c_int, C.c_int, C.c_double, C.c_int, dp]
    L.rg_sky_create.argtypes = [C.POINTER(vp), C.c_char_p, C.POINTER(RgSkyOpts)]
    L.rg_sky_destroy.argtypes = [vp]
    L.rg_sky_destroy.restype = None
    L.rg_sky_error.argtypes = [vp]
    L.rg_sky_error.restype = C.c_char_p
    L.rg_sky_shape.argtypes = [vp, C.POINTER(C.c_int)]
    L.rg_sky_site.argtypes = [vp, dp]
    L.rg_sky_suns.argtypes = [vp, C.POINTER(C.c_int), dp]
    L.rg_sky_matrix.argtypes = [vp, C.POINTER(C.c_float)]
    L.rg_sky_matrix_device.argtypes = [vp, vp]
    L.rg_sky_matrix_device_noalloc.argtypes = [vp, vp]
    L.rg_format_ascii.argtypes = [dp, C.c_longlong, C.c_int, C.c_longlong, C.c_char_p, C.c_char_p, C.c_char_p,
                                  C.c_char_p, C.c_longlong, C.c_int]
    L.rg_format_ascii.restype = C.c_longlong
    _lib = L
    return L


def default_params(rcontrib: bool = False) -> RgParams:
    p = RgParams()
    if rcontrib:
        lib().rg_default_params_rcontrib(C.byref(p))
    else:
        lib().rg_default_params(C.byref(p))
    return p


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def format_ascii(values: np.ndarray, per_record: int, fmt: str, sep: str = "\t", tail: str = "\t",
                 group: int = 0, nthreads: int = 0) -> bytes:
    """ASCII records formatted on the host cores (rg_format_ascii): each of the
    len(values) / per_record records is its values printed with `fmt`, joined
    by `sep`, then `tail` and a newline; a blank line after every `group`
    records."""
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    if nthreads <= 0:
        # the cores this process may run on, not the machine's
        nthreads = min(len(os.sched_getaffinity(0)), 32) if hasattr(os, "sched_getaffinity") else 0
    if v.size % per_record:
        raise ValueError("values do not fill whole records")
    nrec = v.size // per_record
    width = 26 + len(sep.encode())
    cap = v.size * width + nrec * (len(tail.encode()) + 1) + (nrec // group + 1 if group > 0 else 0)
    buf = np.empty(max(cap, 1), dtype=np.uint8)
    n = lib().rg_format_ascii(_dptr(v), nrec, per_record, group, fmt.encode(), sep.encode(), tail.encode(),
                              buf.ctypes.data_as(C.c_char_p), cap, nthreads)
    if n < 0:
        raise ValueError(f"rg_format_ascii failed ({n}) for format {fmt!r}")
    return buf[:n].tobytes()


class Context:
    """One scene on one GPU (an rg_ctx)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        if self._L.rg_ctx_create(C.byref(h), device) != 0:
            raise RuntimeError(self._L.rg_last_error(None).decode())
        self._h = h

    def _check(self, rc: int):
        if rc != 0:
            raise RuntimeError(self._L.rg_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.rg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def load(self, path: str | os.PathLike, allow_exec: bool = False) -> "Context":
        self._check(self._L.rg_scene_load(self._h, str(path).encode(), int(allow_exec)))
        return self

    def load_text(self, text: str | bytes, allow_exec: bool = False) -> "Context":
        data = text if isinstance(text, bytes) else text.encode()
        self._check(self._L.rg_scene_load_text(self._h, data, int(allow_exec)))
        return self

    def commit(self) -> "Context":
        self._check(self._L.rg_scene_commit(self._h))
        return self

    def info(self) -> dict:
        v = (C.c_longlong * 6)()
        self._check(self._L.rg_scene_info(self._h, v, 6))
        keys = ("objects", "primitives", "materials", "sources", "bvh_nodes", "warnings")
        return dict(zip(keys, list(v)))

    def set_params(self, p: RgParams):
        self._check(self._L.rg_set_params(self._h, C.byref(p)))

    def trace(self, rays: np.ndarray, flags: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        out = np.zeros((rays.shape[0], 3))
        self._check(self._L.rg_trace(self._h, _dptr(rays), rays.shape[0], flags, _dptr(out)))
        return out

    def contrib(self, rays: np.ndarray, mods: list, accum: int = 1, flags: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        nrec = rays.shape[0] // accum
        arr = (RgModifier * len(mods))(*mods)
        ncol = 3 * sum(m.nbins for m in mods)
        out = np.zeros((nrec, ncol))
        self._check(self._L.rg_contrib(self._h, _dptr(rays), nrec, accum, flags, arr, len(mods), _dptr(out)))
        return out

    def contrib_device(self, rays_ptr: int, nrec: int, mods: list, out_ptr: int, accum: int = 1,
                       flags: int = 0) -> None:
        """rg_contrib on device buffers: rays_ptr holds nrec*accum x 6 doubles,
        out_ptr receives nrec x (3 * sum nbins) doubles (device addresses of
        this context's device, e.g. torch tensors' data_ptr())."""
        arr = (RgModifier * len(mods))(*mods)
        self._check(self._L.rg_contrib_device(self._h, C.c_void_p(rays_ptr), nrec, accum, flags, arr, len(mods),
                                              C.c_void_p(out_ptr)))

    def set_stream(self, stream: int) -> None:
        """Order device-buffer inputs after the work queued on this HIP stream
        (a raw hipStream_t handle, e.g. torch.cuda.current_stream().cuda_stream;
        0: the null stream)."""
        self._check(self._L.rg_set_stream(self._h, C.c_void_p(stream)))

    def contrib_adjoint(self, rays: np.ndarray, mods: list, dlde: np.ndarray, lcol: np.ndarray, mats: list,
                        accum: int = 1, flags: int = 0) -> np.ndarray:
        """d(loss)/d(material arguments) of materials `mats` (indices),
        (len(mats), 7): red, green, blue, specularity, roughness, and for
        trans transmissivity and transmitted specularity (zero otherwise).

        dlde: (nrec, 3) derivative of the loss with respect to each record's
        E = sum over columns of contribution x lcol; lcol: (3 * sum nbins,)
        radiance of every output column."""
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        nrec = rays.shape[0] // accum
        arr = (RgModifier * len(mods))(*mods)
        ncol = 3 * sum(m.nbins for m in mods)
        dlde = np.ascontiguousarray(dlde, dtype=np.float64).reshape(nrec, 3)
        lcol = np.ascontiguousarray(lcol, dtype=np.float64).reshape(ncol)
        mi = (C.c_int * max(len(mats), 1))(*[int(m) for m in mats])
        grad = np.zeros((len(mats), 7))
        self._check(self._L.rg_contrib_adjoint(self._h, _dptr(rays), nrec, accum, flags, arr, len(mods), _dptr(dlde),
                                               _dptr(lcol), mi, len(mats), _dptr(grad)))
        return grad

    def trace_view(self, view, xres: int, yres: int, pj: float = 0.0, flags: int = 0) -> np.ndarray:
        """A picture of a set-up view (pyradiance_amd.view.View): (yres, xres, 3)
        radiances, top scanline first, rays generated on the GPU."""
        v = RgView()
        v.type = ord(view.type)
        v.vp[:], v.vdir[:], v.hvec[:], v.vvec[:] = view.vp, view.vdir, view.hvec, view.vvec
        v.hn2, v.vn2, v.horiz, v.vert = view.hn2, view.vn2, view.horiz, view.vert
        v.hoff, v.voff, v.vfore, v.vaft = view.hoff, view.voff, view.vfore, view.vaft
        out = np.zeros((yres, xres, 3))
        self._check(self._L.rg_trace_view(self._h, C.byref(v), int(xres), int(yres), float(pj), int(flags),
                                          _dptr(out)))
        return out

    def first_hit(self, rays: np.ndarray):
        """(t, normal, prim) of each ray's first intersection (rtrace -oLN)."""
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        n = rays.shape[0]
        t = np.zeros(n)
        nrm = np.zeros((n, 3))
        prim = np.zeros(n, dtype=np.int32)
        self._check(self._L.rg_first_hit(self._h, _dptr(rays), n, _dptr(t), _dptr(nrm),
                                         prim.ctypes.data_as(C.POINTER(C.c_int))))
        return t, nrm, prim

    def material_index(self, name: str) -> int:
        i = C.c_int()
        self._check(self._L.rg_material_index(self._h, name.encode(), C.byref(i)))
        return int(i.value)

    def material_get(self, index: int) -> tuple:
        """(kind, real arguments) of material `index`."""
        k = C.c_int()
        f = np.zeros(8)
        self._check(self._L.rg_material_get(self._h, index, C.byref(k), _dptr(f), 8))
        n = {0: 5, 1: 5, 2: 7}.get(int(k.value), 8)
        return int(k.value), f[:n].copy()

    def material_set(self, index: int, f) -> None:
        f = np.ascontiguousarray(f, dtype=np.float64).ravel()
        self._check(self._L.rg_material_set(self._h, index, _dptr(f), f.size))

    def set_profiling(self, level: int):
        self._check(self._L.rg_set_profiling(self._h, level))

    def set_tuning(self, **kw):
        """Launch shape: trace_kernel (0/1), refill, window, blocks_per_cu,
        wide, max_big, shadow_cache, child_order, inflight (see rg_set_tuning)."""
        for k, v in kw.items():
            self._check(self._L.rg_set_tuning(self._h, k.encode(), int(v)))
        return self

    def set_bvh_builder(self, kind: int):
        """BVH_AUTO, BVH_SAH or BVH_LBVH for the next commit()."""
        self._check(self._L.rg_set_bvh_builder(self._h, kind))
        return self

    def last_stats(self) -> dict:
        v = (C.c_double * 9)()
        self._check(self._L.rg_last_stats(self._h, v, 9))
        keys = ("rays", "launches", "threads", "trace_ms", "nodes", "prims", "batches", "shade_ms", "node_bytes")
        return dict(zip(keys, list(v)))

    def last_nrays(self) -> int:
        n = C.c_ulonglong()
        self._check(self._L.rg_last_nrays(self._h, C.byref(n)))
        return int(n.value)


def contrib_batch(ctxs: list, rays: list, mods: list, accum: int = 1, flags: int = 0) -> list:
    """rcontrib over several scenes in one wavefront; parameters from ctxs[0].

    rays[i] are scene i's (nrec_i*accum, 6) rays; returns one array per scene.
    """
    if len(ctxs) != len(rays) or not ctxs:
        raise ValueError("one ray array per context")
    arrs = [np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 6) for r in rays]
    nrec = (C.c_longlong * len(arrs))(*[a.shape[0] // accum for a in arrs])
    allr = np.ascontiguousarray(np.concatenate([a[: (a.shape[0] // accum) * accum] for a in arrs]))
    ncol = 3 * sum(m.nbins for m in mods)
    out = np.zeros((sum(nrec), ncol))
    hs = (C.c_void_p * len(ctxs))(*[c._h for c in ctxs])
    arr = (RgModifier * len(mods))(*mods)
    L = lib()
    if L.rg_contrib_batch(hs, len(ctxs), _dptr(allr), nrec, accum, flags, arr, len(mods), _dptr(out)) != 0:
        raise RuntimeError(L.rg_last_error(ctxs[0]._h).decode())
    res, o = [], 0
    for n in nrec:
        res.append(out[o:o + n])
        o += n
    return res


def trace_batch(ctxs: list, rays: list, flags: int = 0) -> list:
    """rtrace over several scenes in one wavefront; parameters from ctxs[0]."""
    if len(ctxs) != len(rays) or not ctxs:
        raise ValueError("one ray array per context")
    arrs = [np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 6) for r in rays]
    n = (C.c_longlong * len(arrs))(*[a.shape[0] for a in arrs])
    allr = np.ascontiguousarray(np.concatenate(arrs))
    out = np.zeros((allr.shape[0], 3))
    hs = (C.c_void_p * len(ctxs))(*[c._h for c in ctxs])
    L = lib()
    if L.rg_trace_batch(hs, len(ctxs), _dptr(allr), n, flags, _dptr(out)) != 0:
        raise RuntimeError(L.rg_last_error(ctxs[0]._h).decode())
    res, o = [], 0
    for k in n:
        res.append(out[o:o + k])
        o += k
    return res


def dcsky(A: np.ndarray, B: np.ndarray, mode: int = 1) -> np.ndarray:
    """C[m,n,:] = sum_k A[m,k,:] * B[k,n,:] per channel on the GPU.
    mode 0: FP32 FMA; 1: BF16 hi/lo split MFMA in the form chosen by K;
    2 / 3 / 4: the split product with operands loaded per wavefront / staged
    through LDS in 128 x 128 tiles / rows of A in registers (K <= 160)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, K, _ = A.shape
    K2, N, _ = B.shape
    if K2 != K or A.shape[2] != 3 or B.shape[2] != 3:
        raise ValueError("matrix dimensions do not match")
    Cm = np.empty((M, N, 3), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    L = lib()
    if L.rg_dcsky(A.ctypes.data_as(fp), B.ctypes.data_as(fp), Cm.ctypes.data_as(fp), M, K, N, mode) != 0:
        raise RuntimeError(L.rg_last_error(None).decode())
    return Cm


def dcsky_device(A_ptr: int, B_ptr: int, C_ptr: int, M: int, K: int, N: int, mode: int = 1) -> None:
    """The same product on device buffers (raw device addresses, e.g. a
    torch tensor's data_ptr()): A (M,K,3), B (K,N,3), C (M,N,3) float32.
    Runs on the null stream and returns when it has finished."""
    L = lib()
    if L.rg_dcsky_device(C.c_void_p(A_ptr), C.c_void_p(B_ptr), C.c_void_p(C_ptr), M, K, N, mode) != 0:
        raise RuntimeError(L.rg_last_error(None).decode())


class Sky:
    """A prepared gendaymtx run (rg_sky): weather data read, steps set up."""

    def __init__(self, weather: str, opts: RgSkyOpts):
        self._L = lib()
        self._h = C.c_void_p()
        if self._L.rg_sky_create(C.byref(self._h), weather.encode(), C.byref(opts)) != 0:
            raise RuntimeError(self._L.rg_sky_error(None).decode())
        v = (C.c_int * 5)()
        self._L.rg_sky_shape(self._h, v)
        self.npatch, self.ncols, self.nsteps, self.ntsteps, self.nsuns = list(v)

    def site(self) -> tuple:
        v = (C.c_double * 2)()
        self._L.rg_sky_site(self._h, v)
        return v[0], v[1]

    def suns(self):
        """(ids, good, altitude, azimuth, direct illuminance) of placed suns."""
        ig = np.zeros((self.nsuns, 2), dtype=np.int32)
        v = np.zeros((self.nsuns, 3))
        if self.nsuns:
            self._L.rg_sky_suns(self._h, ig.ctypes.data_as(C.POINTER(C.c_int)), _dptr(v))
        return ig[:, 0], ig[:, 1], v[:, 0], v[:, 1], v[:, 2]

    def matrix(self) -> np.ndarray:
        """(npatch, ncols, 3) float32, the reference's row order."""
        out = np.empty((self.npatch, self.ncols, 3), dtype=np.float32)
        if self._L.rg_sky_matrix(self._h, out.ctypes.data_as(C.POINTER(C.c_float))) != 0:
            raise RuntimeError(self._L.rg_sky_error(self._h).decode())
        return out

    def matrix_device(self, ptr: int) -> None:
        """Writes the matrix into device memory at `ptr` (npatch*ncols*3 floats)."""
        if self._L.rg_sky_matrix_device(self._h, C.c_void_p(ptr)) != 0:
            raise RuntimeError(self._L.rg_sky_error(self._h).decode())

    def matrix_device_noalloc(self, ptr: int) -> None:
        """matrix_device() without device allocation (its buffers in a fixed
        96 MiB device array; larger plans raise)."""
        if self._L.rg_sky_matrix_device_noalloc(self._h, C.c_void_p(ptr)) != 0:
            raise RuntimeError(self._L.rg_sky_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.rg_sky_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
