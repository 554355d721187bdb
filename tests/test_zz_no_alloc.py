"""The allocation-free gendaymtx entry point (rg_sky_matrix_device_noalloc)
against the allocating one (rg_sky_matrix_device): the same kernels, so the
matrices are identical, on two calls in a row (the second smaller than the
first), through the C ABI; a plan larger than its fixed device array is
refused with an error."""

import ctypes as C
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_noalloc_sky_matrix_matches_allocating_entry_point():
    import torch
    from pyradiance_amd import _lib, gen, scenes
    L = _lib.lib()
    fn = L.rg_sky_matrix_device_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    meta = json.loads((GOLDEN / "sky_cases.json").read_text())
    cases = [("annual tape, Reinhart MF:4, 8760 steps", scenes.annual_wea(3), ["-m", "4"]),
             ("wea1 tape, MF:1 with options", meta["tapes"]["wea1"], meta["cases"]["wea1_m2_opts"]["args"])]
    for name, tape, args in cases:
        _say(f"noalloc gendaymtx: {name}")
        mtx, sky, _, _ = gen.sky_matrix(tape, args)
        try:
            d_alloc = torch.full(mtx.shape, float("nan"), dtype=torch.float32, device="cuda")
            d_free = torch.full(mtx.shape, float("nan"), dtype=torch.float32, device="cuda")
            sky.matrix_device(d_alloc.data_ptr())
            assert fn(sky._h, C.c_void_p(d_free.data_ptr())) == 0, L.rg_sky_error(sky._h).decode()
            torch.cuda.synchronize()
            assert torch.equal(d_free, d_alloc)
            assert np.array_equal(d_free.cpu().numpy(), mtx)
        finally:
            sky.close()
    _say("noalloc gendaymtx: plan beyond the fixed array")
    mtx, sky, _, _ = gen.sky_matrix(scenes.annual_wea(3), ["-m", "4", "-A"])
    try:
        d = torch.empty(mtx.shape, dtype=torch.float32, device="cuda")
        assert fn(sky._h, C.c_void_p(d.data_ptr())) != 0
        assert "allocation-free" in L.rg_sky_error(sky._h).decode()
    finally:
        sky.close()
