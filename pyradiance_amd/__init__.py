"""pyradiance_amd: Radiance's rtrace / rcontrib ray path and the dctimestep
product on AMD MI355X (gfx950), behind pyradiance's call signatures.

    from pyradiance_amd import rtrace, Rcontrib, dctimestep

The compute runs in libradgpu.so (HIP kernels + C ABI, include/radgpu.h);
build it with `python -m pyradiance_amd.build`.
"""

from .gen import gendaymtx
from .rflux import rfluxmtx
from .rt import Rcontrib, rpict, rtrace
from .util import dcsky, dctimestep, read_matrix, write_matrix
from .view import vwrays

__all__ = ["rtrace", "Rcontrib", "rpict", "vwrays", "rfluxmtx", "gendaymtx", "dctimestep", "dcsky", "read_matrix", "write_matrix"]
