"""Host SAH BVH builder: structure and nearest-box walks against brute force.

The builder (pyradiance_amd/csrc/rg_bvh.cpp) is plain C++; it is compiled here
with g++ together with tests/native/bvh_check.cpp, so these checks run without
a GPU.
"""

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bvhcheck(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("bvh") / "bvhcheck.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-pthread", f"-I{ROOT / 'pyradiance_amd/csrc'}",
                    str(ROOT / "tests/native/bvh_check.cpp"), str(ROOT / "pyradiance_amd/csrc/rg_bvh.cpp"),
                    "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    L.check_sah.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.check_bvh4.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.check_sah_threads.argtypes = [C.c_int, C.c_uint, C.c_int]
    return L


@pytest.mark.parametrize("n,clustered", [(1, 0), (2, 0), (3, 0), (257, 0), (20000, 0), (20000, 1)])
def test_sah_tree_is_complete_and_walks_like_brute_force(bvhcheck, n, clustered):
    depth = C.c_int()
    bad = bvhcheck.check_sah(n, 1000, 11, 60, clustered, C.byref(depth))
    assert bad == 0, f"failed checks mask {bad:#x}"
    assert depth.value <= 60


def test_sah_depth_limit_forces_median_splits(bvhcheck):
    # a tight limit still yields a complete, correct tree within it
    depth = C.c_int()
    bad = bvhcheck.check_sah(50000, 300, 3, 30, 0, C.byref(depth))
    assert bad == 0, f"failed checks mask {bad:#x}"
    assert depth.value <= 30


@pytest.mark.parametrize("n,clustered", [(1, 0), (2, 0), (3, 0), (5, 0), (257, 0), (20000, 0), (20000, 1)])
def test_four_wide_tree_holds_the_grandchildren_and_walks_like_brute_force(bvhcheck, n, clustered):
    """The four-wide nodes the persistent traversal walks (rg_bvh.cpp
    collapse_bvh4): slots, boxes, coverage, depth and nearest-box walks."""
    d4, d2 = C.c_int(), C.c_int()
    bad = bvhcheck.check_bvh4(n, 1000, 11, 60, clustered, C.byref(d4))
    assert bad == 0, f"failed checks mask {bad:#x}"
    assert bvhcheck.check_sah(n, 0, 11, 60, clustered, C.byref(d2)) == 0
    if n >= 2:
        # binary depth counts edges below the root, four-wide depth counts nodes
        assert d4.value <= (d2.value + 2) // 2 + 1


@pytest.mark.parametrize("n,clustered", [(70000, 0), (200000, 0), (200000, 1)])
def test_sah_tree_is_the_same_for_any_host_thread_count(bvhcheck, n, clustered):
    """Above 65,536 boxes the builder splits the top of the tree on one
    thread and builds the subtrees below on up to 16: each range is split
    and its nodes numbered as the one-thread build does, so the node arrays
    agree byte for byte."""
    assert bvhcheck.check_sah_threads(n, 7, clustered) == 0
