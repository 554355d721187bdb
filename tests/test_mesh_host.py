"""Mesh ingestion on the CPU: compiled meshes written by the reference
obj2mesh are read, transformed and intersected by this project's host code
(scene_io.cpp, rg_scene.cpp and the host/device primitive tests of
rg_geom.hpp, built here with g++), and compared with the reference rtrace's
-oLNn output for the same rays: same hit or miss, same distance and both
normals (unperturbed and interpolated from the vertex normals)."""

import ctypes as C
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

CSRC = ROOT / "pyradiance_amd" / "csrc"


@pytest.fixture(scope="module")
def sc(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("scene") / "scenecheck.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", f"-I{CSRC}",
                    str(ROOT / "tests/native/scene_check.cpp"), str(CSRC / "scene_io.cpp"),
                    str(CSRC / "rg_scene.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    L.scene_compile.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_int]
    L.scene_trace.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    L.scene_prims.argtypes = [C.POINTER(C.c_int)]
    L.scene_mat_name.restype = C.c_char_p
    L.scene_text_fargs.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_char_p, C.c_int]
    return L


def compile_files(L, files):
    arr = (C.c_char_p * len(files))(*[str(f).encode() for f in files])
    err = C.create_string_buffer(512)
    n = L.scene_compile(arr, len(files), err, 512)
    if n < 0:
        raise RuntimeError(err.value.decode())
    return n


def trace(L, rays):
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    out = np.zeros((len(rays), 8))
    L.scene_trace(rays.ctypes.data_as(C.POINTER(C.c_double)), len(rays), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def test_mesh_hits_match_reference_rtrace(sc):
    g = np.load(GOLDEN / "mesh_rtrace.npz")
    n = compile_files(sc, [GOLDEN / "mesh" / "scene.rad"])
    assert n > 3000                      # 3 meshes of triangles + floor + lamp
    out = trace(sc, g["rays"])
    hit_ref = g["dist"] < 1e9
    hit = out[:, 7] >= 0
    assert np.array_equal(hit, hit_ref)
    m = hit
    # the reference prints 7 significant digits
    assert np.allclose(out[m, 0], g["dist"][m], rtol=1e-6, atol=1e-6)
    assert np.allclose(out[m, 1:4], g["normal"][m], atol=1e-6)
    assert np.allclose(out[m, 4:7], g["pnormal"][m], atol=1e-6)
    # smooth shading really perturbs: the interpolated normals differ
    assert np.abs(g["normal"][m] - g["pnormal"][m]).max() > 1e-2


def test_mesh_materials_and_kinds(sc):
    compile_files(sc, [GOLDEN / "mesh" / "scene.rad"])
    n = compile_files(sc, [GOLDEN / "mesh" / "scene.rad"])
    p = np.zeros((n, 4), dtype=np.int32)
    sc.scene_prims(p.ctypes.data_as(C.POINTER(C.c_int)))
    tri = p[:, 0] == 8
    names = {sc.scene_mat_name(int(m)).decode() for m in p[tri, 1]}
    # triangle materials from inside the meshes, and the torus's own modifier
    assert names == {"red", "chrome", "frost", "glassy"}
    # triangles whose material comes from the mesh are faces (flat); the torus's are not
    flat = {sc.scene_mat_name(int(m)).decode(): bool(f) for m, f in zip(p[tri, 1], p[tri, 2])}
    assert flat == {"red": True, "chrome": True, "frost": True, "glassy": False}


def test_transform_options_follow_xf(sc, tmp_path):
    """-t/-rx/-ry/-rz/-s/-mx/-i compose as row-vector matrices applied in
    order (common/xf.c): a unit triangle placed by a transform lands where
    the same operations applied by hand put it."""
    import math
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    refbin = ROOT / "oracle" / "_ref" / "obj2mesh"
    if not refbin.exists():
        pytest.skip("reference obj2mesh not built")
    subprocess.run([str(refbin), "tri.obj", "tri.rtm"], cwd=tmp_path, check=True)
    args = "-rz 90 -s 2 -t 1 0 0 -i 2 -ry 90"
    (tmp_path / "s.rad").write_text(f"void plastic m 0 0 5 .5 .5 .5 0 0\nm mesh t {len(args.split()) + 1} tri.rtm "
                                    f"{args} 0 0\n")
    compile_files(sc, [tmp_path / "s.rad"])
    geom = np.zeros(64)
    sc.scene_geom(geom.ctypes.data_as(C.POINTER(C.c_double)), 64)

    def rz(a):
        c, s = math.cos(math.radians(a)), math.sin(math.radians(a))
        return np.array([[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])

    def ry(a):
        c, s = math.cos(math.radians(a)), math.sin(math.radians(a))
        return np.array([[c, 0, -s, 0], [0, 1, 0, 0], [s, 0, c, 0], [0, 0, 0, 1]])

    S = np.diag([2.0, 2, 2, 1])
    T = np.eye(4)
    T[3, :3] = (1, 0, 0)
    M = rz(90) @ S @ T @ ry(90) @ ry(90)          # "-i 2" repeats what follows it
    want = []
    for v in ((0, 0, 0), (1, 0, 0), (0, 1, 0)):
        want.append((np.array([*v, 1.0]) @ M)[:3])
    got = geom[:9].reshape(3, 3)
    # obj2mesh quantises positions to its bounding cube (32 bits)
    assert np.allclose(got, np.array(want), atol=1e-6)


def test_instance_hits_match_reference_rtrace(sc):
    """Instanced octrees (frozen with every surface type, file-list, nested;
    rotated, scaled, mirrored, material override) against the reference
    rtrace -oLN."""
    g = np.load(GOLDEN / "inst_rtrace.npz")
    n = compile_files(sc, [GOLDEN / "inst" / "scene.rad"])
    # floor + lamp + 3 placements of a.oct (7 surfaces) + b.oct (2) + c's pad
    assert n == 2 + 3 * 7 + 2 + 1
    out = trace(sc, g["rays"])
    hit_ref = g["dist"] < 1e9
    hit = out[:, 7] >= 0
    assert np.array_equal(hit, hit_ref)
    m = hit
    assert np.allclose(out[m, 0], g["dist"][m], rtol=1e-6, atol=1e-6)
    assert np.allclose(out[m, 1:4], g["normal"][m], atol=1e-6)


def test_instance_materials(sc):
    n = compile_files(sc, [GOLDEN / "inst" / "scene.rad"])
    p = np.zeros((n, 4), dtype=np.int32)
    sc.scene_prims(p.ctypes.data_as(C.POINTER(C.c_int)))
    names = [sc.scene_mat_name(int(m)).decode() for m in p[:, 1]]
    # the glass-modified placement turns all seven of its surfaces to glass
    assert names.count("glassy") == 7
    assert names.count("wood") == 2 * 2 and names.count("steel") == 2 * 4 and names.count("frosted") == 2
    assert names.count("paint") == 2 and names.count("base") == 1


def test_reference_octrees_host(sc):
    """The reference's test octrees (tests/Resources/, kept under
    golden/refres/): contrib.oct (floor, ceiling, glow sky and ground)
    compiles to its two polygons; trace.oct's gendaylit sky glow is modified
    by a brightfunc pattern, which the GPU material set does not have, and is
    refused by name rather than shaded as if unmodified."""
    assert compile_files(sc, [GOLDEN / "refres" / "contrib.oct"]) == 2
    with pytest.raises(RuntimeError, match=r'"skyfunc" of type "brightfunc"'):
        compile_files(sc, [GOLDEN / "refres" / "trace.oct"])


def test_scene_reader_numbers_match_the_c_library(sc):
    """The scene reader takes plain decimal words through std::from_chars and
    anything else (a leading '+', hexadecimal, inf / nan, out-of-range
    magnitudes) through strtod, as before: every value must be the one
    strtod gives (Python's float() is correctly rounded like both)."""
    words = ["1", "-0", "+1.5", "0.1", "1e-320", "2.5E+3", ".5", "5.", "-7.25e-3",
             "0x1p3", "inf", "-inf", "1e400", "-1e400", "123456789012345678901234567890",
             "0.30000000000000004", "4.9406564584124654e-324", "2.2250738585072011e-308"]
    text = "void plastic m 0 0 %d %s\n" % (len(words), " ".join(words))
    out = (C.c_double * len(words))()
    err = C.create_string_buffer(256)
    n = sc.scene_text_fargs(text.encode(), 0, out, len(words), err, 256)
    assert n == len(words), err.value
    for w, v in zip(words, out):
        want = float.fromhex(w) if w.lower().startswith(("0x", "-0x", "+0x")) else float(w)
        assert v == want and np.signbit(v) == np.signbit(want), (w, v, want)
    # a quoted word and a comment still parse as before
    text = "void plastic m 0 0 5 1 2 3 4 5\n# comment\nvoid mixfunc x 4 'a b' c d e 0 0\n"
    assert sc.scene_text_fargs(text.encode(), 0, out, 5, err, 256) == 5
    assert list(out[:5]) == [1.0, 2.0, 3.0, 4.0, 5.0]
