"""CPU-side checks of the C ABI library: it loads, exports every declared entry point,
and its host-only functions (scene generators, camera) behave; no GPU work here."""
import ctypes as C

import numpy as np
import pytest

from research_amd import _lib as L
from research_amd import core


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = L.declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(L.SIGNATURES) <= set(names)


def test_cornell_reference_scene_records():
    sa = core.generate_scene(L.GEN_CORNELL_REF)
    assert len(sa.meshes) == 8 and len(sa.insts) == 8 and sa.tri_count == 36
    # short box: translate(0.35, 0, -0.35) * rotate(18 deg, +y), stored row-major (Scene.cpp:341-346)
    T = sa.insts[6]["transform"].reshape(4, 4)
    a = np.radians(18.0)
    assert np.allclose(T[:3, :3], [[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], atol=1e-6)
    assert np.allclose(T[:3, 3], [0.35, 0.0, -0.35])
    # add_quad winding (0,1,2),(0,2,3) and index_byte_offset bookkeeping
    assert sa.idx[:6].tolist() == [0, 1, 2, 0, 2, 3]
    assert sa.meshes["index_byte_offset"].tolist() == [i * 24 for i in range(6)] + [144, 288]


def test_cornell_random_is_seeded_and_lambertian():
    a = core.generate_scene(L.GEN_CORNELL_RANDOM, seed=1)
    b = core.generate_scene(L.GEN_CORNELL_RANDOM, seed=1)
    c = core.generate_scene(L.GEN_CORNELL_RANDOM, seed=2)
    assert a.tri_count == 36 and len(a.lights) == 0 and a.flags == L.SCENE_TRI_LIGHTS
    assert np.array_equal(a.verts.view(np.uint8), b.verts.view(np.uint8))
    assert not np.array_equal(a.verts.view(np.uint8), c.verts.view(np.uint8))
    nonem = a.mats[a.mats["emissive"].max(1) == 0]
    assert (nonem["flags"] & L.MAT_PURE_LAMBERT).all()
    assert (nonem["base_color"][:, :3] >= 0.1).all() and (nonem["base_color"][:, :3] <= 0.9).all()
    em = a.mats[a.mats["emissive"].max(1) > 0]
    assert len(em) == 1 and np.allclose(em["emissive"], 15.0)


def test_heightfield_full_size_counts():
    sa = core.generate_scene(L.GEN_HEIGHTFIELD, seed=2, scale=1.0)
    # 512^2 cells x 2 + 2000 icospheres x 320 triangles
    assert sa.tri_count == 512 * 512 * 2 + 2000 * 320 == 1_164_288
    em = sa.mats["emissive"].max(1) > 0
    assert em.sum() == 4096
    lum = sa.mats["emissive"][em].max(1)
    assert lum.min() >= 0.6 and lum.max() <= 100.0


def test_indoor_counts_and_window():
    sa = core.generate_scene(L.GEN_INDOOR, seed=4, scale=1.0)
    assert 900_000 < sa.tri_count < 1_100_000
    assert len(sa.lights) == 1 and sa.lights[0]["type"] == 1
    assert (sa.mats["flags"] & L.MAT_PURE_LAMBERT).all()


def test_camera_rejects_bad_arguments():
    cam = L.Camera()
    lib = L.load()
    rc = lib.rrh_camera_look_at(C.byref(L.f3((0, 0, 0))), C.byref(L.f3((0, 0, 0))), 45.0, 10, 10, 0.01, 100.0, 0,
                                C.byref(cam))
    assert rc == -1 and b"eye == center" in lib.rrh_last_error()
    with pytest.raises(L.RRHError):
        core.camera_look_at((0, 0, 3), (0, 0, 0), 45.0, 0, 10)
    with pytest.raises(L.RRHError):
        core.generate_scene(99)


@pytest.mark.parametrize("compiler,std", [("gcc", "-std=c11"), ("g++", "-std=c++17")])
def test_header_layout_checks_compile(tmp_path, compiler, std):
    """include/rrh.h carries static size / offset checks of the GpuScene.h records; a caller
    compiling against it in C11 or C++ builds, and a drifted layout does not."""
    import subprocess
    from pathlib import Path

    inc = Path(__file__).resolve().parents[1] / "include"
    ok = tmp_path / "ok.c"
    ok.write_text('#include "rrh.h"\nint main(void) { return (int)sizeof(rrh_camera) - 304; }\n')
    lang = ["-x", "c++"] if compiler == "g++" else []
    r = subprocess.run([compiler, std, "-Wall", "-Werror", *lang, f"-I{inc}", "-c", str(ok), "-o",
                        str(tmp_path / "ok.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # a record whose layout drifts (an extra field in GpuLight) must not compile
    bad = tmp_path / "rrh.h"
    bad.write_text((inc / "rrh.h").read_text().replace("float emission[3]; float intensity; } rrh_light;",
                                                        "float emission[3]; float intensity; float x; } rrh_light;"))
    r = subprocess.run([compiler, std, *lang, f"-I{tmp_path}", "-c", str(ok), "-o", str(tmp_path / "bad.o")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "rrh_light must be 48 B" in r.stderr
