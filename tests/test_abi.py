"""CPU checks of the C ABI boundary: the library loads and exports every
entry point include/b2g.h declares; record layouts match the header."""
import ctypes
import os
import re

from util import REPO, load_pkg


def header_functions():
    text = open(os.path.join(REPO, "include", "b2g.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(b2[gas]_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    eng = load_pkg().engine
    lib = eng.lib()
    declared = header_functions()
    assert len(declared) >= 25
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(eng.EXPORTS) == declared


def test_record_sizes_match_header():
    eng = load_pkg().engine
    # 4-byte fields only: sizes are the field counts times four
    assert eng.BODY_DTYPE.itemsize == 4 * 25
    assert eng.SHAPE_DTYPE.itemsize == 4 * (4 + 2 + 16 + 16 + 3 + 4)
    assert eng.JOINT_DTYPE.itemsize == 4 * (2 + 4 + 5 + 1 + 5)
    assert eng.CONTACT_DTYPE.itemsize == 4 * (5 + 2 + 2 + 4 + 2 + 2 + 2 + 4)
    assert ctypes.sizeof(eng.Capacity) == 20
    assert ctypes.sizeof(eng.WorldParams) == 28
    assert ctypes.sizeof(eng.Counts) == 4 * 12
    # serving records: natural alignment, as the C compiler lays them out
    assert eng.REQUEST_DTYPE.itemsize == 8 + 4 * 4
    assert eng.RESULT_DTYPE.itemsize == 3 * 8 + 2 * 4
    assert ctypes.sizeof(eng.ServeConfig) == 8 + 4 * 4
    assert ctypes.sizeof(eng.ServeStats) == 3 * 8 + 3 * 8 + 2 * 4 + 8


def test_errors_without_device_are_reported_not_faked():
    eng = load_pkg().engine
    if eng.device_count() > 0:
        return
    import pytest
    with pytest.raises(eng.B2GError):
        eng.Batch()


def _example():
    return os.path.join(REPO, "box2d-optimized_amd", "build", "falling_box")


def test_drop_in_example_links_and_fails_loudly_without_device():
    """User code written against the b2World API links against the library;
    without a device it reports that, rather than stepping on the CPU."""
    import subprocess
    import pytest
    exe = _example()
    if not os.path.exists(exe):
        pytest.skip("example not built (run __graft_entry__.build())")
    eng = load_pkg().engine
    if eng.device_count() > 0:
        pytest.skip("device present: covered by the GPU test")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "device" in r.stdout


def test_scene_records_match_the_restatement_and_offsets():
    """b2a_scene_records (the batch assembler's per-world record builder)
    produces the same shapes and body mass data as the oracle's own setup of
    the scene, tagged with the world and shifted by the body offset."""
    import numpy as np
    import oracle
    from util import bits
    pkg = load_pkg()
    sc = pkg.scenes.small_mixed(seed=3, n=90)
    sc.add_revolute(1, 2, (0.0, 5.0))
    b, s, j = pkg.engine.scene_records(sc, world=7, body_offset=1000)
    w = oracle.RestatedWorld(sc)
    rs, rb = w.shapes(), w.bodies()
    for f in ("type", "count", "radius", "centroid", "vertices", "normals", "friction", "restitution"):
        assert np.array_equal(bits(s[f]), bits(rs[f])), f
    assert np.array_equal(s["body"], rs["body"] + 1000)
    for f in ("xf", "c", "a", "local_center", "inv_mass", "inv_i", "mass"):
        assert np.array_equal(bits(b[f]), bits(rb[f])), f
    assert (b["world"] == 7).all()
    assert len(j) == 1 and j["body_a"][0] == 1001 and j["body_b"][0] == 1002


def test_python_records_match_the_c_compiler_layout(tmp_path):
    """sizeof of every record in include/b2g.h, as gcc lays it out, equals
    the size of its Python mirror."""
    import shutil
    import subprocess
    import pytest
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    eng = load_pkg().engine
    names = {"b2g_body": eng.BODY_DTYPE.itemsize, "b2g_shape": eng.SHAPE_DTYPE.itemsize,
             "b2g_revolute": eng.JOINT_DTYPE.itemsize, "b2g_contact": eng.CONTACT_DTYPE.itemsize,
             "b2g_capacity": ctypes.sizeof(eng.Capacity), "b2g_world_params": ctypes.sizeof(eng.WorldParams),
             "b2g_counts": ctypes.sizeof(eng.Counts), "b2g_profile": ctypes.sizeof(eng.Profile),
             "b2s_request": eng.REQUEST_DTYPE.itemsize, "b2s_result": eng.RESULT_DTYPE.itemsize,
             "b2s_config": ctypes.sizeof(eng.ServeConfig), "b2s_stats": ctypes.sizeof(eng.ServeStats),
             "b2g_contact_event": eng.EVENT_DTYPE.itemsize, "b2g_ray_hit": eng.RAY_HIT_DTYPE.itemsize}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "b2g.h"\nint main(void) {\n' +
                   "".join(f'  printf("{n} %zu\\n", sizeof({n}));\n' for n in names) + "  return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)])
    got = dict(line.split() for line in subprocess.check_output([str(exe)], text=True).splitlines())
    assert {k: int(v) for k, v in got.items()} == names


def test_drop_in_refuses_unsupported_features_before_the_device():
    """User code that turns on what the device does not run (continuous
    physics, b2World's default; a disabled body; a sensor; a PreSolve
    listener; a filter overriding ShouldCollide) gets B2G_ERR_UNSUPPORTED
    with a message from Step, and the world is left as it was; the check
    comes before any device call, so it holds on the CPU. Sleeping (the other
    default) and a body put to sleep are served: those steps get as far as
    the device."""
    import subprocess
    import pytest
    exe = os.path.join(REPO, "box2d-optimized_amd", "build", "unsupported_features")
    if not os.path.exists(exe):
        pytest.skip("example not built (run __graft_entry__.build())")
    eng = load_pkg().engine
    if eng.device_count() > 0:
        pytest.skip("device present: covered by the GPU test")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    lines = r.stdout.strip().splitlines()
    assert r.returncode == 0, r.stdout + r.stderr
    assert len(lines) == 9
    status = [int(ln.split(":")[1].split()[0]) for ln in lines]
    # 5 = B2G_ERR_UNSUPPORTED, 6 = B2G_ERR_NO_DEVICE (it got as far as the device)
    assert status == [5, 6, 6, 6, 5, 5, 5, 5, 6]
    assert "SetContinuousPhysics(false)" in lines[0]
