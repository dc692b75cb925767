"""CPU checks of the C ABI boundary: the library loads and exports every
entry point include/b2g.h declares; record layouts match the header."""
import ctypes
import os
import re

from util import REPO, load_pkg


def header_functions():
    text = open(os.path.join(REPO, "include", "b2g.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(b2[ga]_[a-z_0-9]+)\s*\(", text)))


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
    assert eng.BODY_DTYPE.itemsize == 4 * 23
    assert eng.SHAPE_DTYPE.itemsize == 4 * (4 + 2 + 16 + 16 + 3 + 4)
    assert eng.JOINT_DTYPE.itemsize == 4 * (2 + 4 + 5 + 1 + 5)
    assert eng.CONTACT_DTYPE.itemsize == 4 * (5 + 2 + 2 + 4 + 2 + 2 + 2 + 4)
    assert ctypes.sizeof(eng.Capacity) == 20
    assert ctypes.sizeof(eng.WorldParams) == 24


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
