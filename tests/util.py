"""Shared helpers for the tests (loading the package, golden fixtures)."""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")


def load_pkg():
    name = "box2d_optimized_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(REPO, "box2d-optimized_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def manifold_fields_equal(x, y):
    """Bit-exact equality of the manifold fields that carry meaning for
    pointCount > 0 (type, normal, point, per-point data of the first
    pointCount points); returns a boolean per row."""
    n = len(x)
    ok = x["point_count"] == y["point_count"]
    touching = x["point_count"] > 0
    same_type = x["type"] == y["type"]
    ln = np.all(bits(x["local_normal"]) == bits(y["local_normal"]), axis=1)
    lp = np.all(bits(x["local_point"]) == bits(y["local_point"]), axis=1)
    p0 = np.all(bits(x["point_local"][:, :2]) == bits(y["point_local"][:, :2]), axis=1) & (x["id"][:, 0] == y["id"][:, 0])
    p1 = np.all(bits(x["point_local"][:, 2:]) == bits(y["point_local"][:, 2:]), axis=1) & (x["id"][:, 1] == y["id"][:, 1])
    # circles manifolds ignore localNormal
    ln = ln | (x["type"] == 0)
    body = same_type & ln & lp & p0 & (p1 | (x["point_count"] < 2))
    return ok & (~touching | body)


SCENE_BUILDERS = {}


def scene_for(name):
    """Rebuild a golden scene by name (same generator as make_golden.py)."""
    sys.path.insert(0, GOLDEN)
    import make_golden  # noqa: E402
    mk, steps = make_golden.SCENES[name]
    return mk(), steps
