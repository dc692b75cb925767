"""CPU checkers for the b2World::Step hot path (test infrastructure only).

Two independent checkers live here:

* ``Ref`` -- the reference library itself, compiled from its own sources by
  oracle/Makefile into oracle/_ref/libbox2d_ref.so, driven through
  oracle/ref_driver.cpp. It is the parity anchor and the CPU baseline.
* ``Restated`` -- oracle/restate.cpp, a CPU restatement of the reference's
  step (narrowphase, contact solver, revolute joint, integration, broadphase
  pair set) that can run the velocity/position sweeps in an externally given
  constraint order (the GPU's colour order) as well as in island order.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package; the product path never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libbox2d_ref.so")
ORC_SO = os.path.join(HERE, "build", "liboracle.so")
REPO = os.path.dirname(HERE)


def build(ref: bool = True) -> None:
    """Compile the restatement (and the reference where its sources exist)."""
    target = "all" if ref else "oracle"
    subprocess.check_call(["make", "-s", "-C", HERE, "-j8", target])


# ---- ctypes mirrors of include/b2g.h records ------------------------------

class Shape(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32), ("body", ctypes.c_int32), ("count", ctypes.c_int32),
        ("radius", ctypes.c_float), ("centroid", ctypes.c_float * 2),
        ("vertices", ctypes.c_float * 16), ("normals", ctypes.c_float * 16),
        ("friction", ctypes.c_float), ("restitution", ctypes.c_float),
        ("restitution_threshold", ctypes.c_float),
        ("category", ctypes.c_int32), ("mask", ctypes.c_int32), ("group", ctypes.c_int32),
        ("is_sensor", ctypes.c_int32),
    ]


class Contact(ctypes.Structure):
    _fields_ = [
        ("shape_a", ctypes.c_int32), ("shape_b", ctypes.c_int32),
        ("touching", ctypes.c_int32), ("type", ctypes.c_int32),
        ("point_count", ctypes.c_int32),
        ("local_normal", ctypes.c_float * 2), ("local_point", ctypes.c_float * 2),
        ("point_local", ctypes.c_float * 4),
        ("normal_impulse", ctypes.c_float * 2), ("tangent_impulse", ctypes.c_float * 2),
        ("id", ctypes.c_uint32 * 2),
        ("friction", ctypes.c_float), ("restitution", ctypes.c_float),
        ("restitution_threshold", ctypes.c_float), ("tangent_speed", ctypes.c_float),
    ]


CONTACT_DTYPE = np.dtype([
    ("shape_a", "<i4"), ("shape_b", "<i4"), ("touching", "<i4"), ("type", "<i4"),
    ("point_count", "<i4"), ("local_normal", "<f4", 2), ("local_point", "<f4", 2),
    ("point_local", "<f4", 4), ("normal_impulse", "<f4", 2), ("tangent_impulse", "<f4", 2),
    ("id", "<u4", 2), ("friction", "<f4"), ("restitution", "<f4"),
    ("restitution_threshold", "<f4"), ("tangent_speed", "<f4"),
])
SHAPE_DTYPE = np.dtype([
    ("type", "<i4"), ("body", "<i4"), ("count", "<i4"), ("radius", "<f4"),
    ("centroid", "<f4", 2), ("vertices", "<f4", 16), ("normals", "<f4", 16),
    ("friction", "<f4"), ("restitution", "<f4"), ("restitution_threshold", "<f4"),
    ("category", "<i4"), ("mask", "<i4"), ("group", "<i4"), ("is_sensor", "<i4"),
])
assert CONTACT_DTYPE.itemsize == ctypes.sizeof(Contact)
assert SHAPE_DTYPE.itemsize == ctypes.sizeof(Shape)

_P = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Ref:
    """The compiled reference library (oracle/_ref/libbox2d_ref.so)."""

    _lib = None

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(REF_SO):
                raise RuntimeError(f"reference library missing: {REF_SO}")
            lib = ctypes.CDLL(REF_SO)
            lib.ref_world_from_scene.restype = _P
            lib.ref_world_from_scene.argtypes = [_P, _I, _I, _I]
            lib.ref_world_destroy.argtypes = [_P]
            lib.ref_world_step.argtypes = [_P, _F, _I, _I, _I, _P]
            lib.ref_world_get_bodies.argtypes = [_P, _P]
            lib.ref_world_get_mass.argtypes = [_P, _P]
            lib.ref_world_contact_count.argtypes = [_P]
            lib.ref_world_get_contacts.argtypes = [_P, _I, _P]
            lib.ref_world_get_joint_impulses.argtypes = [_P, _P]
            lib.ref_shape_from_scene.argtypes = [_P, _I, _P]
            lib.ref_collide.argtypes = [_P, _P, _P, _P, _P]
            lib.ref_world_manifold.argtypes = [_P, _P, _F, _P, _F, _P]
            lib.ref_world_profile.argtypes = [_P, _P]
            lib.ref_world_query_aabb.argtypes = [_P, _P, _I, _P]
            lib.ref_world_query_aabb.restype = _I
            lib.ref_world_raycast_all.argtypes = [_P, _P, _I, _P]
            lib.ref_world_raycast_all.restype = _I
            cls._lib = lib
        return cls._lib


class RefWorld:
    """A reference b2World built from a scene (continuous physics off; sleep
    off unless allow_sleep)."""

    def __init__(self, scene, warm_starting=True, allow_sleep=False):
        self.lib = Ref.lib()
        self.scene = scene
        self.cscene = scene.to_c()
        self.h = self.lib.ref_world_from_scene(self.cscene.ptr(), int(warm_starting), int(allow_sleep), 0)
        self.nb = scene.body_count
        self.nj = scene.joint_count

    def close(self):
        if self.h:
            self.lib.ref_world_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1, dt=1.0 / 60.0, vi=8, pi=3, timed=False):
        ms = np.zeros(n, dtype=np.float64) if timed else None
        self.lib.ref_world_step(self.h, dt, vi, pi, n, _ptr(ms) if timed else None)
        return ms

    def bodies(self) -> np.ndarray:
        out = np.zeros((self.nb, 10), dtype=np.float32)
        self.lib.ref_world_get_bodies(self.h, _ptr(out))
        return out

    def awake(self) -> np.ndarray:
        """b2Body::IsAwake per body."""
        out = np.zeros(self.nb, dtype=np.int32)
        self.lib.ref_world_get_awake.argtypes = [_P, _P]
        self.lib.ref_world_get_awake(self.h, _ptr(out))
        return out

    def mass(self) -> np.ndarray:
        out = np.zeros((self.nb, 4), dtype=np.float32)
        self.lib.ref_world_get_mass(self.h, _ptr(out))
        return out

    def contacts(self) -> np.ndarray:
        n = self.lib.ref_world_contact_count(self.h)
        out = np.zeros(n, dtype=CONTACT_DTYPE)
        if n:
            self.lib.ref_world_get_contacts(self.h, n, _ptr(out))
        return out

    def joint_impulses(self) -> np.ndarray:
        out = np.zeros((self.nj, 3), dtype=np.float32)
        if self.nj:
            self.lib.ref_world_get_joint_impulses(self.h, _ptr(out))
        return out

    def query_aabb(self, box) -> np.ndarray:
        """Fixture indices b2World::QueryAABB reports for the box."""
        b = np.ascontiguousarray(box, dtype=np.float32)
        n = self.lib.ref_world_query_aabb(self.h, _ptr(b), 0, None)
        out = np.zeros(max(n, 1), np.int32)
        self.lib.ref_world_query_aabb(self.h, _ptr(b), n, _ptr(out))
        return out[:n]

    def raycast_all(self, seg) -> np.ndarray:
        """Every hit b2World::RayCast reports for the segment (callback
        returning 1): rows of (fixture, fraction, normal.x, normal.y, point.x, point.y)."""
        sg = np.ascontiguousarray(seg, dtype=np.float32)
        n = self.lib.ref_world_raycast_all(self.h, _ptr(sg), 0, None)
        out = np.zeros((max(n, 1), 6), np.float32)
        self.lib.ref_world_raycast_all(self.h, _ptr(sg), n, _ptr(out))
        return out[:n]

    def profile(self):
        out = np.zeros(4, dtype=np.float32)
        self.lib.ref_world_profile(self.h, _ptr(out))
        return out


def ref_shapes(scene) -> np.ndarray:
    """Processed shapes (b2PolygonShape::Set / SetAsBox results) per fixture."""
    lib = Ref.lib()
    cs = scene.to_c()
    out = np.zeros(scene.fixture_count, dtype=SHAPE_DTYPE)
    for i in range(scene.fixture_count):
        lib.ref_shape_from_scene(cs.ptr(), i, _ptr(out[i:i + 1]))
    return out


def ref_collide(shape_a: np.ndarray, xf_a, shape_b: np.ndarray, xf_b) -> np.ndarray:
    """Reference manifold for one pair (records are SHAPE_DTYPE scalars/1-arrays)."""
    lib = Ref.lib()
    sa = np.ascontiguousarray(np.atleast_1d(shape_a))
    sb = np.ascontiguousarray(np.atleast_1d(shape_b))
    xa = np.ascontiguousarray(xf_a, dtype=np.float32)
    xb = np.ascontiguousarray(xf_b, dtype=np.float32)
    out = np.zeros(1, dtype=CONTACT_DTYPE)
    lib.ref_collide(_ptr(sa), _ptr(xa), _ptr(sb), _ptr(xb), _ptr(out))
    return out[0]


def ref_world_manifold(contact: np.ndarray, xf_a, ra, xf_b, rb) -> np.ndarray:
    lib = Ref.lib()
    c = np.ascontiguousarray(np.atleast_1d(contact))
    xa = np.ascontiguousarray(xf_a, dtype=np.float32)
    xb = np.ascontiguousarray(xf_b, dtype=np.float32)
    out = np.zeros(8, dtype=np.float32)
    lib.ref_world_manifold(_ptr(c), _ptr(xa), float(ra), _ptr(xb), float(rb), _ptr(out))
    return out


BODY_DTYPE = np.dtype([
    ("type", "<i4"), ("world", "<i4"), ("xf", "<f4", 4), ("c", "<f4", 2), ("a", "<f4"),
    ("local_center", "<f4", 2), ("v", "<f4", 2), ("w", "<f4"), ("mass", "<f4"),
    ("inv_mass", "<f4"), ("inv_i", "<f4"), ("linear_damping", "<f4"),
    ("angular_damping", "<f4"), ("gravity_scale", "<f4"), ("force", "<f4", 2), ("torque", "<f4"),
    ("flags", "<i4"), ("sleep_time", "<f4"),
])
JOINT_DTYPE = np.dtype([
    ("body_a", "<i4"), ("body_b", "<i4"), ("local_anchor_a", "<f4", 2),
    ("local_anchor_b", "<f4", 2), ("reference_angle", "<f4"), ("lower_angle", "<f4"),
    ("upper_angle", "<f4"), ("motor_speed", "<f4"), ("max_motor_torque", "<f4"),
    ("flags", "<i4"), ("impulse", "<f4", 2), ("motor_impulse", "<f4"),
    ("lower_impulse", "<f4"), ("upper_impulse", "<f4"),
])


def contact_keys(contacts: np.ndarray) -> np.ndarray:
    a = contacts["shape_a"].astype(np.uint64)
    b = contacts["shape_b"].astype(np.uint64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    return (lo << np.uint64(32)) | hi


class Restated:
    """The CPU restatement (oracle/build/liboracle.so)."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORC_SO):
                raise RuntimeError(f"restatement library missing: {ORC_SO}")
            lib = ctypes.CDLL(ORC_SO)
            lib.orc_world_from_scene.restype = _P
            lib.orc_world_from_scene.argtypes = [_P, _I]
            lib.orc_world_destroy.argtypes = [_P]
            for f in ("orc_world_get_shapes", "orc_world_get_bodies", "orc_world_set_bodies",
                      "orc_world_get_joints", "orc_world_set_joints", "orc_world_get_contacts"):
                getattr(lib, f).argtypes = [_P, _P]
            lib.orc_world_body_count.argtypes = [_P]
            lib.orc_world_contact_count.argtypes = [_P]
            lib.orc_world_set_contacts.argtypes = [_P, _I, _P]
            lib.orc_world_set_inv_dt0.argtypes = [_P, _F]
            lib.orc_world_set_colours.argtypes = [_P, _I, _P, _P, _P]
            lib.orc_world_step.argtypes = [_P, _F, _I, _I, ctypes.c_uint]
            lib.orc_collide.argtypes = [_P, _P, _P, _P, _P]
            lib.orc_world_manifold.argtypes = [_P, _P, _F, _P, _F, _P]
            lib.orc_shape_aabb.argtypes = [_P, _P, _P]
            lib.orc_find_pairs.restype = ctypes.c_longlong
            lib.orc_find_pairs.argtypes = [_I, _P, _P, _P, _P, _I, _P, ctypes.c_longlong, _P]
            lib.orc_block_solve.argtypes = [_P, _P, _P, _P]
            cls._lib = lib
        return cls._lib


class RestatedWorld:
    def __init__(self, scene, warm_starting=True):
        self.lib = Restated.lib()
        self.scene = scene
        self.cscene = scene.to_c()
        self.h = self.lib.orc_world_from_scene(self.cscene.ptr(), int(warm_starting))
        self.nb = scene.body_count
        self.nf = scene.fixture_count
        self.nj = scene.joint_count

    def close(self):
        if self.h:
            self.lib.orc_world_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1, dt=1.0 / 60.0, vi=8, pi=3, flags=0):
        for _ in range(n):
            self.lib.orc_world_step(self.h, dt, vi, pi, flags)

    def shapes(self):
        out = np.zeros(self.nf, dtype=SHAPE_DTYPE)
        self.lib.orc_world_get_shapes(self.h, _ptr(out))
        return out

    def bodies(self):
        out = np.zeros(self.nb, dtype=BODY_DTYPE)
        self.lib.orc_world_get_bodies(self.h, _ptr(out))
        return out

    def set_bodies(self, b):
        b = np.ascontiguousarray(b, dtype=BODY_DTYPE)
        self.lib.orc_world_set_bodies(self.h, _ptr(b))

    def joints(self):
        out = np.zeros(self.nj, dtype=JOINT_DTYPE)
        if self.nj:
            self.lib.orc_world_get_joints(self.h, _ptr(out))
        return out

    def set_joints(self, j):
        j = np.ascontiguousarray(j, dtype=JOINT_DTYPE)
        if self.nj:
            self.lib.orc_world_set_joints(self.h, _ptr(j))

    def contacts(self):
        n = self.lib.orc_world_contact_count(self.h)
        out = np.zeros(n, dtype=CONTACT_DTYPE)
        if n:
            self.lib.orc_world_get_contacts(self.h, _ptr(out))
        return out

    def set_contacts(self, c):
        c = np.ascontiguousarray(c, dtype=CONTACT_DTYPE)
        self.lib.orc_world_set_contacts(self.h, len(c), _ptr(c))

    def set_inv_dt0(self, v):
        self.lib.orc_world_set_inv_dt0(self.h, float(v))

    def set_allow_sleep(self, flag):
        self.lib.orc_world_set_allow_sleep.argtypes = [_P, _I]
        self.lib.orc_world_set_allow_sleep(self.h, int(flag))


    def set_collide_order(self, keys):
        """The order the next step's Collide walks the contacts (the
        reference's list order), wake-ups counting at once."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        self.lib.orc_world_set_collide_order.argtypes = [_P, _I, _P]
        self.lib.orc_world_set_collide_order(self.h, len(keys), _ptr(keys))

    def set_colours(self, keys, colours, joint_colours=None):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        colours = np.ascontiguousarray(colours, dtype=np.int32)
        jc = None if joint_colours is None else np.ascontiguousarray(joint_colours, dtype=np.int32)
        self.lib.orc_world_set_colours(self.h, len(keys), _ptr(keys), _ptr(colours),
                                       None if jc is None else _ptr(jc))


def orc_collide(shape_a, xf_a, shape_b, xf_b):
    lib = Restated.lib()
    sa = np.ascontiguousarray(np.atleast_1d(shape_a))
    sb = np.ascontiguousarray(np.atleast_1d(shape_b))
    xa = np.ascontiguousarray(xf_a, dtype=np.float32)
    xb = np.ascontiguousarray(xf_b, dtype=np.float32)
    out = np.zeros(1, dtype=CONTACT_DTYPE)
    lib.orc_collide(_ptr(sa), _ptr(xa), _ptr(sb), _ptr(xb), _ptr(out))
    return out[0]


def orc_world_manifold(contact, xf_a, ra, xf_b, rb):
    lib = Restated.lib()
    c = np.ascontiguousarray(np.atleast_1d(contact))
    xa = np.ascontiguousarray(xf_a, dtype=np.float32)
    xb = np.ascontiguousarray(xf_b, dtype=np.float32)
    out = np.zeros(8, dtype=np.float32)
    lib.orc_world_manifold(_ptr(c), _ptr(xa), float(ra), _ptr(xb), float(rb), _ptr(out))
    return out


def orc_shape_aabbs(shapes, xfs):
    lib = Restated.lib()
    out = np.zeros((len(shapes), 4), dtype=np.float32)
    for i in range(len(shapes)):
        s = np.ascontiguousarray(shapes[i:i + 1])
        x = np.ascontiguousarray(xfs[i], dtype=np.float32)
        lib.orc_shape_aabb(_ptr(s), _ptr(x), _ptr(out[i]))
    return out


def orc_find_pairs(aabbs, shape_body, shape_type, filt, body_type):
    lib = Restated.lib()
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32)
    sb = np.ascontiguousarray(shape_body, dtype=np.int32)
    st = np.ascontiguousarray(shape_type, dtype=np.int32)
    bt = np.ascontiguousarray(body_type, dtype=np.int32)
    f = None if filt is None else np.ascontiguousarray(filt, dtype=np.int32)
    n = len(sb)
    cnt = lib.orc_find_pairs(n, _ptr(aabbs), _ptr(sb), _ptr(st), None if f is None else _ptr(f),
                             len(bt), _ptr(bt), 0, None)
    out = np.zeros(max(cnt, 1), dtype=np.uint64)
    lib.orc_find_pairs(n, _ptr(aabbs), _ptr(sb), _ptr(st), None if f is None else _ptr(f),
                       len(bt), _ptr(bt), cnt, _ptr(out))
    return out[:cnt]


def ref_step_ordered(rw: RefWorld, dt=1.0 / 60.0, vi=8, pi=3):
    """Step the reference once; return (contact keys, joint indices) in the
    order its island solver visited them."""
    lib = Ref.lib()
    lib.ref_world_step_ordered.argtypes = [_P, _F, _I, _I, _I, _P, _P, _P, _P]
    cap = max(4 * lib.ref_world_contact_count(rw.h) + 64, 64)
    keys = np.zeros(cap, dtype=np.uint64)
    joints = np.zeros(max(rw.nj, 1), dtype=np.int32)
    nc = ctypes.c_int(0)
    nj = ctypes.c_int(0)
    rc = lib.ref_world_step_ordered(rw.h, dt, vi, pi, cap, _ptr(keys), ctypes.byref(nc),
                                    _ptr(joints), ctypes.byref(nj))
    assert rc == 0
    return keys[:nc.value], joints[:nj.value]


def ref_last_events(rw: RefWorld):
    """Begin (1) / End (2) contact events the reference's listener received
    during the last ref_step_ordered, in call order: (kinds, fixture keys,
    how many leading events Collide made; the rest are EndContact calls of
    touching contacts the broadphase destroyed)."""
    lib = Ref.lib()
    lib.ref_world_last_events.argtypes = [_P, _I, _P, _P, _P]
    n = lib.ref_world_last_events(rw.h, 0, None, None, None)
    kinds = np.zeros(max(n, 1), np.int32)
    keys = np.zeros(max(n, 1), np.uint64)
    nc = ctypes.c_int(0)
    lib.ref_world_last_events(rw.h, n, _ptr(kinds), _ptr(keys), ctypes.byref(nc))
    return kinds[:n], keys[:n], nc.value


def ref_last_postsolve(rw: RefWorld):
    """PostSolve calls of the last ref_step_ordered: (contact keys, impulse
    counts), in call order."""
    lib = Ref.lib()
    lib.ref_world_last_postsolve.argtypes = [_P, _I, _P, _P]
    n = lib.ref_world_last_postsolve(rw.h, 0, None, None)
    keys = np.zeros(max(n, 1), np.uint64)
    counts = np.zeros(max(n, 1), np.int32)
    lib.ref_world_last_postsolve(rw.h, n, _ptr(keys), _ptr(counts))
    return keys[:n], counts[:n]


def ref_state_to_records(rw: RefWorld, restated_bodies: np.ndarray) -> np.ndarray:
    """Body records with the reference's current sweep/velocity/transform and
    the restatement's (bit-identical) mass properties."""
    rb = rw.bodies()
    out = restated_bodies.copy()
    out["c"] = rb[:, 0:2]
    out["a"] = rb[:, 2]
    out["v"] = rb[:, 3:5]
    out["w"] = rb[:, 5]
    out["xf"] = rb[:, 6:10]
    return out


def ref_aabbs(rw: RefWorld) -> np.ndarray:
    lib = Ref.lib()
    lib.ref_world_get_aabbs.argtypes = [_P, _P]
    out = np.zeros((rw.scene.fixture_count, 4), dtype=np.float32)
    lib.ref_world_get_aabbs(rw.h, _ptr(out))
    return out
