"""ctypes binding of lib/libbox2d_amd.so (the b2g C ABI in include/b2g.h).

The library is the product: HIP kernels for gfx950 behind a C ABI. This
module only marshals arrays; it raises if the library is missing or no HIP
device is visible -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbox2d_amd.so")

# status codes (include/b2g.h)
OK, ERR_HIP, ERR_OOM, ERR_CAPACITY, ERR_INVALID, ERR_UNSUPPORTED, ERR_NO_DEVICE = range(7)
STEP_FIND_NEW_CONTACTS, STEP_PROFILE, STEP_ASYNC, STEP_SKIP_BROADPHASE, STEP_GRAPH, STEP_EVENTS = 1, 2, 4, 8, 16, 32
STEP_COLLIDE_ONLY = 64
EVENT_BEGIN, EVENT_END, EVENT_DESTROYED = 1, 2, 3

BODY_DTYPE = np.dtype([
    ("type", "<i4"), ("world", "<i4"), ("xf", "<f4", 4), ("c", "<f4", 2), ("a", "<f4"),
    ("local_center", "<f4", 2), ("v", "<f4", 2), ("w", "<f4"), ("mass", "<f4"),
    ("inv_mass", "<f4"), ("inv_i", "<f4"), ("linear_damping", "<f4"),
    ("angular_damping", "<f4"), ("gravity_scale", "<f4"), ("force", "<f4", 2), ("torque", "<f4"),
    ("flags", "<i4"), ("sleep_time", "<f4"),
])
BODY_ASLEEP, BODY_NO_AUTOSLEEP = 1, 2
SHAPE_DTYPE = np.dtype([
    ("type", "<i4"), ("body", "<i4"), ("count", "<i4"), ("radius", "<f4"),
    ("centroid", "<f4", 2), ("vertices", "<f4", 16), ("normals", "<f4", 16),
    ("friction", "<f4"), ("restitution", "<f4"), ("restitution_threshold", "<f4"),
    ("category", "<i4"), ("mask", "<i4"), ("group", "<i4"), ("is_sensor", "<i4"),
])
JOINT_DTYPE = np.dtype([
    ("body_a", "<i4"), ("body_b", "<i4"), ("local_anchor_a", "<f4", 2),
    ("local_anchor_b", "<f4", 2), ("reference_angle", "<f4"), ("lower_angle", "<f4"),
    ("upper_angle", "<f4"), ("motor_speed", "<f4"), ("max_motor_torque", "<f4"),
    ("flags", "<i4"), ("impulse", "<f4", 2), ("motor_impulse", "<f4"),
    ("lower_impulse", "<f4"), ("upper_impulse", "<f4"),
])
CONTACT_DTYPE = np.dtype([
    ("shape_a", "<i4"), ("shape_b", "<i4"), ("touching", "<i4"), ("type", "<i4"),
    ("point_count", "<i4"), ("local_normal", "<f4", 2), ("local_point", "<f4", 2),
    ("point_local", "<f4", 4), ("normal_impulse", "<f4", 2), ("tangent_impulse", "<f4", 2),
    ("id", "<u4", 2), ("friction", "<f4"), ("restitution", "<f4"),
    ("restitution_threshold", "<f4"), ("tangent_speed", "<f4"),
])


class Capacity(ctypes.Structure):
    _fields_ = [("worlds", ctypes.c_int32), ("bodies", ctypes.c_int32), ("shapes", ctypes.c_int32),
                ("joints", ctypes.c_int32), ("contacts", ctypes.c_int32)]


class WorldParams(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 2), ("dt", ctypes.c_float), ("dt_ratio", ctypes.c_float),
                ("warm_starting", ctypes.c_int32), ("active", ctypes.c_int32), ("allow_sleep", ctypes.c_int32)]


class Counts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "worlds", "bodies", "shapes", "joints", "contacts", "touching", "colours", "islands",
        "pos_iters_run", "status", "serial", "graph_steps")]


class Profile(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("step", "collide", "solve", "broadphase")]


_P = ctypes.c_void_p
_I = ctypes.c_int32
_F = ctypes.c_float
_lib = None

# every entry point declared in include/b2g.h
EXPORTS = [
    "b2g_device_count", "b2g_last_error", "b2g_batch_create", "b2g_batch_destroy",
    "b2g_set_worlds", "b2g_set_world_params", "b2g_upload_bodies", "b2g_upload_shapes",
    "b2g_upload_joints", "b2g_upload_contacts", "b2g_batch_step", "b2g_batch_wait",
    "b2g_get_counts", "b2g_get_profile", "b2g_download_bodies", "b2g_download_contacts",
    "b2g_download_joints", "b2g_download_contact_colours", "b2g_collide_pairs", "b2g_collide_pairs_noalloc",
    "b2g_find_pairs",
    "b2a_world_from_scene", "b2a_world_destroy", "b2a_world_step", "b2a_world_get_bodies",
    "b2a_world_contact_count", "b2a_world_get_contacts", "b2a_world_batch", "b2a_world_prepare",
    "b2g_time_kernel", "b2g_kernel_time", "b2a_scene_records", "b2a_batch_from_scenes", "b2s_run",
    "b2g_set_option", "b2g_get_option", "b2g_upload_contact_colours", "b2g_snapshot_save",
    "b2g_snapshot_load", "b2g_download_events", "b2g_query_aabbs", "b2g_raycast",
    "b2g_set_counts", "b2g_download_solver_points", "b2a_world_destroy_body", "b2a_world_destroy_joint",
    "b2a_world_destroy_fixture", "b2a_world_set_filter", "b2a_world_body_count", "b2a_world_set_allow_sleep",
]
SOLVER_AUTO, SOLVER_COLOURS, SOLVER_WORLDS, SOLVER_CELLS = 0, 1, 2, 3
OPT_SOLVER, OPT_WORLD_BLOCK, OPT_TUNED_WORLD_BLOCK, OPT_LAST_SOLVER, OPT_PER_WORLD_ELIGIBLE = 1, 2, 3, 4, 5
OPT_WORLD_LDS, OPT_TUNED_WORLD_LDS, OPT_KEEP_IDLE = 6, 7, 8
OPT_PAIR_SEARCH, OPT_LAST_PAIR_SEARCH = 9, 10
KERNEL_NARROWPHASE, KERNEL_WARM_START, KERNEL_CONTACT_VELOCITY, KERNEL_CONTACT_POSITION, \
    KERNEL_PAIRS_WRITE, KERNEL_SOLVER = range(6)


EVENT_DTYPE = np.dtype([("kind", "<i4"), ("solver_points", "<i4"), ("contact", CONTACT_DTYPE)])
RAY_HIT_DTYPE = np.dtype([("shape", "<i4"), ("fraction", "<f4"), ("normal", "<f4", 2)])
REQUEST_DTYPE = np.dtype([("arrival", "<f8"), ("world", "<i4"), ("substeps", "<i4"), ("dt", "<f4"),
                          ("reserved", "<i4")])
RESULT_DTYPE = np.dtype([("arrival", "<f8"), ("admitted", "<f8"), ("completed", "<f8"), ("ticks", "<i4"),
                         ("reserved", "<i4")])


class ServeConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_int64), ("velocity_iterations", ctypes.c_int32),
                ("position_iterations", ctypes.c_int32), ("step_flags", ctypes.c_uint32),
                ("realtime", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("busy_s", ctypes.c_double), ("mean_in_flight", ctypes.c_double),
                ("ticks", ctypes.c_int64), ("world_steps", ctypes.c_int64), ("body_steps", ctypes.c_int64),
                ("max_in_flight", ctypes.c_int32), ("completed", ctypes.c_int32), ("device_s", ctypes.c_double)]


class B2GError(RuntimeError):
    pass


def lib():
    """Load the HIP library; raises when it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise B2GError(f"{LIB_PATH} is missing: build it (make -C box2d-optimized_amd)")
        L = ctypes.CDLL(LIB_PATH)
        L.b2g_last_error.restype = ctypes.c_char_p
        L.b2g_device_count.argtypes = [_P]
        L.b2g_batch_create.argtypes = [_P, _P]
        L.b2g_batch_destroy.argtypes = [_P]
        L.b2g_set_worlds.argtypes = [_P, _I, _P, _P, _P]
        L.b2g_set_world_params.argtypes = [_P, _I, _I, _P]
        for f in ("b2g_upload_bodies", "b2g_upload_shapes", "b2g_upload_joints",
                  "b2g_download_bodies", "b2g_download_contacts", "b2g_download_joints",
                  "b2g_download_contact_colours", "b2g_download_solver_points"):
            getattr(L, f).argtypes = [_P, _I, _I, _P]
        L.b2g_upload_contacts.argtypes = [_P, _I, _P]
        L.b2g_batch_step.argtypes = [_P, _I, _I, ctypes.c_uint32]
        L.b2g_batch_wait.argtypes = [_P]
        L.b2g_get_counts.argtypes = [_P, _P]
        L.b2g_get_profile.argtypes = [_P, _P]
        L.b2g_collide_pairs.argtypes = [_I, _P, _P, _P, _P, _P]
        L.b2g_collide_pairs_noalloc.argtypes = [_I, _P, _P, _P, _P, _P]
        L.b2g_find_pairs.argtypes = [_I, _P, _P, _P, _P, _I, _P, ctypes.c_int64, _P, _P]
        L.b2a_world_from_scene.argtypes = [_P, _P, _I]
        L.b2a_world_destroy.argtypes = [_P]
        L.b2a_world_step.argtypes = [_P, _F, _I, _I]
        L.b2a_world_get_bodies.argtypes = [_P, _P]
        L.b2a_world_contact_count.argtypes = [_P, _P]
        L.b2a_world_get_contacts.argtypes = [_P, _I, _P, _P]
        L.b2a_world_batch.restype = _P
        L.b2a_world_batch.argtypes = [_P]
        L.b2a_world_prepare.argtypes = [_P]
        L.b2g_time_kernel.argtypes = [_P, _I]
        L.b2g_kernel_time.argtypes = [_P, _P, _P, _P]
        L.b2a_scene_records.argtypes = [_P, _I, _I, _P, _P, _P]
        L.b2a_batch_from_scenes.argtypes = [_P, _I, _P, _I, _I]
        L.b2s_run.argtypes = [_P, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P]
        L.b2g_set_option.argtypes = [_P, _I, _I]
        L.b2g_upload_contact_colours.argtypes = [_P, _I, _P]
        L.b2g_snapshot_save.argtypes = [_P, ctypes.c_char_p]
        L.b2g_download_events.argtypes = [_P, _I, _P, _P]
        L.b2g_query_aabbs.argtypes = [_P, _I, _P, _P, _I, _P, _P]
        L.b2g_raycast.argtypes = [_P, _I, _P, _P, _I, _P, _P]
        L.b2g_snapshot_load.argtypes = [_P, ctypes.c_char_p]
        L.b2g_get_option.argtypes = [_P, _I, _P]
        L.b2g_set_counts.argtypes = [_P, _I, _I, _I]
        for f in ("b2a_world_destroy_body", "b2a_world_destroy_joint", "b2a_world_destroy_fixture"):
            getattr(L, f).argtypes = [_P, _I]
        L.b2a_world_set_filter.argtypes = [_P, _I, _I, _I, _I]
        L.b2a_world_body_count.argtypes = [_P, _P]
        L.b2a_world_set_allow_sleep.argtypes = [_P, _I]
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise B2GError(f"b2g error {rc}: {lib().b2g_last_error().decode()}")
    return rc


def device_count() -> int:
    n = ctypes.c_int32(0)
    lib().b2g_device_count(ctypes.byref(n))
    return n.value


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def scene_records(scene, world=0, body_offset=0):
    """Device records (bodies, shapes, joints) of a scene, built through the
    b2World API on the host (no device needed)."""
    L = lib()
    cs = scene.to_c()
    bodies = np.zeros(scene.body_count, dtype=BODY_DTYPE)
    shapes = np.zeros(scene.fixture_count, dtype=SHAPE_DTYPE)
    joints = np.zeros(scene.joint_count, dtype=JOINT_DTYPE)
    check(L.b2a_scene_records(cs.ptr(), world, body_offset, _ptr(bodies),
                              _ptr(shapes) if len(shapes) else None, _ptr(joints) if len(joints) else None))
    return bodies, shapes, joints


class Batch:
    """A set of independent worlds in device memory (b2g_batch)."""

    @classmethod
    def from_scenes(cls, scenes, contacts_per_shape=8, warm_starting=True):
        """World i of the batch is scenes[i], laid out world after world."""
        from .scenes import B2gScene
        L = lib()
        cs = [s.to_c() for s in scenes]
        arr = (B2gScene * len(cs))(*[c.struct for c in cs])
        h = ctypes.c_void_p()
        check(L.b2a_batch_from_scenes(ctypes.byref(h), len(cs), ctypes.cast(arr, ctypes.c_void_p),
                                      contacts_per_shape, int(warm_starting)))
        b = cls(handle=h.value)
        b.owned = True
        b.world_bodies = np.cumsum([0] + [s.body_count for s in scenes])
        b.world_shapes = np.cumsum([0] + [s.fixture_count for s in scenes])
        b.world_joints = np.cumsum([0] + [s.joint_count for s in scenes])
        b.world_dynamic = np.array([s.dynamic_count for s in scenes], dtype=np.int32)
        b.gravity = [tuple(s.gravity) for s in scenes]
        del cs, arr
        return b

    def __init__(self, worlds=1, bodies=64, shapes=64, joints=16, contacts=4096, handle=None):
        self.L = lib()
        self.owned = handle is None
        if handle is None:
            cap = Capacity(worlds, bodies, shapes, joints, contacts)
            h = ctypes.c_void_p()
            check(self.L.b2g_batch_create(ctypes.byref(h), ctypes.byref(cap)))
            self.h = h.value
        else:
            self.h = handle

    def close(self):
        if self.h and self.owned:
            self.L.b2g_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, steps=1, dt=1.0 / 60.0, vi=8, pi=3, flags=0, warm=None, sleep=None):
        """Step every world `steps` times the way b2World::Step sequences
        them (b2_world.cpp:1108-1171): the first step of a fresh batch finds
        contacts first and has dtRatio 0; later steps use inv_dt0 * dt.
        `warm`: per-world b2World::SetWarmStarting flags (default all on);
        `sleep`: per-world b2World::SetAllowSleeping flags (default all off)."""
        dt = np.float32(dt)
        nw = len(self.gravity)
        warm = [1] * nw if warm is None else [1 if w else 0 for w in warm]
        sleep = [0] * nw if sleep is None else [1 if z else 0 for z in sleep]
        for _ in range(steps):
            inv_dt0 = getattr(self, "inv_dt0", np.float32(0.0))
            ratio = np.float32(inv_dt0 * dt)
            self.set_world_params(0, [WorldParams(g, dt, ratio, warm[k], 1, sleep[k])
                                      for k, g in enumerate(self.gravity)])
            first = not getattr(self, "stepped", False)
            self.step(vi, pi, flags | (STEP_FIND_NEW_CONTACTS if first else 0))
            self.stepped = True
            self.inv_dt0 = np.float32(np.float32(1.0) / dt) if dt > 0 else np.float32(0.0)
        return self

    def set_worlds(self, params):
        arr = (WorldParams * len(params))(*params)
        check(self.L.b2g_set_worlds(self.h, len(params), None, None, arr))

    def set_world_params(self, first, params):
        arr = (WorldParams * len(params))(*params)
        check(self.L.b2g_set_world_params(self.h, first, len(params), arr))

    def upload_bodies(self, first, rec):
        rec = np.ascontiguousarray(rec, dtype=BODY_DTYPE)
        check(self.L.b2g_upload_bodies(self.h, first, len(rec), _ptr(rec)))

    def upload_shapes(self, first, rec):
        rec = np.ascontiguousarray(rec, dtype=SHAPE_DTYPE)
        check(self.L.b2g_upload_shapes(self.h, first, len(rec), _ptr(rec)))

    def upload_joints(self, first, rec):
        rec = np.ascontiguousarray(rec, dtype=JOINT_DTYPE)
        if len(rec):
            check(self.L.b2g_upload_joints(self.h, first, len(rec), _ptr(rec)))

    def upload_contacts(self, rec):
        rec = np.ascontiguousarray(rec, dtype=CONTACT_DTYPE)
        check(self.L.b2g_upload_contacts(self.h, len(rec), _ptr(rec) if len(rec) else None))

    def step(self, vi=8, pi=3, flags=0):
        check(self.L.b2g_batch_step(self.h, vi, pi, flags))

    def wait(self):
        check(self.L.b2g_batch_wait(self.h))

    def counts(self) -> Counts:
        c = Counts()
        check(self.L.b2g_get_counts(self.h, ctypes.byref(c)))
        return c

    def profile(self) -> Profile:
        p = Profile()
        check(self.L.b2g_get_profile(self.h, ctypes.byref(p)))
        return p

    def save(self, path):
        """Binary snapshot (b2g_snapshot_save)."""
        check(self.L.b2g_snapshot_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, path, like=None):
        """Batch restored from a snapshot; `like` (a Batch) supplies the
        per-world layout metadata the Python side keeps."""
        L = lib()
        h = ctypes.c_void_p()
        check(L.b2g_snapshot_load(ctypes.byref(h), os.fsencode(path)))
        b = cls(handle=h.value)
        b.owned = True
        if like is not None:
            for k in ("world_bodies", "world_shapes", "world_joints", "world_dynamic", "gravity", "inv_dt0"):
                if hasattr(like, k):
                    setattr(b, k, getattr(like, k))
            b.stepped = getattr(like, "stepped", False)
        return b

    def query_aabbs(self, worlds, aabbs, capacity=64):
        """Shapes overlapping each AABB: list of int arrays (b2g_query_aabbs)."""
        w = np.ascontiguousarray(worlds, dtype=np.int32)
        a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(len(w), 4)
        while True:
            out = np.zeros((len(w), capacity), np.int32)
            cnt = np.zeros(len(w), np.int32)
            check(self.L.b2g_query_aabbs(self.h, len(w), _ptr(w), _ptr(a), capacity, _ptr(out), _ptr(cnt)))
            if len(cnt) == 0 or cnt.max() <= capacity:
                return [out[i, :cnt[i]].copy() for i in range(len(w))]
            capacity = int(cnt.max())

    def raycast(self, worlds, segments, capacity=64):
        """All hits of each segment (p1, p2): list of RAY_HIT_DTYPE arrays (b2g_raycast)."""
        w = np.ascontiguousarray(worlds, dtype=np.int32)
        sg = np.ascontiguousarray(segments, dtype=np.float32).reshape(len(w), 4)
        while True:
            out = np.zeros((len(w), capacity), RAY_HIT_DTYPE)
            cnt = np.zeros(len(w), np.int32)
            check(self.L.b2g_raycast(self.h, len(w), _ptr(w), _ptr(sg), capacity, _ptr(out), _ptr(cnt)))
            if len(cnt) == 0 or cnt.max() <= capacity:
                return [out[i, :cnt[i]].copy() for i in range(len(w))]
            capacity = int(cnt.max())

    def events(self):
        """Contact events of the last step taken with STEP_EVENTS."""
        n = ctypes.c_int32(0)
        check(self.L.b2g_download_events(self.h, 0, None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=EVENT_DTYPE)
        if n.value:
            check(self.L.b2g_download_events(self.h, n.value, _ptr(out), ctypes.byref(n)))
        return out

    def set_option(self, option, value):
        check(self.L.b2g_set_option(self.h, option, value))

    def get_option(self, option):
        v = ctypes.c_int32(0)
        check(self.L.b2g_get_option(self.h, option, ctypes.byref(v)))
        return v.value

    def time_kernel(self, kernel):
        check(self.L.b2g_time_kernel(self.h, kernel))

    def kernel_time(self):
        """(total ms, launches, items) of the timed kernel since time_kernel()."""
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        items = ctypes.c_int64(0)
        check(self.L.b2g_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(items)))
        return ms.value, n.value, items.value

    def bodies(self, first=0, count=None):
        if count is None:
            count = self.counts().bodies - first
        out = np.zeros(count, dtype=BODY_DTYPE)
        if count:
            check(self.L.b2g_download_bodies(self.h, first, count, _ptr(out)))
        return out

    def contacts(self):
        n = self.L.b2g_download_contacts(self.h, 0, 0, None)
        if n < 0:
            check(n)
        out = np.zeros(n, dtype=CONTACT_DTYPE)
        if n:
            check(self.L.b2g_download_contacts(self.h, 0, n, _ptr(out)))
        return out

    def set_counts(self, bodies, shapes, joints):
        check(self.L.b2g_set_counts(self.h, bodies, shapes, joints))

    def solver_points(self, n):
        """Points the last step's solver stored per contact (PostSolve's count; 0 = not solved)."""
        out = np.zeros(n, dtype=np.int32)
        if n:
            check(self.L.b2g_download_solver_points(self.h, 0, n, _ptr(out)))
        return out

    def upload_contact_colours(self, colours):
        c = np.ascontiguousarray(colours, dtype=np.int32)
        check(self.L.b2g_upload_contact_colours(self.h, len(c), _ptr(c) if len(c) else None))

    def contact_colours(self, n):
        out = np.zeros(n, dtype=np.int32)
        if n:
            check(self.L.b2g_download_contact_colours(self.h, 0, n, _ptr(out)))
        return out

    def joints(self, count):
        out = np.zeros(count, dtype=JOINT_DTYPE)
        if count:
            check(self.L.b2g_download_joints(self.h, 0, count, _ptr(out)))
        return out


def world_device_bytes(scene, contacts_per_shape=3):
    """Device working memory of one world while it steps: body state (7
    float4 + 2 ints), shapes (geometry, material, filter, AABB, tree and
    sort scratch) and its contacts (two list buffers + velocity and position
    constraints + compaction scratch)."""
    nb, ns = scene.body_count, scene.fixture_count
    return int(nb * (7 * 16 + 8) + ns * (9 * 16 + 3 * 16 + 16 + 120)
               + contacts_per_shape * ns * (2 * (8 + 8 + 64 + 16 + 4) + 10 * 16 + 5 * 16 + 32))


def serve(pool, requests, memory_budget, world_bytes, inv_dt0=None, vi=8, pi=3, graph=False, realtime=True):
    """Continuous-batching request loop (b2s_run) over `pool`, a Batch built
    by Batch.from_scenes that has taken at least one step. `requests` is a
    REQUEST_DTYPE array sorted by arrival. Returns (results, stats, inv_dt0)."""
    L = lib()
    nw = len(pool.gravity)
    req = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
    res = np.zeros(len(req), dtype=RESULT_DTYPE)
    bodies = np.ascontiguousarray(pool.world_dynamic, dtype=np.int32)
    wbytes = np.ascontiguousarray(world_bytes, dtype=np.int64)
    grav = np.ascontiguousarray(np.asarray(pool.gravity, dtype=np.float32).reshape(nw, 2))
    if inv_dt0 is None:  # every world took the pool's last advance() step
        inv_dt0 = np.full(nw, getattr(pool, "inv_dt0", np.float32(0.0)), dtype=np.float32)
    inv = np.ascontiguousarray(inv_dt0, dtype=np.float32)
    cfg = ServeConfig(int(memory_budget), vi, pi, STEP_GRAPH if graph else 0, 1 if realtime else 0)
    st = ServeStats()
    check(L.b2s_run(pool.h, nw, _ptr(bodies), _ptr(wbytes), _ptr(grav), _ptr(inv), len(req),
                    _ptr(req) if len(req) else None, ctypes.byref(cfg), _ptr(res) if len(res) else None,
                    ctypes.byref(st)))
    return res, st, inv


class World:
    """A b2World (drop-in C++ API) built from a scene, stepped on the GPU."""

    def __init__(self, scene, warm_starting=True):
        self.L = lib()
        self.scene = scene
        self.cscene = scene.to_c()
        h = ctypes.c_void_p()
        check(self.L.b2a_world_from_scene(ctypes.byref(h), self.cscene.ptr(), int(warm_starting)))
        self.h = h.value
        self.nb = scene.body_count

    def close(self):
        if self.h:
            self.L.b2a_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1, dt=1.0 / 60.0, vi=8, pi=3):
        for _ in range(n):
            check(self.L.b2a_world_step(self.h, dt, vi, pi))

    def step_status(self, dt=1.0 / 60.0, vi=8, pi=3) -> int:
        """One b2World::Step; returns its status instead of raising."""
        return self.L.b2a_world_step(self.h, dt, vi, pi)

    def destroy_body(self, body):
        check(self.L.b2a_world_destroy_body(self.h, body))
        self._recount()

    def destroy_joint(self, joint):
        check(self.L.b2a_world_destroy_joint(self.h, joint))

    def destroy_fixture(self, fixture):
        check(self.L.b2a_world_destroy_fixture(self.h, fixture))

    def set_filter(self, fixture, category, mask, group):
        check(self.L.b2a_world_set_filter(self.h, fixture, category, mask, group))

    def set_allow_sleep(self, flag):
        """b2World::SetAllowSleeping."""
        check(self.L.b2a_world_set_allow_sleep(self.h, int(flag)))

    def _recount(self):
        n = ctypes.c_int32(0)
        check(self.L.b2a_world_body_count(self.h, ctypes.byref(n)))
        self.nb = n.value

    def bodies(self):
        """10 floats per body: centre c (2), angle, v (2), w, transform (p 2, q.s, q.c)."""
        out = np.zeros((self.nb, 10), dtype=np.float32)
        check(self.L.b2a_world_get_bodies(self.h, _ptr(out)))
        return out

    def contacts(self):
        n = ctypes.c_int32(0)
        check(self.L.b2a_world_get_contacts(self.h, 0, None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=CONTACT_DTYPE)
        if n.value:
            check(self.L.b2a_world_get_contacts(self.h, n.value, _ptr(out), ctypes.byref(n)))
        return out

    def prepare(self):
        """Upload everything to the device without stepping."""
        check(self.L.b2a_world_prepare(self.h))

    def batch(self) -> Batch:
        return Batch(handle=self.L.b2a_world_batch(self.h))


def collide_pairs(shape_a, xf_a, shape_b, xf_b, noalloc=False):
    """Narrowphase manifolds of n shape pairs (b2g_collide_pairs). With
    noalloc, b2g_collide_pairs_noalloc: no device allocation, at most
    NOALLOC_MAX_PAIRS pairs per call."""
    sa = np.ascontiguousarray(shape_a, dtype=SHAPE_DTYPE)
    sb = np.ascontiguousarray(shape_b, dtype=SHAPE_DTYPE)
    xa = np.ascontiguousarray(xf_a, dtype=np.float32).reshape(-1, 4)
    xb = np.ascontiguousarray(xf_b, dtype=np.float32).reshape(-1, 4)
    n = len(sa)
    out = np.zeros(n, dtype=CONTACT_DTYPE)
    if n or noalloc:
        fn = lib().b2g_collide_pairs_noalloc if noalloc else lib().b2g_collide_pairs
        check(fn(n, _ptr(sa), _ptr(xa), _ptr(sb), _ptr(xb), _ptr(out)))
    return out


NOALLOC_MAX_PAIRS = 4096


def find_pairs(aabbs, shape_body, shape_type, filt, body_type, max_pairs=None):
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32)
    sb = np.ascontiguousarray(shape_body, dtype=np.int32)
    st = np.ascontiguousarray(shape_type, dtype=np.int32)
    bt = np.ascontiguousarray(body_type, dtype=np.int32)
    f = None if filt is None else np.ascontiguousarray(filt, dtype=np.int32)
    n = len(sb)
    if max_pairs is None:
        max_pairs = max(64, 32 * n)
    keys = np.zeros(max_pairs, dtype=np.uint64)
    cnt = ctypes.c_int64(0)
    check(lib().b2g_find_pairs(n, _ptr(aabbs), _ptr(sb), _ptr(st), None if f is None else _ptr(f),
                               len(bt), _ptr(bt), max_pairs, _ptr(keys), ctypes.byref(cnt)))
    return keys[:min(cnt.value, max_pairs)], cnt.value
