"""Seeded scene generators for the b2World::Step workloads.

Every side (the HIP engine, the CPU restatement under oracle/ and the compiled
reference) receives the same float32 arrays from these functions, so a scene
is bit-identical everywhere. Randomness comes from splitmix64 (the k-th draw
of a stream is mix(seed + k * 0x9E3779B97F4A7C15), so whole streams are
generated with numpy); a uniform float is (u64 >> 40) * 2**-24 and a normal
is Box-Muller on two uniforms.

Scenes follow the reference's construction conventions:
  * pyramid: testbed/tests/pyramid.cpp:33-72 (edge ground (-40,0)-(40,0),
    SetAsBox(0.5,0.5), row origin (-7,0.75) + i*(0.5625,1.25), spacing 1.125)
  * polygons: b2PolygonShape::Set over k points on a circle
    (src/collision/b2_polygon_shape.cpp:113-243 welds and hulls them)
  * joints: b2RevoluteJointDef::Initialize(bodyA, bodyB, anchor)
    (src/dynamics/b2_revolute_joint.cpp:41-48)
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

CIRCLE, EDGE, POLYGON, BOX = 0, 1, 2, 3
STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
GEOM = 17


class Stream:
    """splitmix64 stream; draws are consumed in order."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.k = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            ks = np.arange(self.k + 1, self.k + n + 1, dtype=np.uint64)
            z = self.seed + ks * GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.k += n
        return z

    def uniform(self, n: int) -> np.ndarray:
        """float64 values exactly representable in float32, in [0, 1)."""
        return (self.u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)

    def normal(self, n: int) -> np.ndarray:
        u1 = 1.0 - self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Scene:
    """Flat arrays in creation order (mirrors the b2g_scene C struct)."""

    gravity: tuple = (0.0, -10.0)
    body_type: list = field(default_factory=list)
    body_pos: list = field(default_factory=list)
    body_angle: list = field(default_factory=list)
    body_vel: list = field(default_factory=list)
    fix_body: list = field(default_factory=list)
    fix_kind: list = field(default_factory=list)
    fix_geom: list = field(default_factory=list)
    fix_material: list = field(default_factory=list)
    fix_filter: list = field(default_factory=list)
    joint_bodies: list = field(default_factory=list)
    joint_params: list = field(default_factory=list)
    joint_flags: list = field(default_factory=list)
    name: str = ""
    # per body: bit0 fixedRotation, bit1 bullet, bit2 !allowSleep, bit3 !awake
    # (b2BodyDef); empty = all 0
    body_flags: list = field(default_factory=list)

    # ---- building -------------------------------------------------------
    def add_body(self, btype, x, y, angle=0.0, vx=0.0, vy=0.0, w=0.0, flags=0) -> int:
        if flags and not self.body_flags:
            self.body_flags = [0] * len(self.body_type)
        self.body_type.append(btype)
        self.body_pos.append((x, y))
        self.body_angle.append(angle)
        self.body_vel.append((vx, vy, w))
        if self.body_flags:
            self.body_flags.append(flags)
        return len(self.body_type) - 1

    def _add_fix(self, body, kind, geom, density, friction, restitution=0.0,
                 threshold=1.0, filt=(1, 0xFFFF, 0, 0)):
        g = np.zeros(GEOM, dtype=np.float32)
        g[: len(geom)] = geom
        self.fix_body.append(body)
        self.fix_kind.append(kind)
        self.fix_geom.append(g)
        self.fix_material.append((density, friction, restitution, threshold))
        self.fix_filter.append(filt)
        return len(self.fix_body) - 1

    def add_circle(self, body, radius, cx=0.0, cy=0.0, density=1.0, friction=0.6, **kw):
        return self._add_fix(body, CIRCLE, [radius, cx, cy], density, friction, **kw)

    def add_edge(self, body, v1, v2, v0=None, v3=None, density=0.0, friction=0.6, **kw):
        one_sided = v0 is not None
        v0 = v0 if v0 is not None else (0.0, 0.0)
        v3 = v3 if v3 is not None else (0.0, 0.0)
        geom = [1.0 if one_sided else 0.0, *v0, *v1, *v2, *v3]
        return self._add_fix(body, EDGE, geom, density, friction, **kw)

    def add_polygon(self, body, points, density=1.0, friction=0.6, **kw):
        pts = np.asarray(points, dtype=np.float32).reshape(-1, 2)
        assert 3 <= len(pts) <= 8
        return self._add_fix(body, POLYGON, [len(pts), *pts.ravel()], density, friction, **kw)

    def add_box(self, body, hx, hy, center=None, angle=0.0, density=1.0, friction=0.6, **kw):
        if center is None:
            geom = [hx, hy, 0.0, 0.0, 0.0, 0.0]
        else:
            geom = [hx, hy, center[0], center[1], angle, 1.0]
        return self._add_fix(body, BOX, geom, density, friction, **kw)

    def add_revolute(self, body_a, body_b, anchor, collide_connected=False,
                     limit=None, motor=None):
        flags = (1 if collide_connected else 0)
        lower = upper = 0.0
        speed = torque = 0.0
        if limit is not None:
            flags |= 2
            lower, upper = limit
        if motor is not None:
            flags |= 4
            speed, torque = motor
        self.joint_bodies.append((body_a, body_b))
        self.joint_params.append((anchor[0], anchor[1], lower, upper, speed, torque))
        self.joint_flags.append(flags)
        return len(self.joint_flags) - 1

    # ---- views -----------------------------------------------------------
    @property
    def body_count(self):
        return len(self.body_type)

    @property
    def fixture_count(self):
        return len(self.fix_body)

    @property
    def joint_count(self):
        return len(self.joint_flags)

    @property
    def dynamic_count(self):
        return int(np.sum(np.asarray(self.body_type) == DYNAMIC))

    def arrays(self) -> dict:
        n, m, j = self.body_count, self.fixture_count, self.joint_count
        return dict(
            gravity=np.asarray(self.gravity, dtype=np.float32),
            body_type=np.asarray(self.body_type, dtype=np.int32).reshape(n),
            body_pos=np.asarray(self.body_pos, dtype=np.float32).reshape(n, 2),
            body_angle=np.asarray(self.body_angle, dtype=np.float32).reshape(n),
            body_vel=np.asarray(self.body_vel, dtype=np.float32).reshape(n, 3),
            fix_body=np.asarray(self.fix_body, dtype=np.int32).reshape(m),
            fix_kind=np.asarray(self.fix_kind, dtype=np.int32).reshape(m),
            fix_geom=np.asarray(self.fix_geom, dtype=np.float32).reshape(m, GEOM),
            fix_material=np.asarray(self.fix_material, dtype=np.float32).reshape(m, 4),
            fix_filter=np.asarray(self.fix_filter, dtype=np.int32).reshape(m, 4),
            joint_bodies=np.asarray(self.joint_bodies, dtype=np.int32).reshape(j, 2),
            joint_params=np.asarray(self.joint_params, dtype=np.float32).reshape(j, 6),
            joint_flags=np.asarray(self.joint_flags, dtype=np.int32).reshape(j),
            body_flags=np.asarray(self.body_flags if self.body_flags else [0] * n, dtype=np.int32).reshape(n),
        )

    def to_c(self):
        """A B2gScene ctypes struct; the returned holder keeps arrays alive."""
        return CScene(self)


class B2gScene(ctypes.Structure):
    _fields_ = [
        ("gravity", ctypes.c_float * 2),
        ("body_count", ctypes.c_int32),
        ("body_type", ctypes.c_void_p),
        ("body_pos", ctypes.c_void_p),
        ("body_angle", ctypes.c_void_p),
        ("body_vel", ctypes.c_void_p),
        ("body_damping", ctypes.c_void_p),
        ("body_gravity_scale", ctypes.c_void_p),
        ("body_flags", ctypes.c_void_p),
        ("fixture_count", ctypes.c_int32),
        ("fix_body", ctypes.c_void_p),
        ("fix_kind", ctypes.c_void_p),
        ("fix_geom", ctypes.c_void_p),
        ("fix_material", ctypes.c_void_p),
        ("fix_filter", ctypes.c_void_p),
        ("joint_count", ctypes.c_int32),
        ("joint_bodies", ctypes.c_void_p),
        ("joint_params", ctypes.c_void_p),
        ("joint_flags", ctypes.c_void_p),
    ]


class CScene:
    def __init__(self, scene: Scene):
        self.arrays = {k: np.ascontiguousarray(v) for k, v in scene.arrays().items()}
        a = self.arrays
        s = B2gScene()
        s.gravity[0], s.gravity[1] = float(a["gravity"][0]), float(a["gravity"][1])
        s.body_count = scene.body_count
        s.body_type = a["body_type"].ctypes.data
        s.body_pos = a["body_pos"].ctypes.data
        s.body_angle = a["body_angle"].ctypes.data
        s.body_vel = a["body_vel"].ctypes.data
        s.body_damping = None
        s.body_gravity_scale = None
        s.body_flags = a["body_flags"].ctypes.data if scene.body_flags else None
        s.fixture_count = scene.fixture_count
        s.fix_body = a["fix_body"].ctypes.data
        s.fix_kind = a["fix_kind"].ctypes.data
        s.fix_geom = a["fix_geom"].ctypes.data
        s.fix_material = a["fix_material"].ctypes.data
        s.fix_filter = a["fix_filter"].ctypes.data
        s.joint_count = scene.joint_count
        s.joint_bodies = a["joint_bodies"].ctypes.data if scene.joint_count else None
        s.joint_params = a["joint_params"].ctypes.data if scene.joint_count else None
        s.joint_flags = a["joint_flags"].ctypes.data if scene.joint_count else None
        self.struct = s

    def ptr(self):
        return ctypes.byref(self.struct)


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def pyramid(rows: int = 20, density: float = 1.0, friction: float = 0.6,
            origin=(0.0, 0.0), ground_half=40.0, scene: Scene | None = None) -> Scene:
    """Plain single world: edge ground and a `rows`-row box pyramid.

    Follows testbed/tests/pyramid.cpp:33-72 (float32 accumulation of the row
    and column offsets), with density and friction set explicitly.
    """
    s = scene if scene is not None else Scene(name=f"pyramid{rows}")
    ox, oy = np.float32(origin[0]), np.float32(origin[1])
    g = s.add_body(STATIC, float(ox), float(oy))
    s.add_edge(g, (-ground_half, 0.0), (ground_half, 0.0), friction=friction)
    x = np.array([-7.0, 0.75], dtype=np.float32) + np.array([ox, oy], dtype=np.float32)
    dx = np.array([0.5625, 1.25], dtype=np.float32)
    dy = np.array([1.125, 0.0], dtype=np.float32)
    for i in range(rows):
        y = x.copy()
        for _ in range(i, rows):
            b = s.add_body(DYNAMIC, float(y[0]), float(y[1]))
            s.add_box(b, 0.5, 0.5, density=density, friction=friction)
            y = (y + dy).astype(np.float32)
        x = (x + dx).astype(np.float32)
    return s


def _walls(s: Scene, half_w: float, height: float, y0: float = 0.0, friction=0.6, x0=0.0):
    g = s.add_body(STATIC, x0, y0)
    s.add_edge(g, (-half_w, 0.0), (half_w, 0.0), friction=friction)
    s.add_edge(g, (-half_w, 0.0), (-half_w, height), friction=friction)
    s.add_edge(g, (half_w, 0.0), (half_w, height), friction=friction)
    s.add_edge(g, (-half_w, height), (half_w, height), friction=friction)
    return g


def random_pile(n: int = 20000, seed: int = 0xC2, half_w: float = 100.0,
                height: float = 200.0, fill_from: float = 0.25,
                scene: Scene | None = None, x0: float = 0.0) -> Scene:
    """Large random pile: n dynamic shapes in a static box of 4 edges.

    60% convex polygons (k in 3..8 points on a circle of radius U[0.25,1.0],
    sorted random angles, hulled by b2PolygonShape::Set), 40% circles of
    radius U[0.25,0.75]; centres uniform in the upper part of the box, body
    angle U[0,2pi); density 1, friction 0.6, restitution 0. Fixture i is a
    polygon when i % 5 < 3.
    """
    s = scene if scene is not None else Scene(name=f"pile{n}")
    _walls(s, half_w, height, x0=x0)
    rs = Stream(seed)
    margin = 1.0
    xs = rs.uniform(n) * (2 * (half_w - margin)) - (half_w - margin) + x0
    ys = rs.uniform(n) * (height * (1.0 - fill_from) - 2 * margin) + height * fill_from + margin
    angles = rs.uniform(n) * (2.0 * math.pi)
    kinds = rs.uniform(n)  # polygon vertex count draw
    radii = rs.uniform(n)
    ang_pts = rs.uniform(n * 8).reshape(n, 8) * (2.0 * math.pi)
    # bodies (positions and angles rounded to float32, as b2BodyDef holds them)
    b0 = s.body_count
    s.body_type.extend([DYNAMIC] * n)
    s.body_pos.extend(zip(xs.astype(np.float32).astype(np.float64).tolist(),
                          ys.astype(np.float32).astype(np.float64).tolist()))
    s.body_angle.extend(angles.astype(np.float32).astype(np.float64).tolist())
    s.body_vel.extend([(0.0, 0.0, 0.0)] * n)
    # fixture i: polygon when i % 5 < 3 (k points at sorted angles on a circle
    # of radius r), else a circle
    idx = np.arange(n)
    poly = idx % 5 < 3
    k = 3 + (kinds * 6.0).astype(np.int64)
    r_poly = 0.25 + 0.75 * radii
    th = np.sort(np.where(np.arange(8)[None, :] < k[:, None], ang_pts, np.inf), axis=1)
    th = np.where(np.isfinite(th), th, 0.0)
    pts = np.stack([(r_poly[:, None] * np.cos(th)).astype(np.float32),
                    (r_poly[:, None] * np.sin(th)).astype(np.float32)], axis=2).reshape(n, 16)
    geom = np.zeros((n, GEOM), np.float32)
    geom[:, 0] = k
    geom[:, 1:17] = np.where((np.arange(16)[None, :] // 2) < k[:, None], pts, 0.0)
    circ = np.zeros((n, GEOM), np.float32)
    circ[:, 0] = (0.25 + 0.5 * radii).astype(np.float32)
    geom = np.where(poly[:, None], geom, circ)
    s.fix_body.extend((b0 + idx).tolist())
    s.fix_kind.extend(np.where(poly, POLYGON, CIRCLE).tolist())
    s.fix_geom.extend(list(geom))
    s.fix_material.extend([(1.0, 0.6, 0.0, 1.0)] * n)
    s.fix_filter.extend([(1, 0xFFFF, 0, 0)] * n)
    return s


def gas_and_chains(n_circles: int = 100000, n_chains: int = 500, links: int = 30,
                   seed: int = 0xC3, half: float = 500.0) -> Scene:
    """Broadphase-heavy world with joints, zero gravity.

    n_circles circles of radius 0.5*(1+0.1*U[-1,1]) with velocity N(0,2) per
    axis inside a static 2*half box of 4 edges, plus n_chains chains of
    `links` 0.2x1 m boxes joined end to end by revolute joints, link 0 pinned
    to the static body (collideConnected false).
    """
    s = Scene(name=f"gas{n_circles}_chains{n_chains}x{links}", gravity=(0.0, 0.0))
    g = s.add_body(STATIC, 0.0, -half)
    h2 = 2 * half
    s.add_edge(g, (-half, 0.0), (half, 0.0))
    s.add_edge(g, (-half, 0.0), (-half, h2))
    s.add_edge(g, (half, 0.0), (half, h2))
    s.add_edge(g, (-half, h2), (half, h2))
    rs = Stream(seed)
    m = 2.0
    xs = rs.uniform(n_circles) * (h2 - 2 * m) - (half - m)
    ys = rs.uniform(n_circles) * (h2 - 2 * m) - (half - m)
    rr = 0.5 * (1.0 + 0.1 * (2.0 * rs.uniform(n_circles) - 1.0))
    vx = 2.0 * rs.normal(n_circles)
    vy = 2.0 * rs.normal(n_circles)
    for i in range(n_circles):
        b = s.add_body(DYNAMIC, float(np.float32(xs[i])), float(np.float32(ys[i])), 0.0,
                       float(np.float32(vx[i])), float(np.float32(vy[i])))
        s.add_circle(b, float(np.float32(rr[i])), density=1.0, friction=0.6)
    # chains on a grid of anchors, links along +x
    cols = int(math.ceil(math.sqrt(n_chains)))
    pitch_x = (h2 - 2 * m - links - 2) / max(cols, 1)
    pitch_y = (h2 - 2 * m) / max(cols, 1)
    for c in range(n_chains):
        ax = -half + m + 1.0 + (c % cols) * pitch_x
        ay = -half + m + 0.5 * pitch_y + (c // cols) * pitch_y
        prev = g
        for k in range(links):
            b = s.add_body(DYNAMIC, float(np.float32(ax + 0.5 + k)), float(np.float32(ay)))
            s.add_box(b, 0.5, 0.1, density=1.0, friction=0.6)
            s.add_revolute(prev, b, (float(np.float32(ax + k)), float(np.float32(ay))))
            prev = b
    return s


def batched_sizes(worlds: int = 4096, seed: int = 0xC4, median: float = 200.0,
                  sigma: float = 0.95, lo: int = 10, hi: int = 4000) -> np.ndarray:
    rs = Stream(seed)
    z = rs.normal(worlds)
    return np.clip(np.round(median * np.exp(sigma * z)), lo, hi).astype(np.int64)


def batch_world(index: int, size: int, seed_base: int = 0xC4) -> Scene:
    """World `index` of the batched workload: a pyramid (rows 5..40) or a
    random pile with about `size` dynamic bodies."""
    rs = Stream(seed_base + index)
    kind = rs.uniform(1)[0]
    if kind < 0.3:
        rows = int(max(5, min(40, math.floor((math.sqrt(8 * size + 1) - 1) / 2))))
        s = pyramid(rows)
        s.name = f"w{index}_pyramid{rows}"
        return s
    half_w = max(6.0, math.sqrt(size * 3.0) * 0.6)
    s = random_pile(int(size), seed=seed_base * 7919 + index, half_w=half_w,
                    height=2.0 * half_w + 10.0)
    s.name = f"w{index}_pile{size}"
    return s


def serving_sizes(worlds: int = 10000, seed: int = 0xC5, lo: int = 16, hi: int = 4096) -> np.ndarray:
    """Resident pool of the serving workload: body counts log-uniform in
    [lo, hi]."""
    rs = Stream(seed)
    return np.round(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * rs.uniform(worlds))).astype(np.int64)


def serving_trace(worlds: int, rate: float = 20000.0, duration: float = 2.0, seed: int = 0xC5,
                  dts=(1.0 / 60.0, 1.0 / 240.0)):
    """Poisson request arrivals at `rate` per second over `duration` seconds:
    each request names a uniform world, 1-8 substeps and dt 1/60 or 1/240.
    Returns a structured array (arrival, world, substeps, dt, reserved)."""
    rs = Stream(seed ^ 0x5E5)
    n_max = int(rate * duration * 1.5 + 64)
    gaps = -np.log(np.maximum(rs.uniform(n_max), 1e-300)) / rate
    t = np.cumsum(gaps)
    n = int(np.searchsorted(t, duration))
    u = rs.uniform(3 * n).reshape(n, 3) if n else np.zeros((0, 3))
    out = np.zeros(n, dtype=[("arrival", "<f8"), ("world", "<i4"), ("substeps", "<i4"), ("dt", "<f4"),
                             ("reserved", "<i4")])
    out["arrival"] = t[:n]
    out["world"] = np.minimum((u[:, 0] * worlds).astype(np.int64), worlds - 1)
    out["substeps"] = 1 + np.minimum((u[:, 1] * 8).astype(np.int64), 7)
    out["dt"] = np.where(u[:, 2] < 0.5, np.float32(dts[0]), np.float32(dts[1]))
    return out


def small_mixed(seed: int = 7, n: int = 60) -> Scene:
    """A small mixed scene (circles, polygons, boxes, one-sided edges) for
    fast parity checks."""
    s = Scene(name=f"mixed{n}")
    g = s.add_body(STATIC, 0.0, 0.0)
    s.add_edge(g, (-20.0, 0.0), (20.0, 0.0))
    s.add_edge(g, (-20.0, 0.0), (-20.0, 30.0))
    s.add_edge(g, (20.0, 0.0), (20.0, 30.0))
    # a one-sided ramp with ghost vertices
    s.add_edge(g, (-10.0, 5.0), (0.0, 2.0), v0=(-15.0, 6.0), v3=(5.0, 2.0))
    rs = Stream(seed)
    u = rs.uniform(n * 12).reshape(n, 12)
    for i in range(n):
        x = -18.0 + 36.0 * u[i, 0]
        y = 1.0 + 25.0 * u[i, 1]
        b = s.add_body(DYNAMIC, float(np.float32(x)), float(np.float32(y)),
                       float(np.float32(u[i, 2] * 6.28)))
        kind = i % 3
        if kind == 0:
            s.add_circle(b, float(np.float32(0.3 + 0.4 * u[i, 3])))
        elif kind == 1:
            s.add_box(b, float(np.float32(0.3 + 0.5 * u[i, 3])), float(np.float32(0.3 + 0.5 * u[i, 4])))
        else:
            k = 3 + int(u[i, 3] * 6)
            r = 0.3 + 0.6 * u[i, 4]
            th = np.sort(u[i, 4:4 + k] * 2 * math.pi)
            pts = np.stack([r * np.cos(th), r * np.sin(th)], axis=1).astype(np.float32)
            s.add_polygon(b, pts)
    return s


WORKLOADS = {
    "pyramid20": lambda: pyramid(20),
    "pile20k": lambda: random_pile(20000, seed=0xC2),
    "gas100k_chains500": lambda: gas_and_chains(),
}
