"""CPU tests that pin the oracle (oracle/restate.cpp) before it is trusted.

Each check is against something the restatement did not produce itself:
golden vectors written by the compiled reference (tests/golden/), the
reference library where it is built, brute force, or closed forms.
"""
import math

import numpy as np
import pytest

import oracle
from util import bits, golden, load_pkg, manifold_fields_equal, scene_for

S = load_pkg().scenes


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build(ref=True)


def test_narrowphase_matches_reference_golden_bitexact():
    g = golden("narrowphase.npz")
    man = g["manifold"]
    got = np.array([oracle.orc_collide(a, xa, b, xb) for a, xa, b, xb in
                    zip(g["shape_a"], g["xf_a"], g["shape_b"], g["xf_b"])])
    ok = manifold_fields_equal(man, got)
    assert ok.all(), f"{(~ok).sum()} of {len(ok)} manifolds differ"
    # every shape-pair kind is represented, touching and separated
    kinds = set(zip(g["shape_a"]["type"].tolist(), g["shape_b"]["type"].tolist()))
    assert kinds == {(0, 0), (2, 0), (2, 2), (1, 0), (1, 2)}
    assert 0 < (man["point_count"] > 0).sum() < len(man)
    assert (man["point_count"] == 2).sum() > 50


def test_world_manifold_matches_reference_golden_bitexact():
    g = golden("narrowphase.npz")
    man, wm = g["manifold"], g["world_manifold"]
    for i in np.where(man["point_count"] > 0)[0]:
        out = oracle.orc_world_manifold(man[i], g["xf_a"][i], g["shape_a"][i]["radius"], g["xf_b"][i],
                                        g["shape_b"][i]["radius"])
        n = man[i]["point_count"]
        sel = [0, 1, 2, 3, 6] if n == 1 else list(range(8))
        assert np.array_equal(bits(out[sel]), bits(wm[i][sel])), i


def test_shape_setup_and_mass_match_reference_golden_bitexact():
    g = golden("shapes.npz")
    for name, sc in {"mixed90": S.small_mixed(seed=3, n=90),
                     "pile300": S.random_pile(300, seed=77, half_w=20, height=40)}.items():
        w = oracle.RestatedWorld(sc)
        got = w.shapes()
        ref = g[f"{name}_shapes"]
        for f in ("type", "count", "radius", "centroid", "vertices", "normals"):
            assert np.array_equal(bits(got[f]), bits(ref[f])), (name, f)
        body = w.bodies()
        mass = g[f"{name}_mass"]
        assert np.array_equal(bits(body["local_center"]), bits(mass[:, 2:4])), name
        dyn = mass[:, 0] > 0
        inv = np.zeros_like(mass[:, 0])
        inv[dyn] = np.float32(1.0) / mass[dyn, 0]
        assert np.array_equal(bits(body["inv_mass"]), bits(inv)), name
        assert np.allclose(body["inv_i"][dyn], 1.0 / mass[dyn, 1], rtol=1e-5), name
        rb = g[f"{name}_bodies"]
        assert np.array_equal(bits(body["c"]), bits(rb[:, 0:2])), name


@pytest.mark.parametrize("name", ["pyramid", "mixed", "chain", "pile", "pyramid_settled", "pile_settled"])
def test_one_step_in_reference_order_is_bitexact(name):
    """From the reference's state after N steps, one restated step that visits
    the constraints in the reference's island order reproduces the
    reference's state after N+1 steps bit for bit (bodies and manifolds)."""
    sc, _ = scene_for(name)
    g = golden(f"onestep_{name}.npz")
    w = oracle.RestatedWorld(sc)
    bodies = w.bodies()
    s0 = g["state0"]
    bodies["c"], bodies["a"], bodies["v"], bodies["w"], bodies["xf"] = s0[:, 0:2], s0[:, 2], s0[:, 3:5], s0[:, 5], s0[:, 6:10]
    w.set_bodies(bodies)
    w.set_contacts(g["contacts0"])
    if sc.joint_count:
        j = w.joints()
        j["impulse"] = g["joints0"][:, :2]
        w.set_joints(j)
    w.set_inv_dt0(np.float32(1.0) / np.float32(1.0 / 60.0))
    keys = g["order_keys"]
    jrank = np.zeros(sc.joint_count, np.int32)
    jrank[g["order_joints"]] = np.arange(len(g["order_joints"]))
    w.set_colours(keys, np.arange(len(keys), dtype=np.int32), jrank if sc.joint_count else None)
    w.step(1)
    s1 = g["state1"]
    got = w.bodies()
    assert np.array_equal(bits(got["c"]), bits(s1[:, 0:2]))
    assert np.array_equal(bits(got["a"]), bits(s1[:, 2]))
    assert np.array_equal(bits(got["v"]), bits(s1[:, 3:5]))
    assert np.array_equal(bits(got["w"]), bits(s1[:, 5]))
    ref_c = g["contacts1"]
    rk = oracle.contact_keys(ref_c)
    o = np.argsort(rk)
    oc = w.contacts()
    assert np.array_equal(rk[o], oracle.contact_keys(oc))
    same_orient = ref_c[o]["shape_a"] == oc["shape_a"]
    ok = manifold_fields_equal(ref_c[o][same_orient], oc[same_orient])
    assert ok.all()
    # orientation differs only for brand-new same-type contacts (empty manifolds)
    assert (oc[~same_orient]["point_count"] == 0).all()
    if sc.joint_count:
        ji = w.joints()
        assert np.array_equal(bits(ji["impulse"]), bits(g["joints1"][:, :2]))


def _sleep_scene():
    import sys
    from util import GOLDEN
    sys.path.insert(0, GOLDEN)
    import make_golden
    return make_golden.sleep_scene()


def test_sleeping_steps_in_reference_order_are_bitexact():
    """Sleeping on (b2World::SetAllowSleeping(true)), 420 steps in lock-step
    with the reference's island order, its contact-list order (Collide's
    walk: a body woken by a contact counts at once for the contacts after it)
    and its fixture order for each new contact (the reference's free
    choices; the restatement takes them from the fixture): every body's
    awake flag after every step and the
    body states every 5 steps equal the reference's bit for bit.
    The run covers islands falling asleep together (b2_island.cpp:420-454),
    a body that may not sleep, a body created asleep and woken by a new touch
    (b2_contact.cpp:224-228), a joint island, and a sleeping island woken
    when a falling ball reaches it (the depth-first search of
    b2_world.cpp:560-575)."""
    g = golden("sleep_trajectory.npz")
    sc = _sleep_scene()
    w = oracle.RestatedWorld(sc)
    w.set_allow_sleep(True)
    ko, jo, no, every = g["order_offsets"], g["joint_offsets"], g["new_offsets"], int(g["state_every"])
    wo = g["walk_offsets"]
    aw = g["awake"].astype(bool)
    # the fixture holds the events named above: everything asleep but the
    # sleepless box and the falling ball, then the pyramid awake again
    assert aw[0].sum() == sc.body_count - 3 and aw[99].sum() == 4 and aw[-1].sum() > 30
    for k in range(int(g["steps"])):
        keys = g["order_keys"][ko[k]:ko[k + 1]]
        jr = np.zeros(sc.joint_count, np.int32)
        jr[g["order_joints"][jo[k]:jo[k + 1]]] = np.arange(jo[k + 1] - jo[k], dtype=np.int32)
        w.set_colours(keys, np.arange(len(keys), dtype=np.int32), jr)
        listed = oracle.contact_keys(w.contacts())
        assert len(listed) == wo[k + 1] - wo[k]
        w.set_collide_order(listed[g["walk_ranks"][wo[k]:wo[k + 1]]])
        w.step(1)
        c = w.contacts()
        row = dict(zip(oracle.contact_keys(c).tolist(), range(len(c))))
        for key, a in zip(g["new_keys"][no[k]:no[k + 1]].tolist(), g["new_shape_a"][no[k]:no[k + 1]]):
            i = row[key]
            if c["shape_a"][i] != a:
                assert c["point_count"][i] == 0  # only a new, empty contact is turned round
                c["shape_a"][i], c["shape_b"][i] = c["shape_b"][i], c["shape_a"][i]
        assert len(c) == len(row) and len(set(row)) == len(c)
        w.set_contacts(c)
        b = w.bodies()
        awake = (b["flags"] & 1) == 0
        awake[np.asarray(sc.body_type) == S.STATIC] = False  # b2Body::IsAwake of a static body is false
        assert np.array_equal(awake, aw[k]), f"step {k + 1}: awake flags differ at {np.nonzero(awake != aw[k])[0]}"
        if (k + 1) % every == 0:
            st = g["states"][(k + 1) // every - 1]
            got = np.concatenate([b["c"], b["a"][:, None], b["v"], b["w"][:, None], b["xf"]], axis=1)
            assert np.array_equal(bits(got), bits(st)), f"step {k + 1}: body states differ"


@pytest.mark.parametrize("name", ["pile2000", "filtered", "mixed"])
def test_pair_set_matches_reference_golden(name):
    g = golden("pairs.npz")
    keys = oracle.orc_find_pairs(g[f"{name}_aabbs"], g[f"{name}_shape_body"], g[f"{name}_shape_type"],
                                 g[f"{name}_filter"], g[f"{name}_body_type"])
    assert np.array_equal(keys, g[f"{name}_keys"])


def test_reference_golden_excludes_exactly_the_joint_connected_pairs():
    """The chains golden pins the joint rule of the pair filter: the
    reference's pair set equals every allowed overlapping pair (the
    restatement's pair finder, itself checked against brute force below)
    less the pairs of bodies a joint with collideConnected false connects
    (b2Body::ShouldCollide, src/dynamics/b2_body.cpp:459-480) - and the scene
    has hundreds of those, one at every chain joint."""
    g = golden("pairs.npz")
    allp = oracle.orc_find_pairs(g["chains_aabbs"], g["chains_shape_body"], g["chains_shape_type"],
                                 g["chains_filter"], g["chains_body_type"])
    body = g["chains_shape_body"]
    joined = {(min(a, b), max(a, b)) for (a, b), f in zip(g["chains_joint_bodies"].tolist(),
                                                         g["chains_joint_flags"].tolist()) if not (f & 1)}
    ba = body[(allp >> np.uint64(32)).astype(np.int64)]
    bb = body[(allp & np.uint64(0xffffffff)).astype(np.int64)]
    jointed = np.array([(min(x, y), max(x, y)) in joined for x, y in zip(ba.tolist(), bb.tolist())], bool)
    assert jointed.sum() >= 16 * 29
    assert np.array_equal(allp[~jointed], g["chains_keys"])


def test_pair_set_matches_brute_force():
    rs = S.Stream(99)
    n = 700
    lo = (rs.uniform(2 * n).reshape(n, 2) * 40.0).astype(np.float32)
    ext = (0.2 + rs.uniform(2 * n).reshape(n, 2) * 2.0).astype(np.float32)
    # snap some bounds so inclusive touching (==) cases occur
    lo[::7, 0] = np.round(lo[::7, 0])
    hi = (lo + ext).astype(np.float32)
    hi[1::7, 0] = np.round(hi[1::7, 0])
    aabbs = np.concatenate([lo, hi], axis=1)
    body = np.arange(n, dtype=np.int32) // 2          # fixture pairs share a body
    stype = (np.arange(n) % 3).astype(np.int32)
    btype = np.where(np.arange(n // 2) % 5 == 0, 0, 2).astype(np.int32)
    keys = oracle.orc_find_pairs(aabbs, body, stype, None, btype)
    expect = []
    for i in range(n):
        for j in range(i + 1, n):
            a, b = aabbs[i], aabbs[j]
            if not (a[2] >= b[0] and a[0] <= b[2] and a[3] >= b[1] and a[1] <= b[3]):
                continue
            if body[i] == body[j] or (btype[body[i]] != 2 and btype[body[j]] != 2):
                continue
            if stype[i] == 1 and stype[j] == 1:
                continue
            expect.append((i << 32) | j)
    assert np.array_equal(keys, np.array(sorted(expect), dtype=np.uint64))


def test_block_solver_satisfies_complementarity():
    """The 2-point block LCP (b2_contact_solver.cpp:389-631) on random SPD
    systems: when a case is taken, x >= 0, w = K x + b >= 0 and x.w = 0; the
    case matches brute-force enumeration in float64."""
    rng = np.random.default_rng(5)
    taken = 0
    for _ in range(3000):
        m = rng.uniform(0.1, 2.0)
        ra = rng.normal(size=2)
        rb = rng.normal(size=2)
        k11 = m + ra[0] ** 2
        k22 = m + ra[1] ** 2
        k12 = m + ra[0] * ra[1] * rng.uniform(-1, 1)
        if k11 * k22 - k12 * k12 <= 1e-3:
            continue
        K = np.array([k11, k12, k22], np.float32)
        b = rng.normal(size=2).astype(np.float32)
        a = np.zeros(2, np.float32)
        case = np.zeros(1, np.int32)
        oracle.Restated.lib().orc_block_solve(K.ctypes.data, b.ctypes.data, a.ctypes.data, case.ctypes.data)
        Km = np.array([[k11, k12], [k12, k22]], np.float64)
        bb = b.astype(np.float64)
        # brute force: first valid of the four enumerated cases
        cands = [np.linalg.solve(Km, -bb), np.array([-bb[0] / k11, 0.0]), np.array([0.0, -bb[1] / k22]),
                 np.zeros(2)]
        expect = 0
        for ci, x in enumerate(cands, start=1):
            wv = Km @ x + bb
            free = [ci in (1, 2), ci in (1, 3)]
            if (x >= -1e-6).all() and all(wv[k] >= -1e-5 or free[k] for k in range(2)):
                expect = ci
                break
        if case[0] == 0 or expect == 0:
            continue
        taken += 1
        assert case[0] == expect
        x = a.astype(np.float64)
        wv = Km @ x + bb
        assert (x >= 0).all() and (wv >= -1e-4 * (1 + abs(bb).max())).all()
        assert abs(x @ wv) < 1e-4 * (1 + abs(bb).max() * (1 + abs(x).max()))
    assert taken > 1000


def _float32_free_fall(v0, c0, g, steps, h=np.float32(1 / 60)):
    v = np.array(v0, np.float32)
    c = np.array(c0, np.float32)
    m = np.float32(math.pi * 0.5 * 0.5)  # circle r=0.5, density 1 (b2CircleShape::ComputeMass)
    m = np.float32(np.float32(np.float32(1.0) * np.float32(3.14159265359)) * np.float32(0.5)) * np.float32(0.5)
    inv = np.float32(1.0) / m
    for _ in range(steps):
        s1 = np.float32(h * inv)
        s2 = np.float32(np.float32(1.0) * m)
        v = (v + s1 * (s2 * np.asarray(g, np.float32))).astype(np.float32)
        c = (c + np.float32(h) * v).astype(np.float32)
    return v, c


def test_orphan_integration_matches_closed_form():
    sc = S.Scene(name="freefall")
    b = sc.add_body(S.DYNAMIC, 1.0, 50.0, 0.0, 0.5, 0.0)
    sc.add_circle(b, 0.5)
    w = oracle.RestatedWorld(sc)
    w.step(30)
    v, c = _float32_free_fall((0.5, 0.0), (1.0, 50.0), (0.0, -10.0), 30)
    got = w.bodies()[0]
    assert np.array_equal(bits(got["v"]), bits(v))
    assert np.array_equal(bits(got["c"]), bits(c))
    # and the textbook values to float32 accuracy
    t = 30 / 60
    assert abs(got["v"][1] - (-10.0 * t)) < 1e-4
    assert abs(got["c"][1] - (50.0 - 10.0 * (30 * 31 / 2) / 3600.0)) < 1e-3


def hello_world_scene():
    s = S.Scene(name="hello")
    g = s.add_body(S.STATIC, 0.0, -10.0)
    s.add_box(g, 50.0, 10.0, density=0.0, friction=0.2)
    b = s.add_body(S.DYNAMIC, 0.0, 4.0)
    s.add_box(b, 1.0, 1.0, density=1.0, friction=0.3)
    return s


def test_hello_world_unit_test_on_restatement():
    """unit-test/hello_world.cpp:109-111 of the reference: after 60 steps at
    6/2 iterations the box rests at y = 1.01 +- 0.01."""
    w = oracle.RestatedWorld(hello_world_scene())
    w.step(60, vi=6, pi=2)
    b = w.bodies()[1]
    x, y = b["xf"][0], b["xf"][1]
    assert abs(x) < 0.01 and abs(y - 1.01) < 0.01 and abs(b["a"]) < 0.01


def joint_reaction_scene():
    s = S.Scene(name="reaction")
    g = s.add_body(S.STATIC, 0.0, 0.0)
    b = s.add_body(S.DYNAMIC, -2.0, 3.0)
    s.add_circle(b, 1.0, density=1.0, friction=0.2, filt=(1, 0, 0, 0))
    s.add_revolute(g, b, (-2.0, 3.0))
    return s


def test_revolute_reaction_unit_test_on_restatement():
    """unit-test/joint_test.cpp:82-105: after one step the revolute reaction
    force is -m g within 1e-5 and the reaction torque is 0."""
    w = oracle.RestatedWorld(joint_reaction_scene())
    w.step(1, vi=6, pi=2)
    j = w.joints()[0]
    m = np.float32(np.float32(1.0 * np.float32(3.14159265359)) * np.float32(1.0)) * np.float32(1.0)
    F = j["impulse"] * np.float32(60.0)
    assert F[0] == 0.0
    assert abs(F[1] + m * -10.0) < 1e-5 * 60 * 10
    assert j["motor_impulse"] + j["lower_impulse"] - j["upper_impulse"] == 0.0


def test_restatement_matches_reference_live_when_built():
    if not oracle.Ref.available():
        pytest.skip("reference library not built here")
    sc = S.small_mixed(seed=13, n=40)
    r = oracle.RefWorld(sc)
    w = oracle.RestatedWorld(sc)
    r.step(5)
    w.step(5)
    rb = r.bodies()
    ob = w.bodies()
    # the first steps have no contact-order ambiguity between the two
    assert np.abs(rb[:, 0:2] - ob["c"]).max() < 1e-4


# ---- closed forms for the collide functions --------------------------------
# Each case has a manifold fixed by geometry alone; the restatement's world
# manifold (b2WorldManifold::Initialize, b2_collision.cpp:26-90) must match
# it to float32 accuracy. polygonRadius (b2_common.h) is the skin of polygons
# and edges.
POLY_R = 0.01


def _shapes(build):
    s = S.Scene(name="closed")
    g = s.add_body(S.STATIC, 0.0, 0.0)
    b = s.add_body(S.DYNAMIC, 0.0, 0.0)
    build(s, g, b)
    return oracle.RestatedWorld(s).shapes()


def _xf(x, y, a=0.0):
    return np.array([x, y, math.sin(a), math.cos(a)], np.float32)


def _world(sa, xa, sb, xb):
    c = oracle.orc_collide(sa, xa, sb, xb)
    n = int(c["point_count"])
    wm = oracle.orc_world_manifold(c, xa, sa["radius"], xb, sb["radius"])
    return n, wm[0:2], wm[2:6].reshape(2, 2)[:n], wm[6:8][:n]


def test_circle_circle_closed_form():
    sh = _shapes(lambda s, g, b: (s.add_circle(g, 0.5), s.add_circle(b, 0.3)))
    ca, cb = np.array([0.1, -0.2]), np.array([0.75, 0.15])
    n, nrm, pts, sep = _world(sh[0], _xf(*ca), sh[1], _xf(*cb))
    d = cb - ca
    L = np.linalg.norm(d)
    u = d / L
    assert n == 1
    assert np.allclose(nrm, u, atol=1e-6)
    assert abs(sep[0] - (L - 0.8)) < 2e-6
    mid = 0.5 * ((ca + 0.5 * u) + (cb - 0.3 * u))
    assert np.allclose(pts[0], mid, atol=2e-6)
    # separated by more than the radii: no point
    n, *_ = _world(sh[0], _xf(0, 0), sh[1], _xf(0.81, 0))
    assert n == 0


def test_polygon_circle_face_and_vertex_regions():
    sh = _shapes(lambda s, g, b: (s.add_box(g, 0.5, 0.5), s.add_circle(b, 0.25), s.add_circle(b, 0.3)))
    box, c25, c30 = sh[0], sh[1], sh[2]
    # face region of the top face
    n, nrm, pts, sep = _world(box, _xf(0, 0), c25, _xf(0.2, 0.7))
    assert n == 1 and np.allclose(nrm, [0, 1], atol=1e-7)
    assert abs(sep[0] - (0.7 - 0.5 - 0.25 - POLY_R)) < 2e-6
    # world point: midway between the polygon's skin and the circle's surface
    assert np.allclose(pts[0], [0.2, 0.5 * ((0.5 + POLY_R) + (0.7 - 0.25))], atol=2e-6)
    # vertex region of the corner (0.5, 0.5)
    cc = np.array([0.7, 0.72])
    n, nrm, pts, sep = _world(box, _xf(0, 0), c30, _xf(*cc))
    d = cc - np.array([0.5, 0.5])
    L = np.linalg.norm(d)
    assert n == 1 and np.allclose(nrm, d / L, atol=1e-6)
    assert abs(sep[0] - (L - 0.3 - POLY_R)) < 2e-6
    # rotating the box by 90 degrees about its centre changes nothing
    n2, nrm2, _, sep2 = _world(box, _xf(0, 0, math.pi / 2), c30, _xf(*cc))
    assert n2 == 1 and np.allclose(nrm2, nrm, atol=1e-6) and abs(sep2[0] - sep[0]) < 2e-6


def test_box_on_box_axis_aligned_closed_form():
    sh = _shapes(lambda s, g, b: (s.add_box(g, 0.5, 0.5), s.add_box(b, 0.5, 0.5)))
    pen = 0.02
    xb = 0.3
    n, nrm, pts, sep = _world(sh[0], _xf(0, 0), sh[1], _xf(xb, 1.0 - pen))
    assert n == 2
    assert np.allclose(nrm, [0, 1], atol=1e-7)
    assert np.allclose(sep, -pen - 2 * POLY_R, atol=2e-6)
    # reference face = A's top face (ties go to A), incident = B's bottom
    # edge; its corner x = xb - 0.5 lies inside A's side planes, the other end
    # is clipped at the side plane, pushed out by the total skin radius
    # (b2_collide_polygon.cpp side offsets + totalRadius)
    assert np.allclose(sorted(pts[:, 0]), [xb - 0.5, 0.5 + 2 * POLY_R], atol=2e-6)
    assert np.allclose(pts[:, 1], 0.5 - pen / 2, atol=2e-6)


def test_box_on_edge_ground_closed_form():
    sh = _shapes(lambda s, g, b: (s.add_edge(g, (-5.0, 0.0), (5.0, 0.0)), s.add_box(b, 0.5, 0.5)))
    edge, box = sh[0], sh[1]
    gap = -0.01                                  # box bottom 1 cm below the edge line
    n, nrm, pts, sep = _world(edge, _xf(0, 0), box, _xf(0.3, 0.5 + gap))
    assert n == 2
    assert np.allclose(nrm, [0, 1], atol=1e-7)
    assert np.allclose(sep, gap - 2 * POLY_R, atol=2e-6)
    assert np.allclose(sorted(pts[:, 0]), [-0.2, 0.8], atol=2e-6)
    assert np.allclose(pts[:, 1], 0.5 * ((0.0 + POLY_R) + (gap - POLY_R)), atol=2e-6)


def test_edge_circle_regions_closed_form():
    sh = _shapes(lambda s, g, b: (s.add_edge(g, (-1.0, 0.0), (1.0, 0.0)), s.add_circle(b, 0.5)))
    edge, circ = sh[0], sh[1]
    # region AB: the segment interior
    n, nrm, pts, sep = _world(edge, _xf(0, 0), circ, _xf(0.3, 0.4))
    assert n == 1 and np.allclose(nrm, [0, 1], atol=1e-7)
    assert abs(sep[0] - (0.4 - 0.5 - POLY_R)) < 2e-6
    # region A: beyond vertex (-1, 0)
    cc = np.array([-1.3, 0.3])
    n, nrm, pts, sep = _world(edge, _xf(0, 0), circ, _xf(*cc))
    d = cc - np.array([-1.0, 0.0])
    L = np.linalg.norm(d)
    assert n == 1 and np.allclose(nrm, d / L, atol=1e-6)
    assert abs(sep[0] - (L - 0.5 - POLY_R)) < 2e-6


# ---- user queries of the reference world, pinned by brute force / closed forms
def test_reference_query_aabb_matches_brute_force():
    """b2World::QueryAABB of the compiled reference reports exactly the
    fixtures whose AABB overlaps the box (inclusive), on a mixed scene."""
    if not oracle.Ref.available():
        pytest.skip("reference library not built here")
    sc = S.random_pile(500, seed=61, half_w=15.0, height=30.0)
    w = oracle.RefWorld(sc)
    aabbs = oracle.ref_aabbs(w)
    rs = S.Stream(62)
    for k in range(40):
        u = rs.uniform(4)
        lo = np.array([-15 + 30 * u[0], 30 * u[1]], np.float32)
        box = np.array([lo[0], lo[1], lo[0] + 0.5 + 6 * u[2], lo[1] + 0.5 + 6 * u[3]], np.float32)
        got = set(w.query_aabb(box).tolist())
        want = {i for i, a in enumerate(aabbs)
                if a[2] >= box[0] and a[0] <= box[2] and a[3] >= box[1] and a[1] <= box[3]}
        assert got == want, k


def test_reference_raycast_closed_forms():
    """b2World::RayCast on shapes with known hits: a circle of radius r at
    distance D along the ray (fraction (D - r) / L, normal -ray), and an
    axis-aligned box face (fraction, outward normal); a ray that misses
    reports nothing."""
    if not oracle.Ref.available():
        pytest.skip("reference library not built here")
    s = S.Scene(name="rays", gravity=(0.0, 0.0))
    b = s.add_body(S.DYNAMIC, 10.0, 0.0)
    s.add_circle(b, 1.0)
    b2 = s.add_body(S.DYNAMIC, 0.0, 10.0)
    s.add_box(b2, 2.0, 1.0)
    w = oracle.RefWorld(s)
    h = w.raycast_all(np.array([0.0, 0.0, 20.0, 0.0], np.float32))
    assert len(h) == 1 and int(h[0, 0]) == 0
    assert abs(h[0, 1] - 9.0 / 20.0) < 1e-6 and np.allclose(h[0, 2:4], [-1.0, 0.0], atol=1e-6)
    assert np.allclose(h[0, 4:6], [9.0, 0.0], atol=1e-5)
    h = w.raycast_all(np.array([0.5, 0.0, 0.5, 20.0], np.float32))
    assert len(h) == 1 and int(h[0, 0]) == 1
    assert abs(h[0, 1] - 9.0 / 20.0) < 1e-6 and np.allclose(h[0, 2:4], [0.0, -1.0], atol=1e-6)
    assert len(w.raycast_all(np.array([0.0, 5.0, 20.0, 5.0], np.float32))) == 0
