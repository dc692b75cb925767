"""The drop-in b2World on the GPU: what user code does between steps.

Each test drives the C++ drop-in objects (through the scene loader's C ABI)
and checks the device against the CPU restatement started from the device's
own state, solving in the device's colour order (1e-5 m / 1e-5 rad, as the
one-step parity tests), or against exact invariants:

  * Step(0) runs Collide alone (b2_world.cpp:1124-1141): bodies are kept,
    stored impulses follow their contact ids, and the next step equals a run
    without the Step(0);
  * DestroyBody / DestroyJoint / DestroyFixture (b2_world.cpp:178-418,
    b2_body.cpp:280-350) remove exactly the destroyed records and their
    contacts; the surviving contacts keep manifolds, impulses and colours;
  * SetFilterData (b2_fixture.cpp:142-146) drops the fixture's contacts that
    the new filter rejects at the next step's Collide (b2_contact_manager.cpp:76-92);
  * the device keeps PostSolve's point count per contact (b2_island.cpp:621-647).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from util import REPO, bits, load_pkg, scene_for

pytestmark = pytest.mark.gpu
S = load_pkg().scenes
DT = np.float32(1.0 / 60.0)


def scene_without(sc, bodies=(), fixtures=(), joints=()):
    """The scene as the drop-in holds it after destroying `bodies` (with their
    fixtures and joints), `fixtures` and `joints` (indices of the original
    scene); survivors keep their creation order."""
    dead_b = set(bodies)
    dead_f = set(fixtures) | {i for i, b in enumerate(sc.fix_body) if b in dead_b}
    dead_j = set(joints) | {q for q, (a, b) in enumerate(sc.joint_bodies) if a in dead_b or b in dead_b}
    out = S.Scene(gravity=sc.gravity, name=sc.name + "-edited")
    bmap = {}
    for i in range(sc.body_count):
        if i in dead_b:
            continue
        bmap[i] = out.add_body(sc.body_type[i], *sc.body_pos[i], sc.body_angle[i], *sc.body_vel[i])
    for i in range(sc.fixture_count):
        if i in dead_f:
            continue
        out.fix_body.append(bmap[sc.fix_body[i]])
        out.fix_kind.append(sc.fix_kind[i])
        out.fix_geom.append(sc.fix_geom[i])
        out.fix_material.append(sc.fix_material[i])
        out.fix_filter.append(sc.fix_filter[i])
    for q in range(sc.joint_count):
        if q in dead_j:
            continue
        a, b = sc.joint_bodies[q]
        out.joint_bodies.append((bmap[a], bmap[b]))
        out.joint_params.append(sc.joint_params[q])
        out.joint_flags.append(sc.joint_flags[q])
    return out, sorted(dead_f), sorted(dead_j)


def _joint_colours(sc):
    """Greedy joint colouring in joint order over dynamic bodies (the order
    the engine builds at upload)."""
    used = {}
    col = np.zeros(sc.joint_count, np.int32)
    bt = np.asarray(sc.body_type)
    for q, (a, b) in enumerate(sc.joint_bodies):
        m = (used.get(a, 0) if bt[a] == 2 else 0) | (used.get(b, 0) if bt[b] == 2 else 0)
        c = 0
        while (m >> c) & 1:
            c += 1
        col[q] = c
        for x in (a, b):
            if bt[x] == 2:
                used[x] = used.get(x, 0) | (1 << c)
    return col


def _restated_step(sc, bodies0, contacts0, joints0, gpu_contacts_after, gpu_colours, flags):
    """One restatement step from the device's state, in the device's colour order."""
    r = oracle.RestatedWorld(sc)
    rb = r.bodies()
    for f in ("c", "a", "v", "w", "xf"):
        rb[f] = bodies0[f]
    r.set_bodies(rb)
    r.set_contacts(contacts0)
    if sc.joint_count:
        j = r.joints()
        for f in ("impulse", "motor_impulse", "lower_impulse", "upper_impulse"):
            j[f] = joints0[f]
        r.set_joints(j)
    r.set_inv_dt0(np.float32(1.0) / DT)
    r.set_colours(oracle.contact_keys(gpu_contacts_after), gpu_colours,
                  _joint_colours(sc) if sc.joint_count else None)
    r.step(1, flags=flags)
    return r


def _compare(sc, gb, ob, label):
    dyn = np.asarray(sc.body_type) == 2
    dc = np.abs(gb["c"] - ob["c"])[dyn].max()
    da = np.abs(gb["a"] - ob["a"])[dyn].max()
    dv = np.abs(gb["v"] - ob["v"])[dyn].max()
    print(f"{label}: max |dc| {dc:.3g} |da| {da:.3g} |dv| {dv:.3g}", flush=True)
    assert dc < 1e-5 and da < 1e-5 and dv < 1e-3


def test_gpu_unsupported_features_are_refused_and_the_rest_steps(engine):
    """The refusal example on a device: every unsupported feature is refused
    with the world untouched (status 5), and the served cases (sleeping on,
    a body put to sleep, sleeping off, a PostSolve listener) step (status 0)."""
    exe = os.path.join(REPO, "box2d-optimized_amd", "build", "unsupported_features")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    status = [int(ln.split(":")[1].split()[0]) for ln in r.stdout.strip().splitlines()]
    assert status == [5, 0, 0, 0, 5, 5, 5, 5, 0]


def test_gpu_step_zero_is_collide_only(engine):
    """Step(dt = 0): bodies are kept and the contacts' stored impulses are
    carried by contact id (the solve, which would scale warm starts by
    dtRatio = 0, does not run), the
    next Step(1/60) is bit-identical to a run without the Step(0), and the
    restatement's Step(0) gives the same contact list bit for bit."""
    sc = S.pyramid(12)
    a = engine.World(sc)
    b = engine.World(sc)
    a.step(40)
    b.step(40)
    ba, ca = a.batch().bodies(), a.contacts()
    assert np.abs(ca["normal_impulse"]).max() > 0.0
    a.step(1, dt=0.0)
    ba0, ca0 = a.batch().bodies(), a.contacts()
    for f in ("c", "a", "v", "w", "xf"):
        assert np.array_equal(bits(ba[f]), bits(ba0[f])), f
    assert np.array_equal(oracle.contact_keys(ca), oracle.contact_keys(ca0))
    # Collide re-evaluates every manifold at the bodies' current (post-solve)
    # positions, so a manifold's points may come back in another order; each
    # new point takes the stored impulses of the old point with its id, and
    # 0 without one (b2_contact.cpp:158-181)
    carried = 0
    for old, new in zip(ca, ca0):
        for k in range(new["point_count"]):
            j = [j for j in range(old["point_count"]) if old["id"][j] == new["id"][k]]
            en = old["normal_impulse"][j[0]] if j else np.float32(0)
            et = old["tangent_impulse"][j[0]] if j else np.float32(0)
            assert np.float32(en).view(np.uint32) == new["normal_impulse"][k].view(np.uint32)
            assert np.float32(et).view(np.uint32) == new["tangent_impulse"][k].view(np.uint32)
            carried += bool(j)
    assert carried > 0
    # the restatement's Collide from the same state
    r = oracle.RestatedWorld(sc)
    rb = r.bodies()
    for f in ("c", "a", "v", "w", "xf"):
        rb[f] = ba[f]
    r.set_bodies(rb)
    r.set_contacts(ca)
    r.set_inv_dt0(np.float32(1.0) / DT)
    r.step(1, dt=0.0)
    rc = r.contacts()
    assert np.array_equal(oracle.contact_keys(rc), oracle.contact_keys(ca0))
    for f in ("point_count", "id"):
        assert np.array_equal(rc[f], ca0[f]), f
    for f in ("normal_impulse", "tangent_impulse", "point_local", "local_normal"):
        assert np.array_equal(bits(rc[f]), bits(ca0[f])), f
    a.step(1)
    b.step(1)
    ga, gb = a.batch().bodies(), b.batch().bodies()
    for f in ("c", "a", "v", "w"):
        assert np.array_equal(bits(ga[f]), bits(gb[f])), f


def test_gpu_destroy_bodies_joint_fixture_mid_run_vs_restatement(engine):
    """Destroy two bodies (one jointed, so its joints go too), one more joint
    and a fixture mid-run. The device keeps exactly the surviving records and
    contacts (manifolds, impulses and colours bit for bit); the next step
    matches the restatement of the edited scene from the same state in the
    device's colour order."""
    sc, _ = scene_for("chain")
    w = engine.World(sc)
    w.step(45)
    b = w.batch()
    before = b.contacts()
    col_before = b.contact_colours(len(before))
    # In this scene body k holds fixture k: 0 the ground, 1..12 the chain's
    # links (link k hangs on joints k-1 and k), 13..32 the circles. Destroy a
    # link and a circle that are in contact with something, so that contacts
    # go with them, plus another joint and another circle's fixture.
    in_contact = set(before["shape_a"].tolist()) | set(before["shape_b"].tolist())
    link = next(k for k in range(1, 12) if k in in_contact)
    circles = [k for k in range(13, 33) if k in in_contact]
    assert len(circles) >= 2, in_contact
    kill_bodies = [link, circles[0]]
    kill_joint_orig = 9 if link not in (9, 10) else 3
    kill_fixture_orig = circles[-1]  # its body stays, without fixtures
    print(f"destroying bodies {kill_bodies}, joint {kill_joint_orig}, fixture {kill_fixture_orig}", flush=True)
    edited, dead_f, dead_j = scene_without(sc, kill_bodies, [kill_fixture_orig], [kill_joint_orig])
    # destroy through the drop-in (indices in the current creation order)
    w.destroy_body(circles[0])
    w.destroy_body(link)
    joints_left = [q for q in range(sc.joint_count) if sc.joint_bodies[q][0] not in kill_bodies
                   and sc.joint_bodies[q][1] not in kill_bodies]
    w.destroy_joint(joints_left.index(kill_joint_orig))
    fixtures_left = [i for i in range(sc.fixture_count) if sc.fix_body[i] not in kill_bodies]
    w.destroy_fixture(fixtures_left.index(kill_fixture_orig))
    w.prepare()
    b = w.batch()
    # the drop-in's Step sets dtRatio = inv_dt0 * dt itself; stepping the batch
    # directly below, the same ratio is set here
    ratio = np.float32(np.float32(1.0) / DT) * DT
    b.set_world_params(0, [engine.WorldParams((sc.gravity[0], sc.gravity[1]), DT, ratio, 1, 1)])
    assert b.counts().bodies == edited.body_count and b.counts().shapes == edited.fixture_count
    assert b.counts().joints == edited.joint_count
    # survivors of the contact list, renumbered
    fmap = {i: k for k, i in enumerate(i for i in range(sc.fixture_count) if i not in dead_f)}
    keep = np.array([c["shape_a"] in fmap and c["shape_b"] in fmap for c in before], bool)
    expect = before[keep].copy()
    expect["shape_a"] = [fmap[x] for x in expect["shape_a"]]
    expect["shape_b"] = [fmap[x] for x in expect["shape_b"]]
    after = b.contacts()
    print(f"contacts {len(before)} -> {len(after)} (destroyed {int((~keep).sum())})", flush=True)
    assert len(after) == int(keep.sum()) and (~keep).sum() > 0
    assert np.array_equal(after.view(np.uint8), expect.view(np.uint8))
    assert np.array_equal(b.contact_colours(len(after)), col_before[keep])
    bodies0, joints0 = b.bodies(), b.joints(edited.joint_count)
    # one device step, the restatement of the edited scene from the same state
    b.step(8, 3, engine.STEP_SKIP_BROADPHASE)
    gb, gc = b.bodies(), b.contacts()
    r = _restated_step(edited, bodies0, after, joints0, gc, b.contact_colours(len(gc)), engine.STEP_SKIP_BROADPHASE)
    _compare(edited, gb, r.bodies(), "after destroy")
    oc = r.contacts()
    assert np.array_equal(oracle.contact_keys(gc), oracle.contact_keys(oc))
    assert np.array_equal(gc["point_count"], oc["point_count"])
    # and the drop-in keeps stepping the edited world
    w.step(30)
    assert np.isfinite(w.bodies()).all()


def test_gpu_set_filter_data_drops_rejected_contacts_at_the_next_collide(engine):
    """SetFilterData with a mask that rejects everything: the fixture's
    contacts are gone before the next step's solve, the step equals the
    restatement without them, and the broadphase does not re-create them."""
    sc = S.pyramid(8)
    w = engine.World(sc)
    w.step(30)
    b = w.batch()
    before = b.contacts()
    target = 1 + 8 * 7 // 2  # a box in the middle of the pyramid
    mine = (before["shape_a"] == target) | (before["shape_b"] == target)
    assert (mine & (before["point_count"] > 0)).sum() >= 2
    bodies0 = b.bodies()
    w.set_filter(target, 1, 0, 0)
    w.step(1)
    gb, gc = b.bodies(), b.contacts()
    assert not ((gc["shape_a"] == target) | (gc["shape_b"] == target)).any()
    edited = S.pyramid(8)
    edited.fix_filter[target] = (1, 0, 0, 0)
    kept = before[~mine]
    # the device solved exactly the other contacts; the restatement solves
    # them in the device's colour order (colours of the contacts it still holds)
    r = oracle.RestatedWorld(edited)
    rb = r.bodies()
    for f in ("c", "a", "v", "w", "xf"):
        rb[f] = bodies0[f]
    r.set_bodies(rb)
    r.set_contacts(kept)
    r.set_inv_dt0(np.float32(1.0) / DT)
    r.set_colours(oracle.contact_keys(gc), b.contact_colours(len(gc)))
    r.step(1)
    _compare(edited, gb, r.bodies(), "after SetFilterData")
    assert np.array_equal(oracle.contact_keys(gc), oracle.contact_keys(r.contacts()))


def test_gpu_solver_points_name_the_stored_impulses(engine):
    """Per contact, the points the last step's solver stored (PostSolve's
    count): 0 for contacts it did not solve, else 1 or 2 and never more than
    the manifold's points; one-point manifolds store one."""
    sc, _ = scene_for("pile")
    w = engine.World(sc)
    w.step(60)
    b = w.batch()
    c = b.contacts()
    pts = b.solver_points(len(c))
    touching = c["point_count"] > 0
    solved_new = (pts > 0) & ~touching
    print(f"contacts {len(c)}, touching {touching.sum()}, solved {np.sum(pts > 0)}, "
          f"two-point drop {np.sum((pts == 1) & (c['point_count'] == 2))}", flush=True)
    # a contact made by this step's broadphase was not solved in it
    assert not solved_new.any()
    assert np.all(pts <= c["point_count"])
    assert np.all(pts[touching] >= 1)
    assert np.all(pts[c["point_count"] == 1] <= 1)


def test_gpu_body_of_an_undeclared_world_is_rejected(engine):
    """A body tagged with a world the batch does not declare is an error at
    the next step, not an out-of-bounds write."""
    bt = engine.Batch(worlds=2, bodies=8, shapes=8, joints=1, contacts=64)
    bt.set_worlds([engine.WorldParams((0.0, -10.0), DT, 0.0, 1, 1)])
    rec = np.zeros(2, dtype=engine.BODY_DTYPE)
    rec["type"] = 2
    rec["xf"][:, 3] = 1.0
    rec["world"] = [0, 1]  # world 1 is within the capacity but not declared
    bt.upload_bodies(0, rec)
    with pytest.raises(engine.B2GError, match="world"):
        bt.step(8, 3, engine.STEP_FIND_NEW_CONTACTS)
