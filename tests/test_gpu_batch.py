"""GPU tests of the batching pattern: many independent worlds stepped by one
set of launches.

A world's result must not depend on the other worlds in its batch: every
stage works per contact/body/island, and the colouring rank is a function of
the world-local shape pair only. So the comparison with the same world
stepped alone is bit-exact.
"""
import numpy as np
import pytest

from util import bits, load_pkg, scene_for

pytestmark = pytest.mark.gpu
S = load_pkg().scenes


def _mixed_scenes():
    return [
        S.pyramid(6),
        S.random_pile(150, seed=11, half_w=8.0, height=20.0),
        scene_for("chain")[0],
        S.small_mixed(seed=5, n=40),
        S.batch_world(3, 120),
        S.pyramid(12),
        S.random_pile(400, seed=12, half_w=12.0, height=30.0),
        S.batch_world(17, 60),
    ]


def _state(rec):
    return np.concatenate([bits(rec["xf"]).reshape(len(rec), -1), bits(rec["c"]).reshape(len(rec), -1),
                           bits(rec["a"]).reshape(len(rec), -1), bits(rec["v"]).reshape(len(rec), -1),
                           bits(rec["w"]).reshape(len(rec), -1)], axis=1)


def test_gpu_batch_worlds_match_worlds_alone_bitexact(engine):
    scenes = _mixed_scenes()
    steps = 40
    print("batch of", len(scenes), "worlds,", sum(s.body_count for s in scenes), "bodies", flush=True)
    batch = engine.Batch.from_scenes(scenes).advance(steps)
    got = batch.bodies()
    for i, sc in enumerate(scenes):
        alone = engine.Batch.from_scenes([sc]).advance(steps)
        ref = alone.bodies()
        part = got[batch.world_bodies[i]:batch.world_bodies[i + 1]]
        assert (part["world"] == i).all()
        same = _state(part) == _state(ref)
        print(f"world {i} ({sc.name}): {same.all(axis=1).sum()} of {len(ref)} bodies bit-identical", flush=True)
        assert same.all(), sc.name


@pytest.mark.parametrize("solver", ["colours", "worlds"])
def test_gpu_mixed_warm_starting_matches_worlds_alone_bitexact(engine, solver):
    """Worlds with and without warm starting in one batch: the warm-start
    pass (a per-body gather on the colour-major path, in-kernel on the
    per-world path) reads each world's flag; each world still equals itself
    stepped alone with the same flag, and the flag changes the result."""
    scenes = _mixed_scenes()
    if solver == "worlds":  # the per-world solver takes joint-free batches
        scenes = [sc for sc in scenes if sc.joint_count == 0]
    warm = [k % 2 == 0 for k in range(len(scenes))]
    steps = 30
    opt = engine.SOLVER_WORLDS if solver == "worlds" else engine.SOLVER_COLOURS
    print("batch of", len(scenes), "worlds, solver", solver, "warm", warm, flush=True)
    batch = engine.Batch.from_scenes(scenes)
    batch.set_option(engine.OPT_SOLVER, opt)
    got = batch.advance(steps, warm=warm).bodies()
    changed = 0
    for i, sc in enumerate(scenes):
        alone = engine.Batch.from_scenes([sc]).advance(steps, warm=[warm[i]]).bodies()
        part = got[batch.world_bodies[i]:batch.world_bodies[i + 1]]
        same = _state(part) == _state(alone)
        print(f"world {i} ({sc.name}, warm {warm[i]}): {same.all(axis=1).sum()} of {len(alone)} bit-identical",
              flush=True)
        assert same.all(), sc.name
        if not warm[i]:
            other = engine.Batch.from_scenes([sc]).advance(steps, warm=[True]).bodies()
            changed += int(not (_state(other) == _state(alone)).all())
    assert changed > 0  # switching warm starting off is visible


def test_gpu_per_world_solver_with_large_worlds_bitexact(engine):
    """Worlds above 1024 bodies get a per-world solver launch of their own,
    beside the launch of the smaller worlds (on a second stream); every world
    still equals itself stepped alone."""
    scenes = [S.random_pile(1300, seed=31, half_w=22.0, height=60.0), S.pyramid(8),
              S.random_pile(150, seed=32, half_w=8.0, height=20.0), S.random_pile(1100, seed=33, half_w=20.0, height=55.0),
              S.batch_world(5, 90), S.batch_world(6, 300)]
    assert sum(sc.body_count > 1024 for sc in scenes) == 2
    steps = 25
    batch = engine.Batch.from_scenes(scenes)
    batch.set_option(engine.OPT_SOLVER, engine.SOLVER_WORLDS)
    batch.set_option(engine.OPT_WORLD_BLOCK, 128)  # the small worlds' launch: 128-thread workgroups
    got = batch.advance(steps).bodies()
    assert batch.get_option(engine.OPT_LAST_SOLVER) == 2
    for i, sc in enumerate(scenes):
        alone = engine.Batch.from_scenes([sc]).advance(steps).bodies()
        part = got[batch.world_bodies[i]:batch.world_bodies[i + 1]]
        same = _state(part) == _state(alone)
        print(f"world {i} ({sc.name}, {sc.body_count} bodies): {same.all(axis=1).sum()} of {len(alone)} bit-identical",
              flush=True)
        assert same.all(), sc.name


@pytest.mark.parametrize("solver", ["colours", "worlds"])
def test_gpu_idle_worlds_keep_pairs_bitexact(engine, solver):
    """A world that sits out a step keeps its pairs (its segments of the
    contact list), its boxes and its part of the tree (leaf order, nodes,
    range table) instead of being queried and rebuilt again
    (B2G_OPT_KEEP_IDLE). On a schedule where worlds go idle for several steps
    (across tree rebuilds) and come back, the contact list after every step,
    the final bodies and AABB queries on the final tree equal, bit for bit,
    those of the same schedule with every world queried and rebuilt each
    step."""
    scenes = [sc for sc in _mixed_scenes() if solver == "colours" or sc.joint_count == 0]
    steps = 24
    rng = np.random.default_rng(7)
    sched = rng.random((steps, len(scenes))) < 0.45
    sched[:3, 0] = True
    sched[3:12, 0] = False  # world 0: idle for 9 steps, then back
    sched[13, :] = False    # a step with every world idle (nothing to sort or query)
    sched[14, :] = True     # and one with every world active
    dt = np.float32(1.0 / 60.0)
    ratio = np.float32(np.float32(np.float32(1.0) / dt) * dt)
    opt = engine.SOLVER_WORLDS if solver == "worlds" else engine.SOLVER_COLOURS
    runs = []
    for keep in (1, 0):
        b = engine.Batch.from_scenes(scenes)
        b.set_option(engine.OPT_SOLVER, opt)
        b.set_option(engine.OPT_KEEP_IDLE, keep)
        assert b.get_option(engine.OPT_KEEP_IDLE) == keep
        b.advance(2)
        lists = []
        for t in range(steps):
            b.set_world_params(0, [engine.WorldParams(g, dt, ratio, 1, int(sched[t, w]))
                                   for w, g in enumerate(b.gravity)])
            b.step(8, 3, 0)
            lists.append(b.contacts())
        # the kept tree answers queries like a rebuilt one: boxes over each world
        qrng = np.random.default_rng(8)
        qw = np.repeat(np.arange(len(scenes), dtype=np.int32), 16)
        lo = qrng.uniform(-15.0, 15.0, (len(qw), 2)).astype(np.float32)
        boxes = np.concatenate([lo, lo + qrng.uniform(0.5, 8.0, (len(qw), 2)).astype(np.float32)], axis=1)
        found = [np.sort(r) for r in b.query_aabbs(qw, boxes)]
        runs.append((lists, b.bodies(), found))
    (kl, kb, kq), (fl, fb, fq) = runs
    for t in range(steps):
        assert len(kl[t]) == len(fl[t]) and len(kl[t]) > 0, t
        assert kl[t].tobytes() == fl[t].tobytes(), f"contact list differs after step {t}"
    assert (_state(kb) == _state(fb)).all()
    assert sum(len(r) for r in kq) > 0
    for a, c in zip(kq, fq):
        assert np.array_equal(a, c)
    print(f"{len(scenes)} worlds, {sched.sum()} world-steps of {sched.size} active, "
          f"{len(kl[-1])} contacts at the end", flush=True)


def test_gpu_upload_to_idle_world_requeries_its_pairs(engine):
    """An upload makes the next broadphase query and rebuild every world: a
    world teleported while idle gets the pairs of its new positions at once,
    exactly as with idle keeping off."""
    scenes = _mixed_scenes()
    dt = np.float32(1.0 / 60.0)
    ratio = np.float32(np.float32(np.float32(1.0) / dt) * dt)

    def step(b, active):
        b.set_world_params(0, [engine.WorldParams(g, dt, ratio, 1, int(active[w])) for w, g in enumerate(b.gravity)])
        b.step(8, 3, 0)

    runs = []
    for keep in (1, 0):
        b = engine.Batch.from_scenes(scenes)
        b.set_option(engine.OPT_KEEP_IDLE, keep)
        b.advance(3)
        idle0 = [w != 1 for w in range(len(scenes))]
        step(b, idle0)
        step(b, idle0)
        # world 1 (a pile), idle: every dynamic body shifted by 0.4 m
        lo, hi = b.world_bodies[1], b.world_bodies[2]
        rec = b.bodies(lo, hi - lo)
        dyn = rec["type"] == 2
        rec["xf"][dyn, 0] += np.float32(0.4)
        rec["c"][dyn, 0] += np.float32(0.4)
        b.upload_bodies(lo, rec)
        lists = []
        step(b, idle0)                      # world 1 still idle: the upload alone must refresh its pairs
        lists.append(b.contacts())
        for _ in range(3):
            step(b, [True] * len(scenes))  # then everyone moves
            lists.append(b.contacts())
        runs.append((lists, b.bodies()))
    (kl, kb), (fl, fb) = runs
    for t in range(len(kl)):
        assert len(kl[t]) > 0 and kl[t].tobytes() == fl[t].tobytes(), t
    assert (_state(kb) == _state(fb)).all()


def test_gpu_many_live_batches(engine):
    """Each batch runs its side stream on a hardware queue of its own when
    the runtime can give one, and shares one otherwise: 24 batches alive at
    once all step, each exactly like a lone batch of the same world."""
    sc = S.random_pile(120, seed=41, half_w=8.0, height=20.0)
    ref = engine.Batch.from_scenes([sc, S.pyramid(5)]).advance(6).bodies()
    live = [engine.Batch.from_scenes([sc, S.pyramid(5)]) for _ in range(24)]
    for b in live:
        b.advance(6)
    for b in live:
        assert (_state(b.bodies()) == _state(ref)).all()


def test_gpu_batch_world_matches_drop_in_world(engine):
    """A batch of one built by b2a_batch_from_scenes steps exactly like the
    drop-in b2World built from the same scene."""
    sc = S.random_pile(300, seed=21, half_w=10.0, height=25.0)
    b = engine.Batch.from_scenes([sc]).advance(30).bodies()
    w = engine.World(sc)
    w.step(30)
    st = w.bodies()
    assert np.array_equal(bits(b["c"]), bits(st[:, 0:2]))
    assert np.array_equal(bits(b["a"]), bits(st[:, 2]))
    assert np.array_equal(bits(b["v"]), bits(st[:, 3:5]))


def test_gpu_batch_4096_worlds_steps(engine):
    """The batched workload's shape at a reduced world count: worlds of
    lognormal sizes (pyramids and piles) step together; no world loses
    bodies, penetration stays bounded and nothing becomes non-finite."""
    sizes = S.batched_sizes(256)
    scenes = [S.batch_world(i, int(sizes[i])) for i in range(256)]
    batch = engine.Batch.from_scenes(scenes).advance(60)
    c = batch.counts()
    print("256 worlds:", c.bodies, "bodies,", c.contacts, "contacts,", c.touching, "touching,", c.colours, "colours",
          flush=True)
    st = batch.bodies()
    assert np.isfinite(st["c"]).all() and np.isfinite(st["v"]).all()
    assert c.touching > 0 and c.colours <= 65


def test_gpu_dense_pair_queries_match_worlds_alone_bitexact(engine):
    """From 65536 shapes on, the pair queries keep 8 pair slots per lane
    instead of 16 (more shapes are traversed a second time, for occupancy).
    A batch above that size steps its worlds exactly as each world stepped
    alone, whose few shapes take the 16-slot queries."""
    sizes = S.batched_sizes(320)
    scenes = [S.batch_world(i, int(sizes[i])) for i in range(320)]
    shapes = sum(sc.fixture_count for sc in scenes)
    assert shapes >= 65536, shapes
    steps = 20
    got = engine.Batch.from_scenes(scenes).advance(steps)
    print(f"320 worlds, {shapes} shapes, {got.counts().contacts} contacts", flush=True)
    got_b = got.bodies()
    sample = [0, 37, 101, 202, 319] + [int(np.argmax(sizes[:320]))]
    for i in sample:
        alone = engine.Batch.from_scenes([scenes[i]]).advance(steps).bodies()
        part = got_b[got.world_bodies[i]:got.world_bodies[i + 1]]
        same = _state(part) == _state(alone)
        print(f"world {i} ({scenes[i].body_count} bodies): {same.all(axis=1).sum()} of {len(alone)} bit-identical",
              flush=True)
        assert same.all(), i


def test_gpu_full_batch_worlds_match_worlds_alone_bitexact(engine):
    """The batched benchmark at full size: 4096 worlds of lognormal sizes
    (about 1.14M bodies) step together; sampled worlds, the largest among
    them, equal themselves stepped alone bit for bit."""
    sizes = S.batched_sizes(4096)
    scenes = [S.batch_world(i, int(sizes[i])) for i in range(4096)]
    steps = 20
    batch = engine.Batch.from_scenes(scenes).advance(steps)
    c = batch.counts()
    print(f"4096 worlds: {c.bodies} bodies, {c.contacts} contacts, {c.touching} touching, "
          f"solver path {batch.get_option(engine.OPT_LAST_SOLVER)}", flush=True)
    got = batch.bodies()
    for i in [0, 1234, 2048, 3333, 4095, int(np.argmax(sizes))]:
        alone = engine.Batch.from_scenes([scenes[i]]).advance(steps).bodies()
        part = got[batch.world_bodies[i]:batch.world_bodies[i + 1]]
        same = _state(part) == _state(alone)
        print(f"world {i} ({scenes[i].body_count} bodies): {same.all(axis=1).sum()} of {len(alone)} bit-identical",
              flush=True)
        assert same.all(), i


def test_gpu_graph_steps_match_eager_steps(engine):
    """Replaying the captured step graph gives the eager step's results:
    bit-exact while no constraint falls back to the serial group (the colour
    limit and claim rounds of a captured step only differ from the eager
    step's for such constraints)."""
    scenes = [S.random_pile(600, seed=31, half_w=14.0, height=30.0), S.pyramid(10), scene_for("chain")[0]]
    eager = engine.Batch.from_scenes(scenes).advance(25)
    graph = engine.Batch.from_scenes(scenes).advance(25)
    for k in range(40):
        eager.step(8, 3, 0)
        graph.step(8, 3, engine.STEP_GRAPH)
        c = graph.counts()
        assert c.serial == 0, f"step {k}: {c.serial} constraints in the serial group"
    assert graph.counts().graph_steps == 40
    a, b = eager.bodies(), graph.bodies()
    for f in ("xf", "c", "a", "v", "w"):
        assert np.array_equal(bits(a[f]), bits(b[f])), f
    ka = eager.contacts()
    kb = graph.contacts()
    assert np.array_equal(ka["shape_a"], kb["shape_a"]) and np.array_equal(bits(ka["normal_impulse"]),
                                                                            bits(kb["normal_impulse"]))


def test_gpu_graph_steps_match_eager_steps_one_world(engine):
    """The same for a batch of one world (a captured step of one world runs
    the eager step's two device-wide claim rounds, then the finishing
    workgroup): replays equal eager steps bit for bit."""
    sc = S.random_pile(1500, seed=37, half_w=20.0, height=45.0)
    eager = engine.Batch.from_scenes([sc]).advance(25)
    graph = engine.Batch.from_scenes([sc]).advance(25)
    for k in range(40):
        eager.step(8, 3, 0)
        graph.step(8, 3, engine.STEP_GRAPH)
        assert graph.counts().serial == 0, f"step {k}"
    assert graph.counts().graph_steps == 40
    a, b = eager.bodies(), graph.bodies()
    for f in ("xf", "c", "a", "v", "w"):
        assert np.array_equal(bits(a[f]), bits(b[f])), f
    assert eager.contacts().tobytes() == graph.contacts().tobytes()


def test_gpu_graph_steps_async_batch(engine):
    """Async graph replays of a batch (no host sync between steps), then one
    sync: same state as the same number of synchronous graph steps."""
    scenes = [S.batch_world(i, 80 + 40 * i) for i in range(12)]
    a = engine.Batch.from_scenes(scenes).advance(10)
    b = engine.Batch.from_scenes(scenes).advance(10)
    for _ in range(30):
        a.step(8, 3, engine.STEP_GRAPH)
        b.step(8, 3, engine.STEP_GRAPH | engine.STEP_ASYNC)
    b.wait()
    assert b.counts().status == 0
    x, y = a.bodies(), b.bodies()
    assert np.array_equal(bits(x["c"]), bits(y["c"])) and np.array_equal(bits(x["v"]), bits(y["v"]))


def test_gpu_per_world_solver_matches_colour_sweeps_bitexact(engine):
    """The one-workgroup-per-world solver gives the colour-major sweeps'
    results bit for bit, for every workgroup size the autotuner may pick."""
    scenes = _mixed_scenes()[:2] + [S.batch_world(i, 50 + 30 * i) for i in range(10)]
    scenes = [s for s in scenes if s.joint_count == 0]
    ref = engine.Batch.from_scenes(scenes)
    ref.set_option(engine.OPT_SOLVER, engine.SOLVER_COLOURS)
    ref.advance(30)
    assert ref.get_option(engine.OPT_LAST_SOLVER) == 1
    want = _state(ref.bodies())
    for block, lds in ((64, 0), (256, 0), (1024, 0), (128, 1), (512, 1)):
        b = engine.Batch.from_scenes(scenes)
        assert b.get_option(engine.OPT_PER_WORLD_ELIGIBLE) == 1
        b.set_option(engine.OPT_SOLVER, engine.SOLVER_WORLDS)
        b.set_option(engine.OPT_WORLD_BLOCK, block)
        b.set_option(engine.OPT_WORLD_LDS, lds)
        b.advance(30)
        assert b.get_option(engine.OPT_LAST_SOLVER) == 2
        assert b.get_option(engine.OPT_TUNED_WORLD_LDS) == lds
        same = (_state(b.bodies()) == want).all(axis=1)
        print(f"block {block} lds {lds}: {same.sum()} of {len(same)} bodies bit-identical", flush=True)
        assert same.all()


def test_gpu_per_world_autotune_and_graph(engine):
    """Auto mode picks the per-world solver for a batch of small worlds, the
    autotuner settles on a block size within ten steps, and graph replays of
    the per-world step match eager steps."""
    scenes = [S.batch_world(i, 60 + 20 * i) for i in range(16)]
    a = engine.Batch.from_scenes(scenes).advance(12)
    assert a.get_option(engine.OPT_LAST_SOLVER) == 2
    tuned = a.get_option(engine.OPT_TUNED_WORLD_BLOCK)
    print("autotuned world block:", tuned, flush=True)
    assert tuned in (64, 128, 256, 512, 1024)
    b = engine.Batch.from_scenes(scenes).advance(12)
    for k in range(20):
        a.step(8, 3, 0)
        b.step(8, 3, engine.STEP_GRAPH)
        assert b.counts().serial == 0, f"step {k}: constraints left to the serial group"
    assert b.counts().graph_steps == 20
    assert np.array_equal(_state(a.bodies()), _state(b.bodies()))


def test_gpu_snapshot_resume_bitexact(engine, tmp_path):
    """Save a batch mid-run, keep stepping; load the snapshot into a new
    batch and step the same number of times: bit-identical bodies and
    contacts (the contact colours and warm-start impulses are part of the
    snapshot)."""
    scenes = [S.random_pile(300, seed=41, half_w=10.0, height=25.0), scene_for("chain")[0], S.pyramid(8)]
    a = engine.Batch.from_scenes(scenes).advance(25)
    path = str(tmp_path / "batch.b2gsnap")
    a.save(path)
    a.advance(20)
    b = engine.Batch.load(path, like=a)
    b.inv_dt0 = a.inv_dt0
    b.advance(20)
    x, y = a.bodies(), b.bodies()
    assert np.array_equal(_state(x), _state(y))
    ca, cb = a.contacts(), b.contacts()
    assert np.array_equal(ca["shape_a"], cb["shape_a"])
    assert np.array_equal(bits(ca["normal_impulse"]), bits(cb["normal_impulse"]))


@pytest.mark.parametrize("name", ["pile", "chain"])
def test_gpu_snapshot_load_then_step_vs_restatement(engine, tmp_path, name):
    """A batch loaded from a snapshot holds the saved state bit for bit
    (bodies, contacts with manifolds, impulses and colours, joint impulses),
    and its next step is the step the restatement takes from that state in
    the loaded batch's colour order (1e-5 m / 1e-5 rad, as the one-step
    parity tests). The reference's own b2World::Dump
    (src/dynamics/b2_world.cpp:1469-1531) drops the contact state, so the
    restatement is the checker."""
    import oracle
    sc = scene_for(name)[0]
    a = engine.Batch.from_scenes([sc]).advance(70)
    path = str(tmp_path / "one.b2gsnap")
    a.save(path)
    b = engine.Batch.load(path, like=a)
    bodies0, contacts0 = b.bodies(), b.contacts()
    assert np.array_equal(_state(bodies0), _state(a.bodies()))
    assert np.array_equal(contacts0.view(np.uint8), a.contacts().view(np.uint8))
    assert np.array_equal(b.contact_colours(len(contacts0)), a.contact_colours(len(contacts0)))
    joints0 = b.joints(sc.joint_count) if sc.joint_count else None
    if sc.joint_count:
        assert np.array_equal(joints0.view(np.uint8), a.joints(sc.joint_count).view(np.uint8))
    b.step(8, 3, engine.STEP_SKIP_BROADPHASE)
    gb, gc = b.bodies(), b.contacts()
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
    r.set_inv_dt0(np.float32(1.0) / np.float32(1.0 / 60.0))
    from test_gpu_parity import _joint_colours
    r.set_colours(oracle.contact_keys(gc), b.contact_colours(len(gc)), _joint_colours(sc) if sc.joint_count else None)
    r.step(1, flags=engine.STEP_SKIP_BROADPHASE)
    ob = r.bodies()
    dyn = np.asarray(sc.body_type) == 2
    dc = np.abs(gb["c"] - ob["c"])[dyn].max()
    da = np.abs(gb["a"] - ob["a"])[dyn].max()
    print(f"{name}: step after load vs restatement max |dc| {dc:.3g} |da| {da:.3g}; "
          f"touching {(gc['point_count'] > 0).sum()}", flush=True)
    assert dc < 1e-5 and da < 1e-5
    assert np.array_equal(oracle.contact_keys(gc), oracle.contact_keys(r.contacts()))


def _touching_keys(contacts):
    t = contacts[contacts["point_count"] > 0]
    a = np.minimum(t["shape_a"], t["shape_b"]).astype(np.uint64)
    b = np.maximum(t["shape_a"], t["shape_b"]).astype(np.uint64)
    return set(((a << np.uint64(32)) | b).tolist())


def _event_keys(ev, kinds):
    e = ev[np.isin(ev["kind"], kinds)]["contact"]
    a = np.minimum(e["shape_a"], e["shape_b"]).astype(np.uint64)
    b = np.maximum(e["shape_a"], e["shape_b"]).astype(np.uint64)
    return ((a << np.uint64(32)) | b).tolist()


def test_gpu_contact_events_match_touching_transitions(engine):
    """The device event stream (b2ContactListener Begin/EndContact) equals
    the difference of the touching sets before and after each step; Begin
    events carry a touching manifold, End events an empty one (Collide) or
    the last touching one (destroyed by the broadphase); Collide's events
    come before the broadphase's."""
    sc = S.random_pile(400, seed=51, half_w=10.0, height=30.0)
    b = engine.Batch.from_scenes([sc]).advance(3)
    begins = ends = destroyed = 0
    for k in range(60):
        before = _touching_keys(b.contacts())
        b.step(8, 3, engine.STEP_EVENTS)
        after = _touching_keys(b.contacts())
        ev = b.events()
        beg = _event_keys(ev, [engine.EVENT_BEGIN])
        end = _event_keys(ev, [engine.EVENT_END, engine.EVENT_DESTROYED])
        assert len(beg) == len(set(beg)) and len(end) == len(set(end))
        # a contact may begin and end within one step only if destroyed after beginning
        assert set(beg) - set(end) == after - before, f"step {k}"
        assert set(end) - set(beg) == before - after, f"step {k}"
        assert (ev[ev["kind"] == engine.EVENT_BEGIN]["contact"]["point_count"] > 0).all()
        assert (ev[ev["kind"] == engine.EVENT_END]["contact"]["point_count"] == 0).all()
        assert (ev[ev["kind"] == engine.EVENT_DESTROYED]["contact"]["point_count"] > 0).all()
        kinds = ev["kind"].tolist()
        if engine.EVENT_DESTROYED in kinds:
            first = kinds.index(engine.EVENT_DESTROYED)
            assert all(x == engine.EVENT_DESTROYED for x in kinds[first:])
        begins += len(beg)
        ends += int((ev["kind"] == engine.EVENT_END).sum())
        destroyed += int((ev["kind"] == engine.EVENT_DESTROYED).sum())
    print(f"events over 60 steps: {begins} begin, {ends} end, {destroyed} destroyed", flush=True)
    assert begins > 0


def test_gpu_steps_are_deterministic(engine):
    """The same world stepped twice gives the same bits at every step. Island
    labels come from a read-only root walk after the lock-free union: a
    labelling pass that also shortened paths could overwrite a finished label
    with a non-root ancestor and split an island's early exit."""
    pile = S.random_pile(400, seed=12, half_w=12.0, height=30.0)
    a = engine.Batch.from_scenes([pile])
    b = engine.Batch.from_scenes([pile])
    for k in range(40):
        a.advance(1)
        b.advance(1)
        assert np.array_equal(_state(a.bodies()), _state(b.bodies())), f"step {k + 1}"
