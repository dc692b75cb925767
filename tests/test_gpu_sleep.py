"""GPU tests of sleeping (b2World::SetAllowSleeping(true), b2_island.cpp:
420-454; b2_world.cpp:526-575).

The scene: an 8-row pyramid that settles and sleeps, a ball dropped onto it
later (it wakes the island), a resting box that may not sleep, a box created
asleep in the air with a box falling onto it, and a 4-link chain hanging at
rest (a joint island that sleeps).

The device updates every contact from the awake flags at the start of
Collide (in parallel); the CPU restatement does the same when no walk order
is given, so one device step and one restated step from the same state, in
the device's constraint order, agree: awake flags and sleep timers exactly,
bodies to the bound of the other one-step tests (1e-5 m / 1e-5 rad).
"""
import sys

import numpy as np
import pytest

import oracle
from util import GOLDEN, bits, golden, load_pkg

pytestmark = pytest.mark.gpu
S = load_pkg().scenes
DT = np.float32(1.0 / 60.0)


def _sleep_scene():
    sys.path.insert(0, GOLDEN)
    import make_golden
    return make_golden.sleep_scene()


def _joint_colours(sc):
    from test_gpu_parity import _joint_colours as jc
    return jc(sc)


def _awake(rec, sc):
    a = (rec["flags"] & 1) == 0
    a[np.asarray(sc.body_type) == S.STATIC] = False
    return a


# steps compared one by one: the chain and the pyramid falling asleep, the
# created-asleep box woken, the ball reaching the pyramid, and a few between
CHECKED = set(range(2, 50)) | set(range(74, 90)) | set(range(140, 155)) | set(range(220, 232)) | \
    set(range(326, 345)) | {100, 200, 300, 400, 420}


def test_gpu_sleep_steps_vs_restatement_in_gpu_order(engine):
    sc = _sleep_scene()
    b = engine.Batch.from_scenes([sc])
    r = oracle.RestatedWorld(sc)
    r.set_allow_sleep(True)
    jc = _joint_colours(sc)
    dyn = np.asarray(sc.body_type) == S.DYNAMIC
    worst = 0.0
    transitions = 0
    for k in range(1, 421):
        check = k in CHECKED
        if check:
            bodies0, contacts0 = b.bodies(), b.contacts()
            joints0 = b.joints(sc.joint_count)
        b.advance(1, sleep=[1])
        if not check:
            continue
        gb, gc = b.bodies(), b.contacts()
        colours = b.contact_colours(len(gc))
        rb = r.bodies()
        for f in ("c", "a", "v", "w", "xf", "flags", "sleep_time", "force", "torque"):
            rb[f] = bodies0[f]
        r.set_bodies(rb)
        r.set_contacts(contacts0)
        j = r.joints()
        for f in ("impulse", "motor_impulse", "lower_impulse", "upper_impulse"):
            j[f] = joints0[f]
        r.set_joints(j)
        r.set_inv_dt0(np.float32(1.0) / DT)
        r.set_colours(oracle.contact_keys(gc), colours, jc)
        r.step(1)
        ob = r.bodies()
        a_dev, a_cpu = _awake(gb, sc), _awake(ob, sc)
        assert np.array_equal(a_dev, a_cpu), f"step {k}: awake flags differ at {np.nonzero(a_dev != a_cpu)[0]}"
        assert np.array_equal(gb["flags"], ob["flags"]), f"step {k}"
        assert np.array_equal(bits(gb["sleep_time"]), bits(ob["sleep_time"])), \
            f"step {k}: sleep timers differ at {np.nonzero(gb['sleep_time'] != ob['sleep_time'])[0]}"
        transitions += int((_awake(bodies0, sc) != a_dev).sum())
        d = max(float(np.abs(gb["c"] - ob["c"])[dyn].max()), float(np.abs(gb["a"] - ob["a"])[dyn].max()))
        worst = max(worst, d)
        assert d < 1e-5, f"step {k}: max |dc|,|da| {d:.3g}"
        asleep = ~a_dev & dyn
        still = asleep & ~_awake(bodies0, sc)
        # a sleeping body keeps its transform and has no velocity (b2Body::SetAwake(false))
        assert np.array_equal(bits(gb["c"][still]), bits(bodies0["c"][still]))
        assert not np.any(gb["v"][asleep]) and not np.any(gb["w"][asleep])
        oc = r.contacts()
        assert np.array_equal(oracle.contact_keys(gc), oracle.contact_keys(oc)), f"step {k}: contact lists differ"
    print(f"sleep scene: {len(CHECKED)} steps compared, {transitions} awake-flag changes among them, "
          f"max |dc|,|da| {worst:.3g}", flush=True)
    assert transitions >= 40  # the chain, the pyramid, the boxes and the wake-up all happened in checked steps


def test_gpu_drop_in_world_sleeps_and_wakes_like_the_reference(engine):
    """b2World with sleeping on (SetAllowSleeping(true)) over 420 steps: at
    steps well away from any transition the set of awake bodies equals the
    reference's (from its recorded run), and bodies asleep between two
    checkpoints did not move."""
    g = golden("sleep_trajectory.npz")
    sc = _sleep_scene()
    w = engine.World(sc)
    w.set_allow_sleep(True)
    ref_awake = g["awake"].astype(bool)
    checkpoints = (100, 190, 300, 400)
    prev = None
    for k in range(1, 421):
        w.step(1)
        if k in checkpoints:
            bd = w.batch().bodies()
            aw = _awake(bd, sc)
            print(f"step {k}: {int(aw.sum())} awake (reference {int(ref_awake[k - 1].sum())})", flush=True)
            assert np.array_equal(aw, ref_awake[k - 1]), f"step {k}: awake sets differ at {np.nonzero(aw != ref_awake[k - 1])[0]}"
            if prev is not None:
                still = ~prev[1] & ~aw & (np.asarray(sc.body_type) == S.DYNAMIC)
                assert np.array_equal(bits(bd["c"][still]), bits(prev[0]["c"][still]))
            prev = (bd, aw)
