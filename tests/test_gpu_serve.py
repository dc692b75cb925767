"""GPU tests of the serving pattern (b2s_run): continuous batching over a
resident pool of worlds.

The key property: a world's result is the same whether its request runs
alone or overlapped with other requests in the same ticks (every stage of a
step works per world, inactive worlds stay untouched, and the colouring rank
is world-local), and the same as stepping that world in a batch of its own.
"""
import numpy as np
import pytest

from util import bits, load_pkg

pytestmark = pytest.mark.gpu
S = load_pkg().scenes
DT60 = np.float32(1.0 / 60.0)
DT240 = np.float32(1.0 / 240.0)


def _scenes(n=10):
    return [S.batch_world(100 + i, 40 + 37 * i) for i in range(n)]


def _requests(rows):
    E = load_pkg().engine
    r = np.zeros(len(rows), dtype=E.REQUEST_DTYPE)
    for k, (t, w, n, dt) in enumerate(rows):
        r[k] = (t, w, n, dt, 0)
    return r


def _world_state(pool, w):
    b = pool.bodies()[pool.world_bodies[w]:pool.world_bodies[w + 1]]
    return np.concatenate([bits(b["xf"]).reshape(len(b), -1), bits(b["c"]).reshape(len(b), -1),
                           bits(b["v"]).reshape(len(b), -1), bits(b["w"]).reshape(len(b), -1)], axis=1)


def _mixed_trace(target_world, target_tick, nworlds, seed=3):
    rng = np.random.default_rng(seed)
    rows = [(float(target_tick), target_world, 6, DT60)]
    for t in range(0, 12):
        for w in rng.choice([x for x in range(nworlds) if x != target_world], size=3, replace=False):
            rows.append((float(t), int(w), int(rng.integers(1, 9)), DT60 if rng.random() < 0.5 else DT240))
    rows.sort(key=lambda r: r[0])
    return _requests(rows)


def _pool(engine, scenes, solver):
    b = engine.Batch.from_scenes(scenes)
    if solver == "colours":  # colour-major sweeps (the warm-start gather over all bodies)
        b.set_option(engine.OPT_SOLVER, engine.SOLVER_COLOURS)
    return b.advance(2)


@pytest.mark.parametrize("solver", ["auto", "colours"])
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_serve_same_result_alone_and_in_mix(engine, graph, solver):
    scenes = _scenes()
    wbytes = np.array([engine.world_device_bytes(s) for s in scenes], np.int64)
    target = 3
    alone = _pool(engine, scenes, solver)
    mix = _pool(engine, scenes, solver)
    own = _pool(engine, [scenes[target]], solver)
    ra, sa, _ = engine.serve(alone, _requests([(0.0, target, 6, DT60)]), 1 << 40, wbytes, graph=graph,
                             realtime=False)
    budget = int(wbytes.max() * 4)  # about four worlds in flight: requests queue for memory
    trace = _mixed_trace(target, 2, len(scenes))
    rm, sm, _ = engine.serve(mix, trace, budget, wbytes, graph=graph, realtime=False)
    print(f"mix: {sm.completed} requests in {sm.ticks} ticks, max {sm.max_in_flight} in flight, "
          f"mean {sm.mean_in_flight:.2f}", flush=True)
    assert sm.completed == len(trace) and sa.completed == 1
    assert sm.max_in_flight > 1                       # requests really overlapped
    assert (rm["completed"] >= rm["admitted"]).all() and (rm["admitted"] >= rm["arrival"]).all()
    assert (rm["ticks"] == trace["substeps"]).all()
    # the target ran 6 substeps in both; its state is bit-identical
    assert np.array_equal(_world_state(alone, target), _world_state(mix, target))
    # ... and equal to that world stepped in a batch of its own
    own.advance(6)
    assert np.array_equal(_world_state(alone, target), _world_state(own, 0))
    # worlds without requests did not move
    untouched = _pool(engine, scenes, solver)
    for w in range(len(scenes)):
        if w != target:
            assert np.array_equal(_world_state(alone, w), _world_state(untouched, w))


def test_gpu_serve_realtime_poisson(engine):
    """Poisson arrivals on the wall clock: every request completes, the
    timestamps are ordered, and latency percentiles are finite."""
    scenes = [S.batch_world(200 + i, int(s)) for i, s in enumerate(S.serving_sizes(64, seed=0xC5))]
    pool = engine.Batch.from_scenes(scenes).advance(1)
    wbytes = np.array([engine.world_device_bytes(s) for s in scenes], np.int64)
    trace = S.serving_trace(len(scenes), rate=2000.0, duration=0.3, seed=0xC5)
    res, st, _ = engine.serve(pool, trace, 64 << 30, wbytes, realtime=True)
    lat = (res["completed"] - res["arrival"]) * 1e3
    print(f"{st.completed} requests, {st.ticks} ticks, p50 {np.percentile(lat, 50):.2f} ms, "
          f"p99 {np.percentile(lat, 99):.2f} ms, {st.body_steps / st.wall_s:.3g} body-steps/s, "
          f"mean in flight {st.mean_in_flight:.1f}", flush=True)
    assert st.completed == len(trace)
    assert (res["completed"] >= res["admitted"]).all() and (res["admitted"] >= res["arrival"]).all()
    assert np.isfinite(lat).all() and lat.min() > 0
    # the ticks' device time (events around each eager step) lies within
    # the host's time inside the step calls, which lies within the wall time
    print(f"device {st.device_s:.4f} s, host in steps {st.busy_s:.4f} s, wall {st.wall_s:.4f} s", flush=True)
    assert 0.0 < st.device_s <= st.busy_s <= st.wall_s


def test_gpu_serve_full_pool_same_result_as_alone(engine):
    """The serving workload's full resident pool: 10,000 worlds (16-4096
    bodies, log-uniform; about 6.3M bodies). Worlds that sit out ticks keep
    their pairs and their part of the tree, queries take the dense path, and
    tree rebuilds sort only the active worlds. A target world's request runs
    among several hundred others; its state equals the world stepped alone."""
    sizes = S.serving_sizes(10000, seed=0xC5)
    print("building the 10,000-world pool", flush=True)
    scenes = [S.batch_world(200 + i, int(s)) for i, s in enumerate(sizes)]
    wbytes = np.array([engine.world_device_bytes(s) for s in scenes], np.int64)
    pool = engine.Batch.from_scenes(scenes).advance(2)
    targets = [int(np.argmax(sizes)), 17, 5003]
    rng = np.random.default_rng(11)
    rows = [(2.0, t, 6, DT60) for t in targets]
    for t in range(0, 10):
        for w in rng.choice(10000, size=40, replace=False):
            if int(w) not in targets:
                rows.append((float(t), int(w), int(rng.integers(1, 9)), DT60 if rng.random() < 0.5 else DT240))
    rows.sort(key=lambda r: r[0])
    trace = _requests(rows)
    res, st, _ = engine.serve(pool, trace, 64 << 30, wbytes, realtime=False)
    print(f"{st.completed} requests in {st.ticks} ticks, max {st.max_in_flight} in flight", flush=True)
    assert st.completed == len(trace) and st.max_in_flight > 100
    for t in targets:
        own = engine.Batch.from_scenes([scenes[t]]).advance(2)
        own.advance(6)
        same = _world_state(pool, t) == _world_state(own, 0)
        print(f"world {t} ({scenes[t].body_count} bodies): {same.all(axis=1).sum()} of {len(same)} bit-identical",
              flush=True)
        assert same.all(), t
