"""GPU slot pools (many galaxies per launch) against the same galaxies run
alone, through the C ABI.

A slot's particles map to the same wavefront lanes as on a handle of its own
(slot ranges start on 1024-particle boundaries) and every per-slot stage is
the single-galaxy algorithm on slot-local indices, so the results are
bit-identical to running alone — in the exact and in the rsqrt force path,
with and without graph replay, and after a slot is emptied and refilled."""
import numpy as np
import pytest

import galaxy_hip
from galaxy_hip import ic
from gpu_util import banner, oracle_tree
from oracle import oracle

pytestmark = pytest.mark.gpu

# (capacity, n, preset, bucket, params)
JOBS = [
    (4096, 3000, "disk", 16, dict(theta=0.6, epsilon=0.01)),
    (20000, 17000, "plummer", 32, dict(theta=0.7, epsilon=0.02)),
    (5000, 0, "disk", 16, dict(theta=0.6, epsilon=0.01)),  # empty slot
    (65536, 40000, "disk", 8, dict(theta=0.5, epsilon=0.005, haloMass=2.0, haloA=1.5,
                                   gasDrag=0.25, haloCenters=[(0.1, 0.0, 0.0)])),
    (1024, 1, "plummer", 16, dict(theta=0.6, epsilon=0.01)),
    (9000, 8191, "plummer", 16, dict(theta=0.0, epsilon=0.01)),
]


def _galaxy(preset, n, seed):
    if n == 0:
        return np.zeros((0, 8), np.float32)
    if preset == "plummer":
        return ic.plummer(n, seed=ic.SEED_BASE + 100 + seed)
    return ic.disk(n, seed=ic.SEED_BASE + 200 + seed, gas_fraction=0.3)


def _params(extra):
    p = dict(dt=0.005, G=1.0)
    p.update(extra)
    return p


def _alone(data, bucket, params, exact, nsteps, sort_every=1):
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=exact)
    try:
        sim.update_params(**params)
        sim.step(nsteps, sort_every=sort_every)
        return sim.particles(), sim.energy(), sim.reduce()
    finally:
        sim.destroy()


def _make_pool(exact, graph=True, jobs=JOBS, timing=False):
    pool = galaxy_hip.GalaxyPool([j[0] for j in jobs], buckets=[j[3] for j in jobs],
                                 exact_force=exact, graph=graph, timing=timing)
    datas = []
    for k, (cap, n, preset, bucket, extra) in enumerate(jobs):
        data = _galaxy(preset, n, k)
        datas.append(data)
        pool.upload(k, data, **_params(extra))
    return pool, datas


@pytest.mark.parametrize("exact", [True, False])
def test_pool_slots_equal_alone(exact):
    banner(f"pool vs alone exact={exact}")
    pool, datas = _make_pool(exact)
    try:
        pool.step(3, sort_every=1)
        got = [(pool.particles(k), pool.energy(k) if datas[k].shape[0] else None,
                pool.reduce(k) if datas[k].shape[0] else None) for k in range(len(JOBS))]
    finally:
        pool.destroy()
    for k, (cap, n, preset, bucket, extra) in enumerate(JOBS):
        if n == 0:
            assert got[k][0].shape == (0, 8)
            continue
        banner(f"  slot {k} n={n} {preset} bucket={bucket}")
        parts, energy, red = _alone(datas[k], bucket, _params(extra), exact, 3)
        assert np.array_equal(got[k][0].view(np.uint32), parts.view(np.uint32))
        assert np.array_equal(got[k][1].view(np.uint32), energy.view(np.uint32))
        for key in red:
            assert got[k][2][key] == pytest.approx(red[key], rel=1e-12, abs=1e-300)


def test_pool_slot_matches_oracle_steps():
    """An exact-force pool slot follows the oracle's sort+build+walk steps."""
    banner("pool slot vs oracle")
    pool, datas = _make_pool(True)
    try:
        pool.step(2, sort_every=1)
        parts = pool.particles(1)
        energy = pool.energy(1)
    finally:
        pool.destroy()
    cap, n, preset, bucket, extra = JOBS[1]
    op = oracle.pack_params(n, dt=0.005, G=1.0, epsilon=extra["epsilon"], theta=extra["theta"])
    cur = datas[1].reshape(-1).copy()
    went = None
    for _ in range(2):
        ref = oracle.sort_and_build(cur, n, bucket=bucket)
        cur, went, _, _, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
    assert np.array_equal(parts.reshape(-1).view(np.uint32), cur.view(np.uint32))
    assert np.array_equal(energy.view(np.uint32), went.view(np.uint32))


def test_pool_tree_bit_exact():
    """Per-slot permutation and topology equal mortonTree.js (via the oracle)."""
    banner("pool trees vs oracle")
    pool, datas = _make_pool(False)
    try:
        pool.sort()
        pool.refresh()
        for k, (cap, n, preset, bucket, extra) in enumerate(JOBS):
            if n == 0:
                continue
            ref = oracle_tree(datas[k], bucket=bucket)
            assert np.array_equal(pool.perm(k), ref["perm"])
            assert np.array_equal(pool.particles(k).reshape(-1), ref["sorted"])
            info = pool.tree_info(k)
            assert info["num_nodes"] == ref["num_nodes"]
            assert np.array_equal(pool.nodes(k), ref["nodes"])
    finally:
        pool.destroy()


def test_pool_graph_equals_direct_launches():
    banner("pool graph vs direct")
    out = []
    for graph in (True, False):
        pool, datas = _make_pool(False, graph=graph)
        try:
            pool.step(4, sort_every=2)
            out.append([pool.particles(k) for k in range(len(JOBS))])
        finally:
            pool.destroy()
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_timed_replays_equal_plain_replays():
    """A timing handle replays each step as two graphs (rebuild + refresh,
    walk) with events between them: same particles as one graph per step and
    as direct launches, for a pool and for a single galaxy; the walk and the
    whole step are timed once per step, the other stages by direct launches."""
    banner("timed replays")
    out = []
    for graph, timing in ((True, False), (True, True), (False, True)):
        pool, datas = _make_pool(False, graph=graph, timing=timing)
        try:
            for _ in range(3):  # stage times are folded in at each sync
                pool.step(1, sort_every=2)
                pool.sync()
            out.append([pool.particles(k) for k in range(len(JOBS))])
            st = pool.stage_times()
        finally:
            pool.destroy()
        if timing:
            assert st["walk"][1] == 3 and st["step"][1] == 3 and st["walk"][0] > 0
            assert (st["sort"][1] > 0) == (not graph)
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    data = _galaxy("disk", 20000, 7)
    res = []
    for graph, timing in ((True, False), (True, True), (False, True)):
        sim = galaxy_hip.NBodyGPU(data, bucket=16, graph=graph, timing=timing)
        try:
            sim.update_params(**_params(dict(theta=0.6, epsilon=0.01)))
            sim.step(3, sort_every=2)
            res.append(sim.particles())
        finally:
            sim.destroy()
    for r in res[1:]:
        assert np.array_equal(res[0].view(np.uint32), r.view(np.uint32))


def test_pool_slot_reuse():
    """Empty a slot mid-run and refill it with a galaxy of another size; the
    new galaxy and the untouched slots still equal their runs alone."""
    banner("pool slot reuse")
    pool, datas = _make_pool(False)
    try:
        pool.step(2, sort_every=1)
        pool.clear(3)
        fresh = ic.plummer(61000, seed=ic.SEED_BASE + 999)
        fparams = _params(dict(theta=0.6, epsilon=0.01))
        pool.upload(3, fresh, **fparams)
        pool.set_tree(3, 16)
        pool.step(2, sort_every=1)
        got_fresh = pool.particles(3)
        got_1 = pool.particles(1)
        assert pool.count(3) == 61000
    finally:
        pool.destroy()
    parts, _, _ = _alone(fresh, 16, fparams, False, 2)
    assert np.array_equal(got_fresh.view(np.uint32), parts.view(np.uint32))
    cap, n, preset, bucket, extra = JOBS[1]
    parts1, _, _ = _alone(datas[1], bucket, _params(extra), False, 4)
    assert np.array_equal(got_1.view(np.uint32), parts1.view(np.uint32))


def test_pool_radix10_equals_radix8():
    """The 10-bit-digit sort gives the same permutation, so whole steps agree."""
    banner("pool radix 10 vs 8")
    out = []
    for radix in (8, 10):
        pool = galaxy_hip.GalaxyPool([j[0] for j in JOBS], buckets=[j[3] for j in JOBS],
                                     radix_bits=radix)
        try:
            for k, (cap, n, preset, bucket, extra) in enumerate(JOBS):
                pool.upload(k, _galaxy(preset, n, k), **_params(extra))
            pool.step(3, sort_every=1)
            out.append([(pool.particles(k), pool.perm(k) if JOBS[k][1] else None)
                        for k in range(len(JOBS))])
        finally:
            pool.destroy()
    for (pa, qa), (pb, qb) in zip(*out):
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
        if qa is not None:
            assert np.array_equal(qa, qb)


def test_pool_rejects_bad_sizes():
    banner("pool argument checks")
    pool = galaxy_hip.GalaxyPool([1024, 2048])
    try:
        with pytest.raises(galaxy_hip.GalaxyError):
            pool.upload(0, np.zeros((1025, 8), np.float32))
        with pytest.raises(galaxy_hip.GalaxyError):
            pool.upload(2, np.zeros((10, 8), np.float32))
        with pytest.raises(galaxy_hip.GalaxyError):
            pool.set_tree(1, 16, max_nodes=10 ** 6)
        assert pool.info()["nslots"] == 2
    finally:
        pool.destroy()
    with pytest.raises(galaxy_hip.GalaxyError):
        galaxy_hip.GalaxyPool([1024], radix_bits=9)
