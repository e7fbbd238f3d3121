"""Arena pools on the GPU: galaxies placed in and removed from a byte-budget
arena between steps, every slot's result bit-identical to the same galaxy
stepped alone.

* three galaxies placed, stepped, one removed and a galaxy of another size
  placed in the freed ranges, stepped again (the launch tables and step
  graphs follow the occupied slots);
* a refresh-only step (sort_every 2) right after a slot was emptied: the
  emptied slot's tree scalars are gone, so nothing is refreshed in ranges it
  gave back;
* the arena refuses a galaxy its free ranges cannot hold (GS_E_FULL) and
  leaves the resident ones untouched; an empty arena steps as a no-op;
* arena_info's used bytes are the resident galaxies' ranges."""
import numpy as np
import pytest

import galaxy_hip
from galaxy_hip import ArenaFull, arena_charge, ic
from gpu_util import banner

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=0.005, G=1.0, epsilon=0.01, theta=0.6)


def _galaxies():
    return [ic.disk(30000, seed=11), ic.plummer(9000, seed=12), ic.disk(17000, seed=13),
            ic.plummer(24000, seed=14)]


def _alone(data, schedule, exact):
    """data stepped alone: schedule = [(nsteps, sort_every), ...]."""
    sim = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=exact)
    try:
        sim.update_params(**PARAMS)
        for nsteps, every in schedule:
            sim.step(nsteps, sort_every=every)
        return sim.particles(), sim.reduce()
    finally:
        sim.destroy()


def _same(got, want):
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("exact", [False, True])
def test_arena_place_remove_reuse(exact):
    banner(f"arena place / remove / reuse exact={exact}")
    g = _galaxies()
    budget = sum(arena_charge(x.shape[0]) for x in g[:3]) + (1 << 20)
    with galaxy_hip.GalaxyArena(budget * 4, 4, exact_force=exact) as ar:
        assert ar.arena_info()["device_bytes"] <= budget * 4
        for k in range(3):
            ar.upload(k, g[k], **PARAMS)
        assert ar.arena_info()["used_bytes"] == sum(arena_charge(x.shape[0]) for x in g[:3])
        ar.step(2, sort_every=1)
        red0 = ar.reduce(0)
        got = {k: ar.particles(k) for k in range(3)}
        for k in range(3):
            want, _ = _alone(g[k], [(2, 1)], exact)
            _same(got[k], want)
        ar.clear(1)  # the Plummer sphere leaves; slot 0's reductions stay readable
        assert ar.reduce(0) == red0
        ar.upload(1, g[3], **PARAMS)  # a larger galaxy in the freed slot
        ar.step(3, sort_every=1)
        w0, r0 = _alone(g[0], [(5, 1)], exact)
        w1, r1 = _alone(g[3], [(3, 1)], exact)
        w2, _ = _alone(g[2], [(5, 1)], exact)
        _same(ar.particles(0), w0)
        _same(ar.particles(1), w1)
        _same(ar.particles(2), w2)
        for key, v in r1.items():
            assert ar.reduce(1)[key] == pytest.approx(v, rel=1e-12, abs=1e-300)


def test_refresh_only_step_after_a_slot_is_emptied():
    banner("arena refresh-only step after a clear")
    g = _galaxies()
    with galaxy_hip.GalaxyArena(1 << 30, 3) as ar:
        ar.upload(0, g[0], **PARAMS)
        ar.upload(1, g[1], **PARAMS)
        ar.upload(2, g[2], **PARAMS)
        ar.step(1, sort_every=1)
        ar.clear(0)  # slot 0 held the arena's first ranges
        ar.step(1, sort_every=2)  # step count 1: refresh the stale trees, no rebuild
        ar.step(2, sort_every=1)
        for k in (1, 2):
            want, _ = _alone(g[k], [(1, 1), (1, 2), (2, 1)], False)
            _same(ar.particles(k), want)


def test_full_arena_refuses_and_keeps_residents():
    banner("arena full")
    g = _galaxies()
    budget = arena_charge(30000) + arena_charge(17000) + (4 << 20)
    with galaxy_hip.GalaxyArena(budget, 4) as ar:
        ar.step(1)  # empty: nothing to launch
        ar.upload(0, g[0], **PARAMS)
        ar.upload(1, g[2], **PARAMS)
        fit = ar.arena_info()["largest_fit"]
        assert fit < 24000
        with pytest.raises(ArenaFull):
            ar.upload(2, g[3], **PARAMS)
        assert ar.count(2) == 0
        ar.step(2, sort_every=1)
        _same(ar.particles(0), _alone(g[0], [(2, 1)], False)[0])
        _same(ar.particles(1), _alone(g[2], [(2, 1)], False)[0])
        ar.clear(0)
        ar.upload(2, g[3], **PARAMS)  # fits in the range slot 0 gave back
        ar.step(1, sort_every=1)
        _same(ar.particles(2), _alone(g[3], [(1, 1)], False)[0])
        _same(ar.particles(1), _alone(g[2], [(3, 1)], False)[0])
