"""GPU parity of the tree walk and integrator (nbody_bh.wgsl `main`) against
the FP32 transliteration in the oracle, through the C ABI.

EXACT mode (IEEE sqrt and division in every interaction) must be bit-identical
to the transliteration: accelerations, potential, density, integrated
particles and specific energies.  The default rsqrt path must stay within
the stated force tolerance, 1e-5 relative per particle."""
import numpy as np
import pytest

import galaxy_hip
from galaxy_hip import ic
from gpu_util import banner, oracle_tree, rel_err
from oracle import oracle

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5  # relative force error of the default (rsqrt) path


def _params(n, **kw):
    base = dict(dt=0.005, G=1.0, epsilon=0.02, theta=0.7)
    base.update(kw)
    halo = base.pop("haloCenters", ())
    op = oracle.pack_params(n, dt=base["dt"], G=base["G"], epsilon=base["epsilon"],
                            theta=base["theta"], halo_mass=base.get("haloMass", 0.0),
                            halo_a=base.get("haloA", 1.0), gas_drag=base.get("gasDrag", 0.0),
                            halo_centers=halo)
    return base, halo, op


def _gpu_force(data, bucket, exact, **kw):
    base, halo, _ = _params(data.shape[0], **kw)
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=exact)
    try:
        sim.update_params(haloCenters=halo, **base)
        sim.sort()
        sim.refresh()
        sim.force()
        acc, dens = sim.accel()
        sorted_ = sim.particles()
    finally:
        sim.destroy()
    return acc, dens, sorted_


# eps 0.5: nodes below eps/2 in size are accepted by the s^2 < eps^2/4 test
@pytest.mark.parametrize("theta,eps", [(0.7, 0.02), (0.5, 0.005), (0.0, 0.02), (0.7, 0.5)])
def test_exact_force_bit_identical(theta, eps):
    banner(f"exact force theta={theta} eps={eps}")
    data = ic.make("cpu_disk_16k")
    ref = oracle_tree(data, bucket=16)
    acc, dens, sorted_ = _gpu_force(data, 16, True, theta=theta, epsilon=eps)
    assert np.array_equal(sorted_.reshape(-1), ref["sorted"])
    _, _, op = _params(data.shape[0], theta=theta, epsilon=eps)
    _, _, racc, rdens, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
    assert np.array_equal(acc.view(np.uint32), racc.view(np.uint32))
    assert np.array_equal(dens.view(np.uint32), rdens.view(np.uint32))


def test_fast_force_within_tolerance():
    banner("fast force tolerance")
    data = ic.make("cpu_disk_16k")
    ref = oracle_tree(data, bucket=16)
    acc, dens, _ = _gpu_force(data, 16, False, theta=0.7, epsilon=0.02)
    _, _, op = _params(data.shape[0], theta=0.7, epsilon=0.02)
    _, _, racc, rdens, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
    err = rel_err(acc[:, :3], racc[:, :3])
    banner(f"fast path: max rel err {err.max():.3e}, p99 {np.quantile(err, 0.99):.3e}")
    assert err.max() <= FAST_TOL
    perr = np.abs(acc[:, 3].astype(np.float64) - racc[:, 3]) / np.abs(racc[:, 3].astype(np.float64))
    assert perr.max() <= FAST_TOL  # pair potential (feeds the energy diagnostics)
    assert np.array_equal(dens, rdens)


def _oracle_steps(data, nsteps, bucket, op):
    cur = data.reshape(-1).copy()
    energy = None
    for _ in range(nsteps):
        ref = oracle.sort_and_build(cur, data.shape[0], bucket=bucket)
        pout, energy, _, _, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
        cur = pout
    return cur, energy


@pytest.mark.parametrize("name,n,extra", [
    ("cpu_disk_16k", 16384, {}),
    ("disk_500k", 30000, dict(haloMass=5.0, haloA=3.0, gasDrag=0.25,
                              haloCenters=[(0.01, -0.02, 0.0)])),
    ("collision_4m", 40000, dict(theta=0.5, epsilon=0.005, haloMass=1.0, haloA=2.0,
                                 haloCenters=[(-4.0, 0.0, 0.0), (4.0, 0.0, 0.0)])),
])
def test_exact_steps_bit_identical(name, n, extra):
    """5 full steps (sort + refresh + walk + kick-drift) equal the oracle's."""
    banner(f"exact steps {name} n={n}")
    data = ic.make(name, n=n)
    bucket = 32 if name == "collision_4m" else 16
    base, halo, op = _params(n, **extra)
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=True)
    try:
        sim.update_params(haloCenters=halo, **base)
        sim.step(5, sort_every=1)
        parts = sim.particles()
        energy = sim.energy()
        red = sim.reduce()
    finally:
        sim.destroy()
    want, went = _oracle_steps(data, 5, bucket, op)
    assert np.array_equal(parts.reshape(-1).view(np.uint32), want.view(np.uint32))
    assert np.array_equal(energy.view(np.uint32), went.view(np.uint32))
    # FP64 reductions of the last step
    p = parts.astype(np.float64)
    m = p[:, 3]
    v2 = (p[:, 4] ** 2 + p[:, 5] ** 2) + p[:, 6] ** 2
    assert red["mass"] == pytest.approx(m.sum(), rel=1e-12)
    assert red["px"] == pytest.approx((m * p[:, 4]).sum(), rel=1e-9, abs=1e-12)
    assert red["espec"] == pytest.approx(energy.astype(np.float64).sum(), rel=1e-12)
    assert red["ekin"] == pytest.approx(0.5 * (m * v2).sum(), rel=1e-6)


def test_interaction_count_matches_oracle():
    banner("interaction count")
    data = ic.make("cpu_disk_16k")
    ref = oracle_tree(data, bucket=16)
    base, halo, op = _params(data.shape[0], theta=0.7, epsilon=0.02)
    sim = galaxy_hip.NBodyGPU(data, bucket=16)
    try:
        sim.update_params(**base)
        sim.sort()
        sim.refresh()
        total, per = sim.count_interactions(per_particle=True)
    finally:
        sim.destroy()
    _, _, _, _, inter = oracle.walk(ref["sorted"], ref["nodes"], op)
    assert np.array_equal(per, inter)
    assert total == int(inter.astype(np.uint64).sum())


def test_pair_force_closed_form():
    """Two particles in one leaf: a = G m r / (r^2 + eps^2)^{3/2}."""
    banner("pair closed form")
    data = np.zeros((2, 8), np.float32)
    data[0, :4] = (0.0, 0.0, 0.0, 2.0)
    data[1, :4] = (3.0, 4.0, 0.0, 0.5)
    acc, _, srt = _gpu_force(data, 16, True, theta=0.5, epsilon=0.1)
    for i in range(2):
        j = 1 - i
        r = srt[j, :3].astype(np.float64) - srt[i, :3]
        d2 = (r ** 2).sum() + 0.1 ** 2
        want = srt[j, 3] * r / d2 ** 1.5
        np.testing.assert_allclose(acc[i, :3], want, rtol=2e-7)


def test_far_node_is_its_monopole():
    """A distant cluster enters as one node: a = G M (com - x) / (d^2 + eps^2)^{3/2}."""
    banner("far monopole closed form")
    rng = np.random.default_rng(3)
    n = 200
    data = np.zeros((n + 1, 8), np.float32)
    data[:n, :3] = rng.uniform(-0.01, 0.01, (n, 3)) + np.array([100.0, 0, 0])
    data[:n, 3] = rng.uniform(0.5, 1.5, n)
    data[n, :4] = (0.0, 0.0, 0.0, 1.0)
    acc, _, srt = _gpu_force(data, 16, True, theta=0.5, epsilon=0.01)
    k = int(np.nonzero(srt[:, 0] < 50)[0][0])
    others = np.arange(n + 1) != k
    M = srt[others, 3].astype(np.float64).sum()
    com = (srt[others, :3].astype(np.float64) * srt[others, 3:4]).sum(0) / M
    r = com - srt[k, :3]
    want = M * r / ((r ** 2).sum() + 1e-4) ** 1.5
    np.testing.assert_allclose(acc[k, :3], want, rtol=1e-5)


def test_theta_zero_converges_to_fp64_direct_sum():
    banner("theta 0 vs FP64 direct sum")
    data = ic.make("cpu_disk_16k", n=4096)
    acc, _, srt = _gpu_force(data, 16, False, theta=0.0, epsilon=0.02)
    idx = np.arange(0, 4096, 7, dtype=np.uint32)
    ds = oracle.direct_sum_f64(srt.reshape(-1), 1.0, np.float64(np.float32(0.02 * 0.02)), idx)
    err = rel_err(acc[idx, :3], ds)
    # theta = 0 still takes sub-softening nodes (s < eps/2) as monopoles
    assert np.median(err) < 1e-5
    assert err.max() < 1e-3


def test_hernquist_halo_closed_form():
    banner("halo closed form")
    data = np.zeros((1, 8), np.float32)
    data[0, :4] = (3.0, 0.0, 4.0, 1.0)
    sim = galaxy_hip.NBodyGPU(data, exact_force=True)
    try:
        sim.update_params(dt=1e-3, G=1.0, epsilon=0.1, theta=0.5, haloMass=5.0, haloA=2.0,
                          haloCenters=[(0.0, 0.0, 0.0)])
        sim.step(1, sort_every=1)
        p = sim.particles()
    finally:
        sim.destroy()
    r = np.array([3.0, 0.0, 4.0])
    a = -5.0 * r / (5.0 * (5.0 + 2.0) ** 2)
    np.testing.assert_allclose(p[0, 4:7], a * 1e-3, rtol=1e-6)


@pytest.mark.parametrize("exact", [True, False])
def test_massless_nodes_are_skipped(exact):
    """Zero-mass particles make massless leaves and subtrees, which every lane
    skips (nbody_bh.wgsl:86).  The refresh flags such a tree and the walk then
    runs its variant with the massless-node test: forces against the oracle
    after the on-device sort and after the host-tree upload (no refresh), then
    again once the masses are back (the refresh clears the flag)."""
    banner(f"massless nodes exact={exact}")
    data = ic.make("cpu_disk_16k").copy()
    data[data[:, 0] > 0.8, 3] = 0.0
    ref = oracle_tree(data, bucket=16)
    masses = ref["nodes"].reshape(-1, 16)[:, 3].view(np.float32)
    assert (masses == 0).sum() > 10  # whole leaves and subtrees are massless
    _, _, op = _params(data.shape[0], theta=0.7, epsilon=0.02)
    _, _, racc, rdens, _ = oracle.walk(ref["sorted"], ref["nodes"], op)

    def check(acc, dens):
        if exact:
            assert np.array_equal(acc.view(np.uint32), racc.view(np.uint32))
        else:
            assert rel_err(acc[:, :3], racc[:, :3]).max() <= FAST_TOL
        assert np.array_equal(dens.view(np.uint32), rdens.view(np.uint32))

    acc, dens, _ = _gpu_force(data, 16, exact, theta=0.7, epsilon=0.02)
    check(acc, dens)
    base, _, _ = _params(data.shape[0], theta=0.7, epsilon=0.02)
    sim = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=exact)
    try:
        sim.update_params(**base)
        sim.apply_permutation(ref["perm"])
        sim.upload_tree(ref["nodes"], ref["num_nodes"], ref["level_ranges"])
        sim.force()
        check(*sim.accel())
        # masses back: the same handle's refreshed tree has no massless node
        full = ic.make("cpu_disk_16k")
        ref2 = oracle_tree(full, bucket=16)
        _, _, racc2, rdens2, _ = oracle.walk(ref2["sorted"], ref2["nodes"], op)
        sim.upload(full)
        sim.sort()
        sim.refresh()
        sim.force()
        acc2, dens2 = sim.accel()
    finally:
        sim.destroy()
    if exact:
        assert np.array_equal(acc2.view(np.uint32), racc2.view(np.uint32))
    else:
        assert rel_err(acc2[:, :3], racc2[:, :3]).max() <= FAST_TOL
    assert np.array_equal(dens2.view(np.uint32), rdens2.view(np.uint32))


@pytest.mark.parametrize("name,nsteps", [("disk_500k", 5), ("collision_4m", 1)])
def test_exact_steps_at_full_size(name, nsteps):
    """Full EXACT steps (sort + refresh + walk + halos + gas drag + kick-drift)
    at the workloads' real sizes: config 2, the 500,000-particle disk with 30 %
    gas, a Hernquist halo (M 5, a 3) at the origin and drag 0.25; config 3,
    the 4,194,304-particle collision.  Every particle and specific energy
    equals the oracle's bit for bit."""
    c = ic.CONFIGS[name]
    banner(f"exact full-size steps {name} n={c['n']} steps={nsteps}")
    data = ic.make(name)
    n = data.shape[0]
    extra = dict(theta=c["theta"], epsilon=c["epsilon"])
    if name == "disk_500k":
        extra.update(haloMass=c["halo_mass"], haloA=c["halo_a"], gasDrag=c["gas_drag"],
                     haloCenters=[(0.0, 0.0, 0.0)])
    base, halo, op = _params(n, **extra)
    sim = galaxy_hip.NBodyGPU(data, bucket=c["bucket"], exact_force=True)
    try:
        sim.update_params(haloCenters=halo, **base)
        sim.step(nsteps, sort_every=1)
        parts = sim.particles()
        energy = sim.energy()
    finally:
        sim.destroy()
    want, went = _oracle_steps(data, nsteps, c["bucket"], op)
    assert np.array_equal(parts.reshape(-1).view(np.uint32), want.view(np.uint32))
    assert np.array_equal(energy.view(np.uint32), went.view(np.uint32))


def test_exact_rebuild_cadence():
    """The reference's frame schedule (main.js:401-477): the topology is rebuilt
    every k-th step and the steps in between refresh the stale tree's COMs and
    walk it.  Six EXACT steps with a rebuild every third equal the oracle
    replaying the same schedule bit for bit."""
    banner("exact rebuild cadence, sort every 3 steps")
    data = ic.make("cpu_disk_16k", n=20000)
    n = data.shape[0]
    base, halo, op = _params(n, theta=0.6, epsilon=0.01, gasDrag=0.25)
    sim = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=True)
    try:
        sim.update_params(haloCenters=halo, **base)
        sim.step(6, sort_every=3)
        parts = sim.particles()
        energy = sim.energy()
    finally:
        sim.destroy()
    cur, nodes = data.reshape(-1).copy(), None
    went = None
    for s in range(6):
        if s % 3 == 0:
            ref = oracle.sort_and_build(cur, n, bucket=16)
            cur, nodes = ref["sorted"], ref["nodes"]
        else:
            nodes = oracle.refresh_f64(cur, nodes)
        cur, went, _, _, _ = oracle.walk(cur, nodes, op)
    assert np.array_equal(parts.reshape(-1).view(np.uint32), cur.view(np.uint32))
    assert np.array_equal(energy.view(np.uint32), went.view(np.uint32))


def test_walk_step_is_the_reference_step():
    """gs_walk_step = step() (NBodyGPU.js:321-329): the walk on the tree's
    current COMs with no refresh.  refresh + walk_step equals gs_step(1, 0);
    after updateCentralMass without a refresh it walks the old COMs, equal to
    the oracle's walk of the unrefreshed tree (EXACT, bit for bit).  Walking
    or downloading nodes after a sort and before any refresh is refused."""
    banner("walk_step semantics")
    data = ic.make("cpu_disk_16k", n=12000)
    n = data.shape[0]
    base, halo, op = _params(n, theta=0.6, epsilon=0.01, gasDrag=0.25)
    a = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=True)
    b = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=True)
    try:
        for s in (a, b):
            s.update_params(haloCenters=halo, **base)
            s.sort()
        for call in (a.walk_step, a.force, a.nodes, a.count_interactions):
            with pytest.raises(galaxy_hip.GalaxyError):
                call()  # no COMs since the sort
        a.refresh()
        a.walk_step()
        b.step(1, sort_every=0)
        assert np.array_equal(a.particles().view(np.uint32), b.particles().view(np.uint32))
        assert np.array_equal(a.energy().view(np.uint32), b.energy().view(np.uint32))
        # no refresh: the next walk uses the COMs of the refresh before the drift
        nodes = a.nodes()
        a.update_central_mass(3.0, 11)
        before = a.particles()
        a.walk_step()
        after = a.particles()
    finally:
        a.destroy()
        b.destroy()
    want, _, _, _, _ = oracle.walk(before.reshape(-1), nodes, op)
    assert np.array_equal(after.reshape(-1).view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("theta,eps", [(0.5, 0.005), (0.57, 0.001), (0.7, 0.01)])
def test_fast_step_own_leaf_density(theta, eps):
    """Default path, one rebuilding step: the density of each particle's own
    leaf (seen in the gas pad, -(1 + density)) equals the oracle's bit for
    bit.  At theta^2 < 1/3 the walk tests the own leaf only on leaves some
    lane opens (no particle can accept its own leaf); 0.57 sits just inside
    that bound, 0.7 outside it."""
    n = 30000
    banner(f"fast step own-leaf density theta={theta} eps={eps}")
    data = ic.make("disk_500k", n=n)
    gas = data[:, 7] < -0.5
    assert gas.sum() > 5000
    base, halo, op = _params(n, theta=theta, epsilon=eps)
    sim = galaxy_hip.NBodyGPU(data, bucket=16)
    try:
        sim.update_params(haloCenters=halo, **base)
        sim.step(1, sort_every=1)
        parts = sim.particles()
    finally:
        sim.destroy()
    ref = oracle.sort_and_build(data.reshape(-1), n, bucket=16)
    pout, _, _, _, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
    want = pout.reshape(-1, 8)
    # same particle order (the step's sort of identical positions)
    assert np.array_equal(parts[:, 3].view(np.uint32), want[:, 3].view(np.uint32))
    g = want[:, 7] < -0.5
    assert g.sum() == gas.sum()
    assert np.array_equal(parts[g, 7].view(np.uint32), want[g, 7].view(np.uint32))
    if 3 * theta * theta < 1:  # then every particle visits its own leaf
        assert (want[g, 7] < -1.0).all()
