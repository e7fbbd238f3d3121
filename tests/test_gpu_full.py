"""Full-size parity at the headline workload: the collision of two 2,097,152-
particle disks (N = 4,194,304, theta 0.5, eps 0.005, bucket 32).

Morton codes, permutation and tree words are compared bit-for-bit with the
oracle's sortAndBuild at full size; the default path's forces, potentials and
densities on every particle are compared with the transliterated walk (forces
<= 1e-5 relative where the force is at least 1 % of the RMS force, and within
1e-5 of the RMS force where it cancels below that -- the saddle between the
two galaxies; potentials <= 1e-5 relative; densities bit for bit), EXACT
forces on 4,096 seeded-sampled particles bit for bit, and the FP64
direct-sum error is reported."""
import numpy as np
import pytest

import galaxy_hip
from galaxy_hip import ic
from gpu_util import banner, check_force_bar, oracle_tree, rel_err
from oracle import oracle

pytestmark = pytest.mark.gpu

N = 4194304


@pytest.fixture(scope="module")
def headline():
    banner("building collision_4m inputs + oracle tree")
    data = ic.make("collision_4m")
    ref = oracle_tree(data, bucket=32)
    return data, ref


def _gpu(data, exact):
    sim = galaxy_hip.NBodyGPU(data, bucket=32, exact_force=exact)
    try:
        sim.update_params(dt=0.005, G=1.0, epsilon=0.005, theta=0.5)
        sim.sort()
        perm = sim.perm()
        codes = sim.codes(sorted=True)
        sim.refresh()
        nodes = sim.nodes()
        info = sim.tree_info()
        sim.force()
        acc, dens = sim.accel()
    finally:
        sim.destroy()
    return perm, codes, nodes, info, acc, dens


def test_headline_tree_and_forces(headline):
    data, ref = headline
    banner("collision_4m GPU sort/refresh/force (default path)")
    perm, codes, nodes, info, acc, dens = _gpu(data, exact=False)
    assert np.array_equal(np.sort(perm), np.arange(N, dtype=np.uint32))
    assert (np.diff(codes.astype(np.int64)) >= 0).all()
    assert np.array_equal(perm, ref["perm"])
    assert np.array_equal(codes, ref["codes"])
    assert info["num_nodes"] == ref["num_nodes"]
    assert info["level_ranges"] == ref["level_ranges"]
    assert np.array_equal(nodes, ref["nodes"])
    nu = info["num_nodes"] / N
    banner(f"nodes {info['num_nodes']} (nu = {nu:.4f}), levels {info['num_levels']}")

    # the default path against the transliteration on every particle
    op = oracle.pack_params(N, dt=0.005, G=1.0, epsilon=0.005, theta=0.5)
    _, _, rall, rdens, inter_all = oracle.walk(ref["sorted"], ref["nodes"], op)
    err = rel_err(acc[:, :3], rall[:, :3])
    perr = np.abs(acc[:, 3].astype(np.float64) - rall[:, 3]) / np.abs(rall[:, 3].astype(np.float64))
    worst, rel_big, abs_rms, ncancel = check_force_bar(acc, rall)
    banner(f"default path vs transliteration, all {N} particles: relative max {worst:.3e} "
           f"(p99 {np.quantile(err, 0.99):.3e}, median {np.median(err):.3e}); "
           f"{rel_big:.3e} where |F| >= 1% RMS; max error {abs_rms:.3e} of the RMS force "
           f"({ncancel} particles below 1% RMS); potential max {perr.max():.3e}")
    assert perr.max() <= 1e-5
    assert np.array_equal(dens.view(np.uint32), rdens.view(np.uint32))

    rng = np.random.default_rng(0x6A1A5EED00000003)
    idx = np.sort(rng.choice(N, 4096, replace=False)).astype(np.uint32)
    inter = inter_all[idx]

    ds = oracle.direct_sum_f64(ref["sorted"], 1.0, float(np.float32(0.005 * 0.005)), idx)
    e64 = rel_err(acc[idx, :3], ds)
    banner(f"FP64 direct-sum error (4096 sampled): median {np.median(e64):.3e} "
           f"p99 {np.quantile(e64, 0.99):.3e} max {e64.max():.3e}; "
           f"mean interactions/particle {inter.mean():.0f}")


def test_headline_exact_forces_bit_identical(headline):
    data, ref = headline
    banner("collision_4m EXACT force sample")
    perm, codes, nodes, info, acc, dens = _gpu(data, exact=True)
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(N, 4096, replace=False)).astype(np.uint32)
    op = oracle.pack_params(N, dt=0.005, G=1.0, epsilon=0.005, theta=0.5)
    racc, _ = oracle.walk_subset(ref["sorted"], ref["nodes"], op, idx)
    assert np.array_equal(acc[idx].view(np.uint32), racc.view(np.uint32))
