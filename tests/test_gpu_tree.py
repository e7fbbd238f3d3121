"""GPU parity of the Morton encode, stable sort, gather, BFS build and FP64
refresh against the golden vectors written from the reference's
mortonTree.js, and against the oracle restatement at sizes the fixtures do
not cover.  All calls go through the C ABI (libgalaxy_hip.so).  Bit-exact."""
import numpy as np
import pytest

import cases
import galaxy_hip
from galaxy_hip import ic
from gpu_util import banner, oracle_tree
from oracle import oracle

pytestmark = pytest.mark.gpu


def _sorted_tree_case(case, data, radix=8):
    n = case["N"]
    sim = galaxy_hip.NBodyGPU(data.reshape(n, 8), bucket=16, max_nodes=case["maxNodes"],
                              radix_bits=radix)
    try:
        sim.encode()
        codes = sim.codes(sorted=False)
        sim.sort()
        perm = sim.perm()
        sorted_ = sim.particles()
        scodes = sim.codes(sorted=True)
        info = sim.tree_info()
        sim.refresh()
        nodes = sim.nodes()
    finally:
        sim.destroy()
    return codes, perm, sorted_, scodes, info, nodes


@pytest.mark.parametrize("case", [c for c in cases.cases("stored") if not c.get("codesOnly")],
                         ids=lambda c: c["name"])
def test_tree_matches_mortonTree_js_golden(case):
    banner(f"golden {case['name']} N={case['N']}")
    data = cases.input_of(case)
    codes, perm, sorted_, scodes, info, nodes = _sorted_tree_case(case, data)
    name = case["name"]
    assert np.array_equal(codes, cases.load(name, "codes.u32", np.uint32))
    assert np.float64(info["root_size"]).tobytes().hex() == case["rootSizeHex"]
    assert np.array_equal(perm, cases.load(name, "perm.u32", np.uint32))
    assert cases.sha(sorted_.reshape(-1)) == case["sha"]["sorted"]
    assert np.array_equal(scodes, codes[perm])
    assert info["num_nodes"] == case["numNodes"]
    assert info["level_ranges"] == case["levelRanges"]
    want = cases.load(name, "nodes.u32", np.uint32)
    if not np.array_equal(nodes, want):
        bad = np.nonzero(nodes != want)[0]
        raise AssertionError(f"{bad.size} node words differ, first at node {bad[0] // 16} "
                             f"word {bad[0] % 16}: {nodes[bad[0]]:#x} vs {want[bad[0]]:#x}")


@pytest.mark.parametrize("radix", [8, 10])
@pytest.mark.parametrize("case", cases.cases("digest"), ids=lambda c: c["name"])
def test_tree_matches_mortonTree_js_digests(case, radix):
    banner(f"digest {case['name']} N={case['N']} radix={radix}")
    data = cases.input_of(case)
    codes, perm, sorted_, scodes, info, nodes = _sorted_tree_case(case, data, radix)
    assert cases.sha(codes) == case["sha"]["codes"]
    assert cases.sha(perm) == case["sha"]["perm"]
    assert cases.sha(sorted_.reshape(-1)) == case["sha"]["sorted"]
    assert info["num_nodes"] == case["numNodes"]
    assert info["level_ranges"] == case["levelRanges"]
    assert cases.sha(nodes) == case["sha"]["nodes"]


def test_stride4_codes_host_surface():
    case = next(c for c in cases.cases("stored") if c["name"] == "stride4_2k")
    banner("stride4 codes")
    data = cases.input_of(case)
    codes, size = galaxy_hip.compute_morton_codes(data, case["N"], 4)
    assert np.array_equal(codes, cases.load("stride4_2k", "codes.u32", np.uint32))
    assert np.float64(size).tobytes().hex() == case["rootSizeHex"]


@pytest.mark.parametrize("radix", [8, 10])
@pytest.mark.parametrize("n", [1, 5, 4095, 4096, 4097, 100003, 1 << 20])
def test_radix_sort_is_stable_sort(n, radix):
    banner(f"radix sort n={n} radix={radix}")
    rng = np.random.default_rng(n)
    codes = rng.integers(0, 1 << 30, n, dtype=np.uint32)
    if n > 10:
        codes[: n // 4] = codes[n // 4: 2 * (n // 4)]
        codes[-(n // 8):] = 7  # long run of equal keys across tiles
    perm = galaxy_hip.radix_sort_indices(codes, radix_bits=radix)
    assert np.array_equal(perm, np.argsort(codes, kind="stable").astype(np.uint32))


@pytest.mark.parametrize("name,n,bucket,radix", [("cpu_disk_16k", 16384, 16, 8),
                                                 ("collision_4m", 200000, 32, 8),
                                                 ("disk_500k", 500000, 16, 8),
                                                 ("cpu_disk_16k", 16384, 16, 10),
                                                 ("collision_4m", 200000, 32, 10)])
def test_tree_matches_oracle(name, n, bucket, radix):
    banner(f"oracle tree {name} n={n} bucket={bucket} radix={radix}")
    data = ic.make(name, n=n)
    ref = oracle_tree(data, bucket=bucket)
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, radix_bits=radix)
    try:
        sim.sort()
        perm = sim.perm()
        info = sim.tree_info()
        sim.refresh()
        nodes = sim.nodes()
        sorted_ = sim.particles()
    finally:
        sim.destroy()
    assert np.array_equal(perm, ref["perm"])
    assert np.array_equal(sorted_.reshape(-1), ref["sorted"])
    assert info["num_nodes"] == ref["num_nodes"]
    assert info["level_ranges"] == ref["level_ranges"]
    assert np.array_equal(nodes, ref["nodes"])


def test_tree_build_host_surface_bucket32():
    banner("MortonTree.build bucket 32")
    data = ic.make("collision_4m", n=50000)
    ref = oracle_tree(data, bucket=32)
    srt = ref["sorted"].reshape(-1, 8)
    t = galaxy_hip.MortonTree(50000, bucket=32).build(ref["codes"], srt[:, 0:3].copy(),
                                                      srt[:, 3].copy(), 50000, ref["root_size"])
    assert t.numNodes == ref["num_nodes"]
    assert t.levelRanges == ref["level_ranges"]
    assert np.array_equal(t.words, ref["nodes"])


def test_refresh_after_motion_matches_fp64_oracle():
    """COMs re-summed from moved particles equal the FP64 build-order sums."""
    banner("refresh after motion")
    data = ic.make("cpu_disk_16k")
    sim = galaxy_hip.NBodyGPU(data, bucket=16)
    try:
        sim.update_params(dt=0.005, G=1.0, epsilon=0.02, theta=0.7)
        sim.sort()
        sim.step(2, sort_every=0)  # topology fixed, particles moved
        sim.refresh()
        nodes = sim.nodes()
        parts = sim.particles()
    finally:
        sim.destroy()
    want = oracle.refresh_f64(parts.reshape(-1), nodes)
    assert np.array_equal(nodes, want)


def test_tiny_pool_makes_the_root_a_leaf():
    """maxNodes = 5: the pool rule (idx + 9 > maxNodes) leaves only the root."""
    banner("tiny max_nodes")
    data = ic.make("cpu_disk_16k", n=4096)
    sim = galaxy_hip.NBodyGPU(data, bucket=16, max_nodes=5)
    try:
        sim.sort()
        info = sim.tree_info()
    finally:
        sim.destroy()
    ref = oracle_tree(data, bucket=16, max_nodes=5)
    assert info["num_nodes"] == ref["num_nodes"] == 1


def test_host_tree_upload_flow_matches_device_sort():
    """The reference's flow (host sortAndBuild -> applyPermutation ->
    uploadBHTree -> refresh -> walk) gives the on-device sort's forces."""
    banner("uploadBHTree flow")
    data = ic.make("cpu_disk_16k", n=20000)
    ref = oracle_tree(data, bucket=16)
    params = dict(dt=0.005, G=1.0, epsilon=0.02, theta=0.6)
    a = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=True)
    b = galaxy_hip.NBodyGPU(data, bucket=16, exact_force=True)
    try:
        a.update_params(**params)
        b.update_params(**params)
        a.apply_permutation(ref["perm"])
        a.upload_tree(ref["nodes"], ref["num_nodes"], ref["level_ranges"])
        a.refresh()
        a.force()
        b.sort()
        b.refresh()
        b.force()
        assert np.array_equal(a.particles(), b.particles())
        assert np.array_equal(a.nodes(), b.nodes())
        acc_a, dens_a = a.accel()
        acc_b, dens_b = b.accel()
        assert np.array_equal(acc_a.view(np.uint32), acc_b.view(np.uint32))
        assert np.array_equal(dens_a.view(np.uint32), dens_b.view(np.uint32))
        # a child index that does not follow its parent is refused on the host
        bad = ref["nodes"].copy()
        bad[8:16] = 0
        with pytest.raises(galaxy_hip.GalaxyError):
            a.upload_tree(bad, ref["num_nodes"], ref["level_ranges"])
        # level ranges the refresh would follow out of the tree, or that do not
        # match the nodes' depths, are refused before any device work
        lr = [list(r) for r in ref["level_ranges"]]
        nn = ref["num_nodes"]
        past = [r[:] for r in lr]
        past[-1][1] += 5  # runs past num_nodes
        gap = [r[:] for r in lr]
        gap[1][0] += 1  # not contiguous
        swap = [r[:] for r in lr]
        swap[1], swap[2] = [swap[1][0], swap[2][1]], [swap[1][0] + swap[2][1], swap[1][1]]
        short = lr[:-1]  # leaves the deepest level uncovered
        for bad_lr in (past, gap, swap, short):
            with pytest.raises(galaxy_hip.GalaxyError):
                a.upload_tree(ref["nodes"], nn, bad_lr)
        with pytest.raises(ValueError):
            a.upload_tree(ref["nodes"][:-16], nn, lr)  # shorter than 16 x num_nodes
        # the handle still holds the good tree
        a.upload_tree(ref["nodes"], nn, lr)
        a.force()
        assert np.array_equal(a.accel()[0].view(np.uint32), acc_a.view(np.uint32))
    finally:
        a.destroy()
        b.destroy()


def test_locate_follows_particles_through_the_sort():
    """gs_locate = remapIndex (main.js:436-440): the new index of an old one."""
    banner("locate")
    data = ic.make("collision_4m", n=30000)
    sim = galaxy_hip.NBodyGPU(data, bucket=32)
    try:
        sim.update_params(dt=0.005, G=1.0, epsilon=0.005, theta=0.5)
        sim.sort()
        perm = sim.perm()
        parts = sim.particles()
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size, dtype=np.uint32)
        for old in (0, 1, 15000, 29999):
            new, pos = sim.locate(old)
            assert new == inv[old]
            assert np.array_equal(pos, parts[new, :4])
            assert np.array_equal(pos, data[old, :4])
        with pytest.raises(galaxy_hip.GalaxyError):
            sim.locate(30000)
    finally:
        sim.destroy()


@pytest.mark.parametrize("radix", [8, 10])
def test_top_levels_from_the_digit_histogram(radix):
    """The step's build takes levels 0 and 1 from the encode kernel's digit
    histogram instead of searching the keys; the leaf rules still apply node
    by node there: maxNodes values where the pool rule (idx + 9 > maxNodes)
    makes the root, all level-1 nodes or only the later ones leaves, and N
    around the bucket (root leaf, one level).  Bit-exact against the oracle,
    for 8- and 10-bit digits (10-bit: the padding keys share the top cell's
    last digit)."""
    base = ic.make("cpu_disk_16k", n=4096)
    runs = [(base, mx) for mx in (5, 9, 10, 11, 13, 26, None)]
    runs += [(base[:n].copy(), None) for n in (1, 2, 16, 17, 40, 300)]
    for data, mx in runs:
        n = data.shape[0]
        banner(f"top levels radix {radix} N={n} maxNodes={mx}")
        ref = oracle_tree(data, bucket=16, max_nodes=mx)
        kw = {} if mx is None else {"max_nodes": mx}
        sim = galaxy_hip.NBodyGPU(data, bucket=16, radix_bits=radix, **kw)
        try:
            sim.sort()
            sim.refresh()
            info = sim.tree_info()
            nodes = sim.nodes()
        finally:
            sim.destroy()
        assert info["num_nodes"] == ref["num_nodes"]
        assert info["level_ranges"] == ref["level_ranges"]
        assert np.array_equal(nodes, ref["nodes"])
