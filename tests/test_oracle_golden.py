"""Pins the CPU oracle against golden vectors written from the reference's own
src/physics/mortonTree.js (tests/golden/gen_golden.mjs), and against closed
forms / brute force where the reference has no test (part1by2, stable sort)."""
import numpy as np
import pytest

import cases
from oracle import oracle


def test_part1by2_known_answers_and_bruteforce():
    assert oracle.part1by2(0x3FF) == 0x9249249
    assert oracle.part1by2(1) == 1
    assert oracle.part1by2(2) == 8
    for v in range(1024):
        naive = 0
        for b in range(10):
            naive |= ((v >> b) & 1) << (3 * b)
        assert oracle.part1by2(v) == naive
    # only the low 10 bits take part (mortonTree.js:30)
    assert oracle.part1by2(0x400 | 5) == oracle.part1by2(5)


@pytest.mark.parametrize("d,want", [
    (0.0, 0), (-0.0, 0), (1023.9, 1023), (-1.5, -1), (float("nan"), 0),
    (float("inf"), 0), (-float("inf"), 0), (2.0**31, -(2**31)), (2.0**32 + 7.9, 7),
    (-(2.0**32) - 3.2, -3), (7e12, 7e12 % 2**32 - (2**32 if 7e12 % 2**32 >= 2**31 else 0)),
])
def test_js_toint32(d, want):
    assert oracle.js_toint32(d) == int(want)


@pytest.mark.parametrize("case", cases.cases("stored"), ids=lambda c: c["name"])
def test_oracle_matches_mortonTree_js(case):
    n = case["N"]
    data = cases.input_of(case)
    stride = case["stride"]
    codes, size = oracle.compute_morton_codes(data, n, stride)
    assert np.float64(size).tobytes().hex() == case["rootSizeHex"]
    assert np.array_equal(codes, cases.load(case["name"], "codes.u32", np.uint32))
    if case.get("codesOnly"):
        return
    out = oracle.sort_and_build(data, n, track_idx2=case["trackIdx2"],
                                max_nodes=case["maxNodes"])
    assert np.array_equal(out["perm"], cases.load(case["name"], "perm.u32", np.uint32))
    assert out["num_nodes"] == case["numNodes"]
    assert out["level_ranges"] == case["levelRanges"]
    want_nodes = cases.load(case["name"], "nodes.u32", np.uint32)
    assert np.array_equal(out["nodes"], want_nodes)
    assert cases.sha(out["sorted"]) == case["sha"]["sorted"]
    if case["centralIdx"] >= 0:
        assert out["central_idx"] == case["centralIdx"]
        assert out["central_idx2"] == case["centralIdx2"]


@pytest.mark.parametrize("case", cases.cases("digest"), ids=lambda c: c["name"])
def test_oracle_matches_mortonTree_js_digests(case):
    n = case["N"]
    data = cases.input_of(case)
    codes, size = oracle.compute_morton_codes(data, n, 8)
    assert np.float64(size).tobytes().hex() == case["rootSizeHex"]
    assert cases.sha(codes) == case["sha"]["codes"]
    out = oracle.sort_and_build(data, n, track_idx2=case["trackIdx2"])
    assert cases.sha(out["perm"]) == case["sha"]["perm"]
    assert out["num_nodes"] == case["numNodes"]
    assert out["level_ranges"] == case["levelRanges"]
    assert cases.sha(out["nodes"]) == case["sha"]["nodes"]
    assert cases.sha(out["sorted"]) == case["sha"]["sorted"]
    assert out["central_idx"] == case["centralIdx"]


def test_sort_is_the_stable_sort():
    rng = np.random.default_rng(5)
    for n in (1, 7, 1000, 50000):
        codes = rng.integers(0, 1 << 30, n, dtype=np.uint32)
        codes[: n // 3] = codes[n // 3: 2 * (n // 3)]  # plenty of ties
        perm = oracle.radix_sort_indices(codes)
        assert np.array_equal(perm, np.argsort(codes, kind="stable").astype(np.uint32))


def test_gather_direction_matches_remapIndex_test():
    # test/StarSelection.test.js:51-57: perm [2,0,1] means out[0] = in[2]
    src = np.arange(3 * 8, dtype=np.float32)
    out = oracle.apply_permutation(src, np.array([2, 0, 1], np.uint32), 8)
    assert np.array_equal(out.reshape(3, 8)[0], src.reshape(3, 8)[2])
    assert np.array_equal(out.reshape(3, 8)[1], src.reshape(3, 8)[0])
