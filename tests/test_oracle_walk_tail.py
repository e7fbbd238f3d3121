"""Pins the oracle's integrator tail and potentials (nbody_bh.wgsl:114-116,
130, 144-180 as restated in oracle/walk_ref.c) against the closed forms of
tests/walk_tail_cases.py, and the WGSL f32 refresh (nbody_bh.wgsl:288-313,
or_refresh_f32) against the FP64 refresh that the golden node words pin."""
import numpy as np
import pytest

import walk_tail_cases as wt
from galaxy_hip import ic
from oracle import oracle

RTOL = 2e-6  # f32 arithmetic against float64 closed forms


def _oracle_params(n, p):
    return oracle.pack_params(n, dt=p["dt"], G=p["G"], epsilon=p["epsilon"], theta=p["theta"],
                              halo_mass=p.get("haloMass", 0.0), halo_a=p.get("haloA", 1.0),
                              gas_drag=p.get("gasDrag", 0.0),
                              halo_centers=p.get("haloCenters", ()))


def _oracle_step(data, params, bucket):
    n = data.shape[0]
    ref = oracle.sort_and_build(data.reshape(-1), n, bucket=bucket)
    pout, energy, acc4, dens, _ = oracle.walk(ref["sorted"], ref["nodes"],
                                              _oracle_params(n, params))
    perm = ref["perm"]
    return (wt.in_input_order(pout.reshape(n, 8), perm), wt.in_input_order(energy, perm),
            wt.in_input_order(acc4, perm))


@pytest.mark.parametrize("name", sorted(wt.CASES))
def test_oracle_step_closed_form(name):
    data, params, bucket, want = wt.CASES[name]()
    parts, energy, _ = _oracle_step(data, params, bucket)
    wt.check_step(name, want, parts, energy, RTOL)


def test_oracle_pair_force_and_potential():
    data, params, bucket, want = wt.case_pair()
    _, _, acc4 = _oracle_step(data, params, bucket)
    a = acc4[:, :3].astype(np.float64)
    assert (np.linalg.norm(a - want["acc"], axis=1) <=
            1e-5 * np.linalg.norm(want["acc"], axis=1)).all()
    np.testing.assert_allclose(acc4[:, 3], want["pot"], rtol=RTOL)


def test_oracle_far_monopole_potential():
    data, params, bucket, want = wt.case_far_monopole()
    _, _, acc4 = _oracle_step(data, params, bucket)
    k = want["lone"]
    a = acc4[k, :3].astype(np.float64)
    assert np.linalg.norm(a - want["acc_lone"]) <= 1e-5 * np.linalg.norm(want["acc_lone"])
    np.testing.assert_allclose(acc4[k, 3], want["pot_lone"], rtol=1e-5)


def test_oracle_halo_potential_in_energy():
    """The halo term enters the specific energy once, unhalved (checked through
    the energy of case_halos), and its value is -G M / (r + a) per halo."""
    data, params, bucket, want = wt.case_halos()
    parts, energy, _ = _oracle_step(data, params, bucket)
    v = parts[0, 4:7].astype(np.float64)
    pot_self = -params["G"] * float(data[0, 3]) / params["epsilon"]
    halo = energy[0] - 0.5 * (v @ v) - 0.5 * pot_self
    np.testing.assert_allclose(halo, want["pot_halo"][0], rtol=RTOL)


def test_refresh_f32_within_rounding_of_fp64():
    """The WGSL refresh sums in f32 and divides (com / m); every node's mass and
    COM must agree with the FP64 refresh (golden-pinned summation order) to
    f32 accumulation error, for leaves and internal nodes alike."""
    data = ic.make("cpu_disk_16k", n=3000)
    n = data.shape[0]
    ref = oracle.sort_and_build(data.reshape(-1), n, bucket=16)
    moved = ref["sorted"].reshape(n, 8).copy()
    moved[:, :3] += 0.01 * moved[:, 4:7]  # positions after some motion
    moved = moved.reshape(-1)
    f64 = oracle.refresh_f64(moved, ref["nodes"]).reshape(-1, 16)
    f32 = oracle.refresh_f32(moved, ref["nodes"], ref["level_ranges"]).reshape(-1, 16)
    assert np.array_equal(f32[:, 4:], f64[:, 4:])  # topology words untouched
    m64 = f64[:, 3].view(np.float32).astype(np.float64)
    m32 = f32[:, 3].view(np.float32).astype(np.float64)
    np.testing.assert_allclose(m32, m64, rtol=1e-5)
    assert m32[0] == pytest.approx(moved.reshape(n, 8)[:, 3].astype(np.float64).sum(), rel=1e-5)
    c64 = f64[:, :3].view(np.float32).astype(np.float64)
    c32 = f32[:, :3].view(np.float32).astype(np.float64)
    size = f64[:, 4].view(np.float32).astype(np.float64)
    # a COM lies inside its node; f32 sums err by far less than 1e-4 of its size
    assert (np.abs(c32 - c64).max(1) <= 1e-4 * size + 1e-6).all()


def test_direct_sum_closed_forms():
    """oracle.direct_sum_f64 (the FP64 reference of the force-error figures)
    against closed forms: the softened two-body force G m r / (r^2 + eps^2)^1.5,
    zero net force at the centre of a cube of equal masses, and Newton's third
    law (sum of m a over a random set vanishes)."""
    eps = 0.1
    two = np.zeros((2, 8), np.float32)
    two[0, :4] = (0.0, 0.0, 0.0, 2.0)
    two[1, :4] = (1.5, -0.5, 0.25, 3.0)
    acc = oracle.direct_sum_f64(two.reshape(-1), 0.7, eps * eps, np.array([0, 1], np.uint32))
    r = two[1, :3].astype(np.float64) - two[0, :3]
    inv3 = (r @ r + eps * eps) ** -1.5
    np.testing.assert_allclose(acc[0], 0.7 * 3.0 * r * inv3, rtol=1e-12)
    np.testing.assert_allclose(acc[1], -0.7 * 2.0 * r * inv3, rtol=1e-12)
    cube = np.zeros((9, 8), np.float32)
    k = 0
    for x in (-1, 1):
        for y in (-1, 1):
            for z in (-1, 1):
                cube[k, :4] = (x, y, z, 1.0)
                k += 1
    cube[8, :4] = (0, 0, 0, 5.0)
    a = oracle.direct_sum_f64(cube.reshape(-1), 1.0, 0.0, np.array([8], np.uint32))
    assert np.abs(a).max() < 1e-14
    rng = np.random.default_rng(3)
    pts = np.zeros((300, 8), np.float32)
    pts[:, :3] = rng.normal(size=(300, 3))
    pts[:, 3] = rng.uniform(0.5, 2.0, 300)
    a = oracle.direct_sum_f64(pts.reshape(-1), 1.0, 0.01, np.arange(300, dtype=np.uint32))
    m = pts[:, 3].astype(np.float64)
    net = (m[:, None] * a).sum(0)
    assert np.abs(net).max() < 1e-10 * np.abs(m[:, None] * a).sum()


def test_walk_at_theta_zero_is_the_direct_sum():
    """At theta = 0 no node passes s^2 < theta^2 d^2 (and none passes
    s^2 < eps^2/4 at this eps), so the oracle's walk opens the whole tree: every
    particle interacts with each of the other N - 1 exactly once, and its force
    is the FP64 direct sum to f32 accumulation error."""
    data = ic.make("cpu_disk_16k", n=1500)
    n = data.shape[0]
    ref = oracle.sort_and_build(data.reshape(-1), n, bucket=8)
    op = oracle.pack_params(n, dt=0.005, G=1.0, epsilon=0.05, theta=0.0)
    _, _, acc4, _, inter = oracle.walk(ref["sorted"], ref["nodes"], op)
    assert (inter == n - 1).all()
    ds = oracle.direct_sum_f64(ref["sorted"], 1.0, float(np.float32(0.05 * 0.05)),
                               np.arange(n, dtype=np.uint32))
    a = acc4[:, :3].astype(np.float64)
    err = np.linalg.norm(a - ds, axis=1) / np.linalg.norm(ds, axis=1)
    assert err.max() < 2e-5, err.max()
