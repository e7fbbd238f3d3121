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
