"""The HIP walk's integrator tail and potentials against the closed forms of
tests/walk_tail_cases.py (nbody_bh.wgsl:114-116, 130, 144-180): gas drag,
own-leaf density and gas pad, pair and monopole potentials, Hernquist halo
potential and kick, specific energy.  Both force paths: EXACT to f32 rounding,
the default rsqrt path within its 1e-5 bar."""
import numpy as np
import pytest

import galaxy_hip
import walk_tail_cases as wt
from gpu_util import banner

pytestmark = pytest.mark.gpu

RTOL = {True: 2e-6, False: 1e-5}


def _gpu_step(data, params, bucket, exact):
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=exact)
    try:
        sim.update_params(**params)
        sim.step(1, sort_every=1)
        parts, energy, perm = sim.particles(), sim.energy(), sim.perm()
    finally:
        sim.destroy()
    return wt.in_input_order(parts, perm), wt.in_input_order(energy, perm)


def _gpu_force(data, params, bucket, exact):
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=exact)
    try:
        sim.update_params(**params)
        sim.sort()
        sim.refresh()
        sim.force()
        acc, _ = sim.accel()
        perm = sim.perm()
    finally:
        sim.destroy()
    return wt.in_input_order(acc, perm)


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("name", sorted(wt.CASES))
def test_step_closed_form(name, exact):
    banner(f"walk tail closed form {name} exact={exact}")
    data, params, bucket, want = wt.CASES[name]()
    parts, energy = _gpu_step(data, params, bucket, exact)
    wt.check_step(name, want, parts, energy, RTOL[exact])


@pytest.mark.parametrize("exact", [True, False])
def test_pair_and_monopole_potential(exact):
    banner(f"pair and monopole potential exact={exact}")
    data, params, bucket, want = wt.case_pair()
    acc = _gpu_force(data, params, bucket, exact)
    a = acc[:, :3].astype(np.float64)
    assert (np.linalg.norm(a - want["acc"], axis=1) <=
            1e-5 * np.linalg.norm(want["acc"], axis=1)).all()
    np.testing.assert_allclose(acc[:, 3], want["pot"], rtol=RTOL[exact])
    data, params, bucket, want = wt.case_far_monopole()
    acc = _gpu_force(data, params, bucket, exact)
    k = want["lone"]
    a = acc[k, :3].astype(np.float64)
    assert np.linalg.norm(a - want["acc_lone"]) <= 1e-5 * np.linalg.norm(want["acc_lone"])
    np.testing.assert_allclose(acc[k, 3], want["pot_lone"], rtol=1e-5)
