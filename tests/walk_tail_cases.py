"""Closed-form cases for the tail of the reference's `main` kernel
(nbody_bh.wgsl:144-180: halos, kick-drift, gas drag, pad density, specific
energy) and for the pair/monopole potential (:114-116, :130).

Each case is a small particle set with parameters whose outcome follows from
the physics the shader states, written here in a form of its own (cylindrical
decomposition for the drag, potentials differentiated numerically for the
accelerations, the total pair energy for the halved pair potential), in
float64.  `tests/test_oracle_walk_tail.py` holds the oracle to these, and
`tests/test_gpu_walk_tail.py` the HIP library.  Each check is chosen so that a
plausible slip fails it: damping the azimuthal instead of the non-azimuthal
velocity, a tangent of the wrong orientation, dt missing from the damping,
size^2 for size^3 in the density, the 1 missing from the gas pad, a missing
1/2 on the pair potential or one on the halo potential, (r^2 + a^2) for
(r + a)^2 in the Hernquist halo, G dropped from a potential.
"""
import numpy as np

F32 = np.float32


def _rec(pos, m, vel, pad=0.0):
    r = np.zeros(8, np.float32)
    r[:3] = pos
    r[3] = m
    r[4:7] = vel
    r[7] = pad
    return r


def _root_size(pos):
    """Edge of the root cube (mortonTree.js:53): max extent * 1.0001 + 1e-9."""
    p = pos.astype(np.float64)
    ext = (p.max(0) - p.min(0)).max()
    return float(F32(ext * 1.0001 + 1e-9))


def hernquist_pot(x, c, G, M, a):
    r = np.linalg.norm(np.asarray(x, np.float64) - np.asarray(c, np.float64))
    return -G * M / (r + a)


def plummer_pot(x, y, G, m, eps):
    r2 = ((np.asarray(y, np.float64) - np.asarray(x, np.float64)) ** 2).sum()
    return -G * m / np.sqrt(r2 + eps * eps)


def grad(f, x, h=1e-5):
    """Central difference gradient of a scalar function at x (float64)."""
    x = np.asarray(x, np.float64)
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        g[k] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def drag_expect(x, v, gamma, dt):
    """Gas drag: the azimuthal velocity about the z axis survives, every other
    component decays by exp(-gamma dt).  phi-hat = (-sin phi, cos phi, 0)."""
    phi = np.arctan2(x[1], x[0])
    t = np.array([-np.sin(phi), np.cos(phi), 0.0])
    v = np.asarray(v, np.float64)
    vt = (v @ t) * t
    return vt + (v - vt) * np.exp(-gamma * dt)


def case_gas_drag():
    """One gas particle off-axis, no other mass, no halo.  N = 1: the root is a
    leaf of size 1e-9, below eps/2, so the walk accepts it as a monopole of the
    particle itself: zero force (r = 0) and the self potential -G m / eps."""
    G, eps, dt, gamma = 1.5, 0.1, 0.01, 2.0
    x = np.array([0.6, 0.8, 0.3])
    v = np.array([0.3, -0.2, 0.5])
    m = 0.75
    data = np.stack([_rec(x, m, v, -1.0)])
    params = dict(dt=dt, G=G, epsilon=eps, theta=0.5, gasDrag=gamma)
    xs = F32(x).astype(np.float64)
    vs = F32(v).astype(np.float64)
    v1 = drag_expect(xs, vs, float(F32(gamma)), float(F32(dt)))
    size = _root_size(F32(x)[None])
    dens = 1.0 / size ** 3
    pot = -G * float(F32(m)) / eps
    want = dict(v=v1[None], x=(xs + v1 * dt)[None], pad=np.array([-(1.0 + dens)]),
                energy=np.array([0.5 * (v1 @ v1) + 0.5 * pot]))
    return data, params, 16, want


def case_star_no_drag():
    """The same particle as a star (pad 0): no drag, pad stays 0."""
    data, params, bucket, _ = case_gas_drag()
    data = data.copy()
    data[0, 7] = 0.0
    x = data[0, :3].astype(np.float64)
    v = data[0, 4:7].astype(np.float64)
    pot = -params["G"] * float(data[0, 3]) / params["epsilon"]
    want = dict(v=v[None], x=(x + v * params["dt"])[None], pad=np.array([0.0]),
                energy=np.array([0.5 * (v @ v) + 0.5 * pot]))
    return data, params, bucket, want


def case_leaf_density():
    """Sixteen particles, bucket 16: the root is the leaf of every particle, so
    each density is 16 / size^3 and a gas particle's pad becomes -(1 + that);
    stars keep pad 0.  Pads only (forces are not closed-form here)."""
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 1, (16, 3)) * np.array([1.0, 0.5, 0.25])
    data = np.stack([_rec(pos[k], 0.01 + 0.001 * k, (0, 0, 0), -1.0 if k % 3 == 0 else 0.0)
                     for k in range(16)])
    size = _root_size(data[:, :3])
    dens = 16.0 / size ** 3
    pad = np.where(data[:, 7] < -0.5, -(1.0 + dens), 0.0)
    return data, dict(dt=1e-4, G=1.0, epsilon=0.05, theta=0.5), 16, dict(pad=pad)


def case_two_leaf_density():
    """Seventeen gas particles, bucket 16, nine in the low octant of the root
    cube and eight in the high one: the root opens into two leaves of edge
    size/2, and each particle's density is its own leaf's count / (size/2)^3."""
    rng = np.random.default_rng(12)
    lo = rng.uniform(0.05, 0.45, (9, 3))
    hi = rng.uniform(0.55, 0.95, (8, 3))
    pos = np.concatenate([lo, hi])
    pos[0] = (0.0, 0.0, 0.0)  # pin the cube: min 0, max 1 on x
    pos[9] = (1.0, 0.97, 0.96)
    data = np.stack([_rec(pos[k], 0.02, (0, 0, 0), -1.0) for k in range(17)])
    half = _root_size(data[:, :3]) / 2
    count = np.where(np.arange(17) < 9, 9.0, 8.0)
    return data, dict(dt=1e-4, G=1.0, epsilon=0.05, theta=0.5), 16, dict(
        pad=-(1.0 + count / half ** 3))


def case_pair():
    """Two particles (one leaf): each feels the Plummer potential of the other,
    a = -grad phi with phi = -G m_j / sqrt(r^2 + eps^2); the specific energy is
    1/2 |v'|^2 + 1/2 phi, so sum_i m_i (e_i - 1/2 |v_i'|^2) is the pair's
    potential energy -G m_0 m_1 / sqrt(r^2 + eps^2)."""
    G, eps, dt = 2.0, 0.1, 0.01
    xa, xb = np.array([0.1, -0.2, 0.05]), np.array([0.4, 0.2, -0.07])
    va, vb = np.array([0.05, 0.1, -0.02]), np.array([-0.03, 0.0, 0.04])
    ma, mb = 0.7, 0.3
    data = np.stack([_rec(xa, ma, va), _rec(xb, mb, vb)])
    d = data.astype(np.float64)
    pos, m, vel = d[:, :3], d[:, 3], d[:, 4:7]
    acc, pot = np.zeros((2, 3)), np.zeros(2)
    for i in range(2):
        j = 1 - i
        pot[i] = plummer_pot(pos[i], pos[j], G, m[j], eps)
        acc[i] = -grad(lambda x: plummer_pot(x, pos[j], G, m[j], eps), pos[i])
    v1 = vel + acc * dt
    e = 0.5 * (v1 ** 2).sum(1) + 0.5 * pot
    u_pair = -G * m[0] * m[1] / np.sqrt(((pos[0] - pos[1]) ** 2).sum() + eps * eps)
    return data, dict(dt=dt, G=G, epsilon=eps, theta=0.5), 16, dict(
        acc=acc, pot=pot, v=v1, x=pos + v1 * dt, energy=e, u_pair=u_pair)


def case_far_monopole():
    """A distant cluster enters the lone particle's walk as one node: its
    potential is -G M / sqrt(d^2 + eps^2) with d the distance to the cluster's
    centre of mass."""
    rng = np.random.default_rng(3)
    n, G, eps = 200, 1.25, 0.01
    pos = rng.uniform(-0.01, 0.01, (n, 3)) + np.array([100.0, 0, 0])
    data = np.zeros((n + 1, 8), np.float32)
    data[:n, :3] = pos
    data[:n, 3] = rng.uniform(0.5, 1.5, n)
    data[n, :4] = (0.0, 0.0, 0.0, 1.0)
    d = data.astype(np.float64)
    M = d[:n, 3].sum()
    com = (d[:n, :3] * d[:n, 3:4]).sum(0) / M
    pot = -G * M / np.sqrt((com ** 2).sum() + eps * eps)
    acc = -grad(lambda x: -G * M / np.sqrt(((com - x) ** 2).sum() + eps * eps), np.zeros(3))
    return data, dict(dt=1e-3, G=G, epsilon=eps, theta=0.5), 16, dict(
        lone=n, acc_lone=acc, pot_lone=pot)


def case_halos():
    """One particle in two Hernquist halos: phi_h = -G M / (r_h + a) per halo,
    a = -grad of their sum; the specific energy adds the halo potential
    unhalved: e = 1/2 |v'|^2 + 1/2 phi_self + phi_halo, with phi_self = -G m /
    eps (the N = 1 root accepted below eps/2, as in case_gas_drag)."""
    G, M, a, eps, dt = 1.5, 5.0, 2.0, 0.1, 1e-3
    c1, c2 = (0.5, -1.0, 0.25), (-2.0, 0.5, 1.0)
    x = np.array([3.0, 0.0, 4.0])
    v = np.array([0.1, 0.2, -0.3])
    m = 0.5
    data = np.stack([_rec(x, m, v)])
    xs = F32(x).astype(np.float64)
    vs = F32(v).astype(np.float64)
    halo = lambda p: hernquist_pot(p, c1, G, M, a) + hernquist_pot(p, c2, G, M, a)
    acc = -grad(halo, xs)
    v1 = vs + acc * dt
    pot_self = -G * m / eps
    e = 0.5 * (v1 @ v1) + 0.5 * pot_self + halo(xs)
    params = dict(dt=dt, G=G, epsilon=eps, theta=0.5, haloMass=M, haloA=a,
                  haloCenters=[c1, c2])
    return data, params, 16, dict(v=v1[None], x=(xs + v1 * dt)[None], energy=np.array([e]),
                                  pot_halo=np.array([halo(xs)]))


CASES = {
    "gas_drag": case_gas_drag,
    "star_no_drag": case_star_no_drag,
    "leaf_density": case_leaf_density,
    "two_leaf_density": case_two_leaf_density,
    "pair": case_pair,
    "halos": case_halos,
}


def in_input_order(out, perm):
    """Outputs in the sorted order (out[new], perm[new] = old) -> input order."""
    res = np.empty_like(out)
    res[perm] = out
    return res


def check_step(name, want, parts, energy, rtol):
    """parts (n x 8) and energy (n) in input order, against the closed form."""
    p = parts.astype(np.float64)
    if "v" in want:
        np.testing.assert_allclose(p[:, 4:7], want["v"], rtol=rtol, atol=1e-7,
                                   err_msg=f"{name}: velocity")
    if "x" in want:
        np.testing.assert_allclose(p[:, :3], want["x"], rtol=rtol, atol=1e-7,
                                   err_msg=f"{name}: position")
    if "pad" in want:
        np.testing.assert_allclose(p[:, 7], want["pad"], rtol=rtol, err_msg=f"{name}: pad")
    if "energy" in want:
        np.testing.assert_allclose(energy.astype(np.float64), want["energy"], rtol=rtol,
                                   err_msg=f"{name}: specific energy")
    if "u_pair" in want:
        m = p[:, 3]
        ke = 0.5 * (p[:, 4:7] ** 2).sum(1)
        u = (m * (energy.astype(np.float64) - ke)).sum()
        np.testing.assert_allclose(u, want["u_pair"], rtol=rtol,
                                   err_msg=f"{name}: pair potential energy")
